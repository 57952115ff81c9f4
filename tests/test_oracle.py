"""Pins the CPU restatement (oracle/) against independent formulations.

The reference ships no tests, fixtures or golden vectors, so every check here
is a closed form, an invariant or a brute force (see tests/helpers.py).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import helpers as H


def _build(oracle_lib, edges, V, P=1, S=1, eta=0, m=1, sym=False):
    return oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=eta, m=m, symmetrize=sym, threads=4)


@pytest.mark.parametrize("seed", [0, 1])
def test_cpbt_matches_bruteforce(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    for trial in range(60):
        n = int(rng.integers(1, 9))
        k = int(rng.integers(1, 5))
        a = rng.integers(0, 20, n).astype(np.uint64)
        if trial % 7 == 0:
            a[:] = 0
        if trial % 5 == 0:
            a[rng.integers(0, n)] = 500
        off = oracle_lib.optimal_offset(a, k)
        best, pre = H.cpbt_bruteforce(a, k)
        assert off[0] == 0 and off[k] == n
        assert np.all(np.diff(off.astype(np.int64)) >= 0)
        sums = [pre[off[i + 1]] - pre[off[i]] for i in range(k)]
        # the bottleneck of the chosen split is the brute-force optimum
        assert max(sums) == best
        # greedy maximality: each of the first k-1 groups is the longest run
        # from its start whose sum stays within the optimum
        for i in range(k - 1):
            if off[i + 1] < n:
                assert pre[off[i + 1] + 1] - pre[off[i]] > best


def test_cpbt_hand_cases(oracle_lib):
    # one heavy page, P > pages, all zero
    assert list(oracle_lib.optimal_offset(np.array([1, 1, 100, 1], np.uint64), 2)) == [0, 2, 4]
    assert list(oracle_lib.optimal_offset(np.array([5, 5], np.uint64), 4)) == [0, 1, 2, 2, 2]
    assert list(oracle_lib.optimal_offset(np.zeros(3, np.uint64), 3)) == [0, 3, 3, 3]


@pytest.mark.parametrize("P,S,sym,weighted", [
    (1, 1, False, False), (1, 2, False, False), (2, 1, False, True), (3, 1, True, False),
    (8, 1, False, False), (3, 2, True, True), (8, 3, False, False), (5, 1, False, False),
])
def test_build_matches_closed_form(oracle_lib, P, S, sym, weighted):
    rng = np.random.default_rng(P * 10 + S + 100 * sym)
    V = 3 * 4096 + 1000 if P < 8 else 12 * 4096 + 77  # V not a multiple of the page size
    edges = H.random_edges(rng, V, 20000 + 7, weighted=weighted)
    eta, m = 3, 20
    og = _build(oracle_lib, edges, V, P, S, eta, m, sym)
    recs, ranks = H.logical_list(edges, P, sym)
    assert og.edges_num == recs.shape[0]
    # degree = bincount (R5: out-degree; in-degree kept)
    assert np.array_equal(og.out_degree(), np.bincount(recs["src"], minlength=V).astype(np.uint32))
    assert np.array_equal(og.in_degree(), np.bincount(recs["dst"], minlength=V).astype(np.uint32))
    # page weights: per page (1+(P-1)/16 + eta*m) * out-edges + eta*(P-1) * vertices
    pa = og.page_amount()
    pages = (V + 4095) // 4096
    pg_edges = np.bincount(recs["src"] // 4096, minlength=pages).astype(np.uint64)
    pg_verts = np.minimum(4096, V - np.arange(pages) * 4096).astype(np.uint64)
    assert np.array_equal(pa, (1 + (P - 1) // 16 + eta * m) * pg_edges + eta * (P - 1) * pg_verts)
    po = og.partition_offset()
    assert po[0] == 0 and po[P] == V and np.all(np.diff(po.astype(np.int64)) >= 0)
    assert np.all((po % 4096 == 0) | (po == V))  # page aligned or clamped to V (R9)
    best, pre = H.cpbt_bruteforce(pa, P) if pages <= 16 else (None, None)
    if best is not None:
        page_b = [min(pages, (int(x) + 4095) // 4096) for x in po]
        assert max(pre[page_b[i + 1]] - pre[page_b[i]] for i in range(P)) == best
    los = [og.local_partition_offset(p) for p in range(P)]
    for p in range(P):
        lo = los[p]
        assert lo[0] == po[p] and lo[S] == po[p + 1] and np.all(np.diff(lo.astype(np.int64)) >= 0)
    for pull in (True, False):
        perm, qs = H.expected_order(recs, ranks, P, po, los, pull)
        exp = recs[perm]
        for p in range(P):
            for s in range(S):
                q = p * S + s
                part = exp[qs == q]
                if pull:
                    got, ci = og.csc(p, s)
                    assert np.array_equal(got.view(np.uint8), part.view(np.uint8)), (p, s)
                    idx, vtx = H.compressed_index(part["dst"], V)
                    assert np.array_equal(ci["index"], idx) and np.array_equal(ci["vertex"], vtx)
                else:
                    got, index, bm = og.csr(p, s)
                    assert np.array_equal(got.view(np.uint8), part.view(np.uint8)), (p, s)
                    srcs = part["src"].astype(np.int64)
                    present = np.unique(srcs)
                    bits = np.zeros((V >> 6) + 1, np.uint64)
                    np.bitwise_or.at(bits, present >> 6, np.left_shift(np.uint64(1), (present & 63).astype(np.uint64)))
                    assert np.array_equal(bm, bits)
                    rowptr = np.searchsorted(srcs, np.arange(V + 1))
                    assert np.array_equal(index[present], rowptr[present].astype(np.uint64))
                    assert np.array_equal(index[present + 1], rowptr[present + 1].astype(np.uint64))


def test_empty_sub_partitions(oracle_lib):
    # all edges inside the first page: later partitions get no edges (R8)
    V = 5 * 4096
    rng = np.random.default_rng(7)
    e = np.zeros(500, H.REC8)
    e["src"] = rng.integers(0, 100, 500)
    e["dst"] = rng.integers(0, 100, 500)
    og = _build(oracle_lib, e, V, P=4, S=2, eta=0, m=1)
    total = 0
    for p in range(4):
        for s in range(2):
            recs, ci = og.csc(p, s)
            total += recs.shape[0]
            if recs.shape[0] == 0:
                assert ci.shape[0] == 1 and ci["index"][0] == 0 and ci["vertex"][0] == V
    assert total == 500


def _small_graph(seed, V=1500, E=12000, weighted=False):
    rng = np.random.default_rng(seed)
    return H.random_edges(rng, V, E, weighted=weighted), V


@pytest.mark.parametrize("P,S", [(1, 1), (3, 2)])
@pytest.mark.parametrize("policy", [0, 1])
def test_bfs_depth_and_min_id_parent(oracle_lib, P, S, policy):
    edges, V = _small_graph(11)
    og = _build(oracle_lib, edges, V, P, S)
    src, dst = edges["src"].astype(np.int64), edges["dst"].astype(np.int64)
    root = int(np.argmax(np.bincount(src, minlength=V)))
    parent, depth, levels = og.bfs(root, policy)
    exp_depth = H.bfs_queue(V, src, dst, root)
    assert np.array_equal(depth, exp_depth)
    assert np.array_equal(parent, H.min_id_parents(V, src, dst, exp_depth, root))
    reached = exp_depth != np.iinfo(np.uint32).max
    assert levels == int(exp_depth[reached].max()) + 1


@pytest.mark.parametrize("P,S", [(1, 1), (4, 2)])
def test_cc_min_label(oracle_lib, P, S):
    edges, V = _small_graph(12, V=3000, E=2500)  # sparse: many components
    og = _build(oracle_lib, edges, V, P, S, sym=True)
    label, iters = og.cc()
    A = sp.coo_matrix((np.ones(len(edges)), (edges["src"], edges["dst"])), shape=(V, V))
    ncomp, comp = connected_components(A, directed=True, connection="weak")
    minid = np.full(ncomp, V, np.int64)
    np.minimum.at(minid, comp, np.arange(V))
    assert np.array_equal(label, minid[comp].astype(np.uint32))


@pytest.mark.parametrize("P,S", [(1, 1), (2, 2)])
def test_sssp_matches_dijkstra(oracle_lib, P, S):
    edges, V = _small_graph(13, V=800, E=6000, weighted=True)
    og = _build(oracle_lib, edges, V, P, S)
    src, dst = edges["src"].astype(np.int64), edges["dst"].astype(np.int64)
    root = int(np.argmax(np.bincount(src, minlength=V)))
    dist, _ = og.sssp(root)
    exp = H.dijkstra_f32(V, src, dst, edges["w"], root)
    assert np.array_equal(dist.view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("P,S", [(1, 1), (3, 1), (2, 2)])
def test_pagerank_matches_dense(oracle_lib, P, S):
    edges, V = _small_graph(14, V=600, E=5000)
    og = _build(oracle_lib, edges, V, P, S)
    pr, deltas = og.pagerank(20)
    exp = H.pagerank_dense(V, edges["src"].astype(np.int64), edges["dst"].astype(np.int64), 20)
    assert np.abs(pr - exp).sum() / np.abs(exp).sum() < 1e-13
    assert deltas[0] == 1.0 and deltas[-1] == 0.0


@pytest.mark.parametrize("P,S", [(1, 1), (4, 2)])
def test_pagerank_deltas_match_dense(oracle_lib, P, S):
    """Every delta(i) of the oracle equals the dense formulation's: a dropped
    degree factor, a missing /V or the last iteration divided would show."""
    edges, V = _small_graph(15, V=700, E=6000)
    og = _build(oracle_lib, edges, V, P, S)
    pr, deltas = og.pagerank(12)
    exp, exp_d = H.pagerank_dense(V, edges["src"].astype(np.int64), edges["dst"].astype(np.int64), 12,
                                  with_deltas=True)
    assert deltas.shape == exp_d.shape
    np.testing.assert_allclose(deltas, exp_d, rtol=1e-11, atol=0)
    assert np.abs(pr - exp).sum() / np.abs(exp).sum() < 1e-13


def test_pagerank_dense_formula_pins():
    """The dense formulation itself on a 3-cycle with a sink, by hand: every
    vertex of the cycle has degree 1, the sink (vertex 3) none."""
    src, dst = np.array([0, 1, 2, 2]), np.array([1, 2, 0, 3])
    # degrees 1, 1, 2, 0; x0 = 1, 1, 1/2, 1
    x, d = H.pagerank_dense(4, src, dst, 2, with_deltas=True)
    x0 = np.array([1.0, 1.0, 0.5, 1.0])
    s1 = np.array([x0[2], x0[0], x0[1], x0[2]])
    y1 = 0.15 + 0.85 * s1
    x1 = y1 / np.array([1.0, 1.0, 2.0, 1.0])
    assert d[1] == pytest.approx((np.abs(x1 - x0) * np.array([1, 1, 2, 1])).sum() / 4, rel=1e-15)
    s2 = np.array([x1[2], x1[0], x1[1], x1[2]])
    assert np.allclose(x, 0.15 + 0.85 * s2, rtol=1e-15) and d[2] == 0.0 and d[0] == 1.0


def test_csr_batch_matches_stable_argsort(oracle_lib):
    """The batched restatement of prepare_push_mode at P = S = 1: each graph's
    records stably sorted by src, and its row pointer = searchsorted."""
    rng = np.random.default_rng(16)
    n = 9
    nv = rng.integers(1, 3000, n).astype(np.uint32)
    nv[2] = 1
    ne = (nv.astype(np.float64) * rng.uniform(0, 12, n)).astype(np.uint64)
    ne[4] = 0
    parts = []
    for i in range(n):
        e = np.zeros(int(ne[i]), H.REC8)
        e["src"] = rng.integers(0, nv[i], int(ne[i]))
        e["dst"] = rng.integers(0, nv[i], int(ne[i]))
        parts.append(e)
    allrec = np.concatenate(parts)
    got, rp = oracle_lib.csr_batch(allrec, nv, ne, threads=3)
    eo, vo = 0, 0
    for i in range(n):
        e = parts[i]
        order = np.argsort(e["src"], kind="stable")
        assert np.array_equal(got[eo:eo + e.shape[0]].view(np.uint8), e[order].view(np.uint8))
        exp_rp = np.searchsorted(e["src"][order].astype(np.int64), np.arange(int(nv[i]) + 1))
        assert np.array_equal(rp[vo:vo + int(nv[i]) + 1], exp_rp.astype(np.uint64))
        eo += e.shape[0]
        vo += int(nv[i]) + 1


def test_widest_paths_reference_hand_case():
    """The numpy widest-path reference on a hand graph: two routes to 3
    (bottlenecks 0.5 and 0.25), a dead end 4, an unreachable 5."""
    src = np.array([0, 1, 0, 2, 3, 5])
    dst = np.array([1, 3, 2, 3, 4, 0])
    w = np.array([0.5, 0.75, 0.25, 1.0, 0.125, 1.0], np.float32)
    width = H.widest_paths(6, src, dst, w, 0)
    assert width.tolist() == [1e9, 0.5, 0.25, 0.5, 0.125, 0.0]
