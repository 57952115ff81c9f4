"""The reference's own API, step by step: process_edges / process_vertices as
the C++ facade (Graph<EdgeData>) drives them, through eg_bfs_level,
eg_cc_iteration, eg_sssp_iteration and eg_pr_init / eg_pr_pull / eg_pr_apply,
run to completion from Python over the C ABI.

Every array is compared with the oracle: BFS parents and depths, CC labels and
SSSP distances bit for bit, PageRank ranks and the per-iteration delta(i) of
toolkits/pagerank.cpp within a stated tolerance.  Each step's sparse/dense
decision is checked against the rule of core/graph.hpp:1483-1498 (repairs
R3/R4: use_undirected => sparse iff the active set's out-degree sum <
edges/20, use_iteration => dense, otherwise sparse), evaluated on the active
set the step was given, whose out-degree sum is recomputed here from the
downloaded bitmap and the oracle's degrees.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

REFERENCE, AUTO, PUSH, PULL = 0, 1, 2, 3
PR_FP64_TOL = 1e-12    # relative L1 against the FP64 oracle (summation order only)
PR_FP32C_TOL = 1e-6    # relative L1, compensated FP32 against the FP64 oracle
# delta(i) relative tolerance: FP64 differs from the oracle only in summation
# order; with FP32 gathers each delta (a sum of |x' - x| * deg, which shrinks
# about 0.85x per iteration while the FP32 rounding of x does not) carries
# rounding of relative size ~6e-8 / 0.85^i, about 2e-6 by iteration 19
DELTA_RTOL = {0: 1e-9, 1: 1e-4}


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def rule_sparse(policy, program, active_edges, E):
    """core/graph.hpp:1483-1498 (R3/R4) per policy; `edges / 20` is integer division."""
    if policy == PUSH:
        return True
    if policy == PULL:
        return False
    if policy == REFERENCE and program == "rooted":
        return True
    if policy == REFERENCE and program == "iterative":
        return False
    return active_edges < E // 20


def _graph(eg, oracle_lib, edges, V, P, S, sym=False):
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, 1, 1, edges.dtype.itemsize)
    g.build(buf.ptr, edges.shape[0], sym)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=1, symmetrize=sym, threads=8)
    return g, og, buf


def _rmat(eg, scale, seed, rec=8):
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * rec)
    eg.gen_rmat(buf.ptr, E, scale, seed=seed, rec_bytes=rec, w_lo=0.0, w_hi=1.0)
    return buf.download(eg.REC8 if rec == 8 else eg.REC12, E), V


def _bitmap(eg, V):
    return eg.DeviceBuffer(((V >> 6) + 1) * 8)


def _ids(eg, bm, V):
    return eg.bitmap_ids(bm.download(np.uint64, (V >> 6) + 1), V)


@pytest.mark.parametrize("scale,P,S", [(14, 1, 1), (14, 4, 2)])
@pytest.mark.parametrize("policy", [REFERENCE, AUTO, PUSH, PULL])
def test_bfs_level_loop(eg, oracle_lib, scale, P, S, policy):
    """bfs.cpp:22-112 as the facade runs it: one eg_bfs_level per level, the
    frontier's out-edge sum handed from one level to the next; rows of 33-512
    in-edges (one lane each in the dense step) and larger rows (cut into work
    items) both occur at R-MAT scale 14."""
    print(f"api bfs scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, seed=31)
    g, og, _ = _graph(eg, oracle_lib, edges, V, P, S)
    indeg = og.in_degree()
    assert ((indeg > 32) & (indeg <= 512)).any() and (indeg > 512).any()
    deg = og.out_degree().astype(np.int64)
    root = int(np.argmax(deg))
    E = g.edges_num
    fr, nx, vis = _bitmap(eg, V), _bitmap(eg, V), _bitmap(eg, V)
    parent, depth = eg.DeviceBuffer(V * 4), eg.DeviceBuffer(V * 4)
    eg.device_fill(parent.ptr, V, V, np.uint32)
    eg.device_fill(depth.ptr, V, 0xFFFFFFFF, np.uint32)
    eg.device_fill(parent.ptr + 4 * root, 1, root, np.uint32)
    eg.device_fill(depth.ptr + 4 * root, 1, 0, np.uint32)
    for bm in (fr, vis):
        eg.bitmap_clear(bm, V)
        eg.bitmap_set_bit(bm, root)
    active_edges = g.bitmap_degree_sum(fr)
    assert active_edges == deg[root]
    o_parent, o_depth, o_levels = og.bfs(root, 0)
    level, modes = 0, []
    while True:
        # the frontier handed to this level is the oracle's level set
        front = _ids(eg, fr, V)
        assert np.array_equal(front, np.flatnonzero(o_depth == level)), f"frontier of level {level}"
        assert active_edges == int(deg[front].sum())
        nv, ne, sparse = g.bfs_level(policy, level, fr, nx, vis, parent, depth, active_edges)
        assert sparse == rule_sparse(policy, "rooted", active_edges, E), f"mode of level {level}"
        modes.append(sparse)
        level += 1
        got = _ids(eg, nx, V)
        assert nv == got.size and ne == int(deg[got].sum())
        if nv == 0:
            break
        fr, nx = nx, fr
        active_edges = ne
    assert level == o_levels
    assert np.array_equal(depth.download(np.uint32, V), o_depth)
    assert np.array_equal(parent.download(np.uint32, V), o_parent)
    if policy == AUTO:
        assert any(modes) and not all(modes), "the auto policy should switch modes on R-MAT"


@pytest.mark.parametrize("scale,P,S", [(14, 1, 1), (13, 3, 2)])
@pytest.mark.parametrize("policy", [REFERENCE, AUTO, PUSH, PULL])
def test_cc_iteration_loop(eg, oracle_lib, scale, P, S, policy):
    """cc.cpp:22-112 on load_undirected_from_directed: label = id, all
    active, one eg_cc_iteration per step until no label changes."""
    print(f"api cc scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, seed=32)
    g, og, _ = _graph(eg, oracle_lib, edges, V, P, S, sym=True)
    deg = og.out_degree().astype(np.int64)
    E = g.edges_num
    assert E == 2 * edges.shape[0]
    act, nx, label = _bitmap(eg, V), _bitmap(eg, V), eg.DeviceBuffer(V * 4)
    eg.device_iota(label.ptr, V)
    eg.bitmap_fill(act, V)
    active_edges = g.bitmap_degree_sum(act)
    assert active_edges == E
    modes = []
    for step in range(1000):
        a = _ids(eg, act, V)
        assert active_edges == int(deg[a].sum())
        nv, ne, sparse = g.cc_iteration(policy, act, nx, label, active_edges)
        assert sparse == rule_sparse(policy, "undirected", active_edges, E), f"mode of step {step}"
        modes.append(sparse)
        got = _ids(eg, nx, V)
        assert nv == got.size and ne == int(deg[got].sum())
        if nv == 0:
            break
        act, nx = nx, act
        active_edges = ne
    o_label, _ = og.cc()
    assert np.array_equal(label.download(np.uint32, V), o_label)
    if policy in (REFERENCE, AUTO):
        assert not modes[0] and modes[-1], "all-active first step dense, the last steps sparse"


@pytest.mark.parametrize("scale,P,S", [(13, 1, 1), (13, 2, 2)])
@pytest.mark.parametrize("policy", [REFERENCE, AUTO, PUSH, PULL])
def test_sssp_iteration_loop(eg, oracle_lib, scale, P, S, policy):
    """sssp.cpp:24-108: distance 1e9 but the root, Bellman-Ford steps of
    eg_sssp_iteration until no distance drops; fp32 distances bit-exact."""
    print(f"api sssp scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, seed=33, rec=12)
    g, og, _ = _graph(eg, oracle_lib, edges, V, P, S)
    deg = og.out_degree().astype(np.int64)
    E = g.edges_num
    root = int(np.argmax(deg))
    act, nx, dist = _bitmap(eg, V), _bitmap(eg, V), eg.DeviceBuffer(V * 4)
    eg.device_fill(dist.ptr, V, 1e9, np.float32)
    eg.device_fill(dist.ptr + 4 * root, 1, 0.0, np.float32)
    eg.bitmap_clear(act, V)
    eg.bitmap_set_bit(act, root)
    active_edges = g.bitmap_degree_sum(act)
    for step in range(10000):
        a = _ids(eg, act, V)
        assert active_edges == int(deg[a].sum())
        nv, ne, sparse = g.sssp_iteration(policy, act, nx, dist, active_edges)
        assert sparse == rule_sparse(policy, "rooted", active_edges, E), f"mode of step {step}"
        got = _ids(eg, nx, V)
        assert nv == got.size and ne == int(deg[got].sum())
        if nv == 0:
            break
        act, nx = nx, act
        active_edges = ne
    o_dist, _ = og.sssp(root)
    assert np.array_equal(dist.download(np.float32, V).view(np.uint32), o_dist.view(np.uint32))


@pytest.mark.parametrize("scale,P,S", [(14, 1, 1), (14, 8, 1), (13, 3, 2)])
@pytest.mark.parametrize("prec", [0, 1])
def test_pagerank_step_loop(eg, oracle_lib, scale, P, S, prec):
    """pagerank.cpp:26-103 as the facade runs it: eg_pr_init, then per
    iteration eg_pr_pull (dense: use_iteration) and eg_pr_apply, the last one
    undivided; ranks and every delta(i) against the oracle."""
    print(f"api pagerank scale={scale} P={P} S={S} prec={prec}", flush=True)
    iters = 20
    edges, V = _rmat(eg, scale, seed=34)
    g, og, _ = _graph(eg, oracle_lib, edges, V, P, S)
    assert not rule_sparse(REFERENCE, "iterative", 0, g.edges_num)
    x, nxt = eg.DeviceBuffer(V * 8), eg.DeviceBuffer(V * 8)
    x32, nxt32 = (eg.DeviceBuffer(V * 4), eg.DeviceBuffer(V * 4)) if prec else (None, None)
    delta = eg.DeviceBuffer(8)
    g.pr_init(x, x32)
    deltas = [1.0]  # process_vertices returning 1 per vertex, / V
    for it in range(iters):
        last = it == iters - 1
        g.pr_pull(prec, x32 if prec else x, nxt)
        eg.device_fill(delta.ptr, 1, 0.0, np.float64)
        g.pr_apply(x, nxt, nxt32, last, delta)
        deltas.append(float(delta.download(np.float64, 1)[0]) / V)
        x, nxt = nxt, x
        x32, nxt32 = nxt32, x32
    o_pr, o_deltas = og.pagerank(iters)
    got = x.download(np.float64, V)
    l1 = np.abs(got - o_pr).sum() / np.abs(o_pr).sum()
    assert l1 <= (PR_FP32C_TOL if prec else PR_FP64_TOL), l1
    assert deltas[-1] == 0.0 and o_deltas[-1] == 0.0
    np.testing.assert_allclose(deltas[:-1], o_deltas[:-1], rtol=DELTA_RTOL[prec], atol=1e-15)


def test_mode_rule_boundaries(eg, oracle_lib):
    """The switch at its threshold: a frontier whose out-edge sum is exactly
    edges/20 runs dense, one edge less sparse (the rule is a strict <)."""
    rng = np.random.default_rng(35)
    V = 4000
    edges = H.random_edges(rng, V, 40000, skew=False)
    g, og, _ = _graph(eg, oracle_lib, edges, V, 1, 1, sym=True)
    E = g.edges_num
    act, nx, label = _bitmap(eg, V), _bitmap(eg, V), eg.DeviceBuffer(V * 4)
    for active_edges, want in ((E // 20 - 1, True), (E // 20, False)):
        eg.device_iota(label.ptr, V)
        eg.bitmap_clear(act, V)
        eg.bitmap_set_bit(act, 7)
        for policy in (REFERENCE, AUTO):
            _, _, sparse = g.cc_iteration(policy, act, nx, label, active_edges)
            assert sparse == want
    with pytest.raises(eg.EndGraphError):
        g.cc_iteration(9, act, nx, label, 0)
