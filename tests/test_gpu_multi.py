"""Batched (64-root) BFS and SSSP: every root's column equals the single-root
program from that root, bit for bit (parents under the min-id rule, depths,
levels, traversed edges, fp32 distances), and four columns of every batched
SSSP equal the oracle's distances (sssp.cpp:24-108), so an error the batched
and the single-root programs shared would still show."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _graph(eg, edges, V, P=1, S=1, sym=False):
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, 1, 1, edges.dtype.itemsize)
    g.build(buf.ptr, edges.shape[0], sym)
    return g, buf


def _rmat(eg, scale, seed, rec=12):
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * rec)
    eg.gen_rmat(buf.ptr, E, scale, seed=seed, rec_bytes=rec, w_lo=0.0, w_hi=1.0)
    return buf.download(eg.REC8 if rec == 8 else eg.REC12, E), V


def _check_bfs(g, roots, policy):
    par, dep, lev, edg, steps = g.bfs_multi(roots, policy)
    assert steps >= int(lev.max())
    for r, root in enumerate(roots):
        p1, d1, l1, e1 = g.bfs(int(root), policy)
        assert np.array_equal(par[:, r], p1), f"parent column {r} (root {root})"
        assert np.array_equal(dep[:, r], d1), f"depth column {r} (root {root})"
        assert lev[r] == l1 and edg[r] == e1, f"levels/edges of root {root}"
    for r in range(len(roots), 64):
        assert (par[:, r] == g.V).all() and (dep[:, r] == 0xFFFFFFFF).all()


def _check_sssp(g, roots, policy, og=None):
    dist, steps = g.sssp_multi(roots, policy)
    for r, root in enumerate(roots):
        d1, _ = g.sssp(int(root), policy)
        assert np.array_equal(dist[:, r].view(np.uint32), d1.view(np.uint32)), f"dist column {r} (root {root})"
    for r in range(len(roots), 64):
        assert (dist[:, r] == np.float32(1e9)).all()
    if og is not None:
        for r in sorted({0, len(roots) // 3, 2 * len(roots) // 3, len(roots) - 1}):
            o_dist, _ = og.sssp(int(roots[r]))
            assert np.array_equal(dist[:, r].view(np.uint32), o_dist.view(np.uint32)), f"oracle column {r}"


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (14, 1, 1), (13, 4, 2)])
@pytest.mark.parametrize("policy", [0, 1, 2, 3])
def test_bfs_multi_equals_single(eg, scale, P, S, policy):
    print(f"bfs_multi scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, seed=21)
    g, _ = _graph(eg, edges, V, P, S)
    roots = np.random.default_rng(scale * 10 + policy).integers(0, V, 64).astype(np.uint32)
    _check_bfs(g, roots, policy)


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (14, 1, 1), (13, 4, 2)])
@pytest.mark.parametrize("policy", [0, 1, 3])
def test_sssp_multi_equals_single(eg, oracle_lib, scale, P, S, policy):
    """Scales 13 and 14 have more than 4096 vertices, so the sources span
    several words of the single-root program's active-set summary."""
    print(f"sssp_multi scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, seed=22)
    g, _ = _graph(eg, edges, V, P, S)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=1, threads=8)
    roots = np.random.default_rng(scale * 7 + policy).integers(0, V, 64).astype(np.uint32)
    _check_sssp(g, roots, policy, og)


def test_multi_few_and_duplicate_roots(eg):
    """Fewer than 64 roots, repeated roots, a root without out-edges, a
    symmetrised graph."""
    rng = np.random.default_rng(4)
    V = 3000
    edges = H.random_edges(rng, V, 20000, weighted=True, skew=True)
    deg = np.bincount(edges["src"], minlength=V)
    lonely = int(np.flatnonzero(deg == 0)[0])
    g, _ = _graph(eg, edges, V, sym=False)
    roots = np.array([5, 5, lonely, 17, 5], dtype=np.uint32)
    _check_bfs(g, roots, 1)
    _check_sssp(g, roots, 1)
    gs, _ = _graph(eg, edges, V, sym=True)
    _check_bfs(gs, roots[:3], 1)


def test_multi_errors(eg):
    rng = np.random.default_rng(5)
    V = 500
    edges = H.random_edges(rng, V, 3000, weighted=False)
    g, _ = _graph(eg, edges, V)
    with pytest.raises(eg.EndGraphError):
        g.bfs_multi(np.zeros(65, np.uint32))
    with pytest.raises(eg.EndGraphError):
        g.bfs_multi(np.array([V], np.uint32))
    with pytest.raises(eg.EndGraphError):
        g.sssp_multi(np.array([1], np.uint32))  # 8-byte records carry no weights


def test_multi_step_graph_replays(eg):
    """The steps run as a captured HIP graph that is kept between calls with
    the same buffers: replays for new root sets, with single-root programs
    (eager launches on the caller's stream) and the other batched program
    (a recapture) in between, all equal the single-root results."""
    edges, V = _rmat(eg, 12, seed=23)
    g, _ = _graph(eg, edges, V)
    parent, depth = eg.DeviceBuffer(V * 64 * 4), eg.DeviceBuffer(V * 64 * 4)
    dist = eg.DeviceBuffer(V * 64 * 4)
    rng = np.random.default_rng(23)
    for rnd in range(2):
        for i in range(3):
            print(f"step graph replay round={rnd} bfs call={i}", flush=True)
            roots = rng.integers(0, V, 64 if i != 2 else 17).astype(np.uint32)
            par, dep, lev, edg, _ = g.bfs_multi(roots, 1, parent, depth)
            for r, root in enumerate(roots):
                p1, d1, l1, e1 = g.bfs(int(root), 1)
                assert np.array_equal(par[:, r], p1) and np.array_equal(dep[:, r], d1), f"call {i} root {root}"
                assert lev[r] == l1 and edg[r] == e1
        for i in range(3):
            print(f"step graph replay round={rnd} sssp call={i}", flush=True)
            roots = rng.integers(0, V, 64).astype(np.uint32)
            d, _ = g.sssp_multi(roots, 1, dist)
            for r, root in enumerate(roots):
                d1, _ = g.sssp(int(root), 1)
                assert np.array_equal(d[:, r].view(np.uint32), d1.view(np.uint32)), f"call {i} root {root}"
