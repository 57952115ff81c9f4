"""GPU parity of the engine: BFS parents/depths, CC labels and SSSP distances
bit-exact against the oracle; PageRank within a stated L1 bound."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

PR_FP64_TOL = 1e-12    # relative L1, FP64 GPU vs FP64 oracle (only summation order differs)
PR_FP32C_TOL = 1e-6    # relative L1, compensated FP32 vs FP64 oracle


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _pair(eg, oracle_lib, edges, V, P, S, sym=False, eta=1, m=1):
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, eta, m, edges.dtype.itemsize)
    g.build(buf.ptr, edges.shape[0], sym)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=eta, m=m, symmetrize=sym, threads=8)
    return g, og, buf


def _rmat(eg, scale, ef=16, seed=1, rec=8, wlo=0.0, whi=1.0):
    V, E = 1 << scale, ef << scale
    buf = eg.DeviceBuffer(E * rec)
    eg.gen_rmat(buf.ptr, E, scale, seed=seed, rec_bytes=rec, w_lo=wlo, w_hi=whi)
    return buf.download(eg.REC8 if rec == 8 else eg.REC12, E), V


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (14, 1, 1), (14, 8, 1), (13, 3, 2)])
@pytest.mark.parametrize("policy", [0, 1])
def test_bfs_bit_exact(eg, oracle_lib, scale, P, S, policy):
    print(f"bfs scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, S)
    root = int(np.argmax(og.out_degree()))
    parent, depth, levels, _ = g.bfs(root, policy)
    o_parent, o_depth, o_levels = og.bfs(root, policy)
    assert np.array_equal(depth, o_depth)
    assert np.array_equal(parent, o_parent)
    assert levels == o_levels


def _hub_chain_edges(eg):
    """A 300-vertex chain from the root (one push level per link), then a
    hub of 19,000 out-edges (19 push work items of 1024 edges) whose level
    pushes at least V/4 out-edges, vertices of out-degree 255, 256, 4096
    and 4097 (the byte degree's saturation and the hub threshold) behind
    it, and 50,000 random edges among the hub's targets."""
    V = 20000
    rng = np.random.default_rng(11)
    src, dst = [], []
    chain = np.arange(100, 400)
    src += list(chain[:-1]); dst += list(chain[1:])
    src.append(399); dst.append(5)
    src += [5] * (V - 1000); dst += list(range(1000, V))
    for v, d in ((6, 255), (7, 256), (8, 4096), (9, 4097)):
        src.append(5); dst.append(v)
        src += [v] * d; dst += list(rng.choice(np.arange(400, V), d, replace=False))
    src += list(rng.integers(1000, V, 50000)); dst += list(rng.integers(400, V, 50000))
    e = np.zeros(len(src), eg.REC8)
    e["src"], e["dst"] = src, dst
    return rng.permutation(e), V


@pytest.mark.parametrize("P", [1, 4])
@pytest.mark.parametrize("policy", [0, 1, 2, 3])
def test_bfs_hub_and_deep_levels(eg, oracle_lib, P, policy):
    print(f"bfs hub/chain P={P} policy={policy}", flush=True)
    edges, V = _hub_chain_edges(eg)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, 1)
    assert int(og.out_degree()[5]) == V - 1000 + 4
    parent, depth, levels, _ = g.bfs(100, policy)
    o_parent, o_depth, o_levels = og.bfs(100, 0)
    assert levels == o_levels and levels > 300
    assert np.array_equal(depth, o_depth)
    assert np.array_equal(parent, o_parent)


@pytest.mark.parametrize("policy", [0, 1, 2, 3])
def test_sssp_hub_and_deep_levels(eg, oracle_lib, policy):
    print(f"sssp hub/chain policy={policy}", flush=True)
    e8, V = _hub_chain_edges(eg)
    edges = np.zeros(e8.shape[0], eg.REC12)
    edges["src"], edges["dst"] = e8["src"], e8["dst"]
    edges["w"] = np.random.default_rng(12).random(e8.shape[0], dtype=np.float32)
    g, og, _ = _pair(eg, oracle_lib, edges, V, 2, 1)
    dist, iters = g.sssp(100, policy)
    o_dist, _ = og.sssp(100)
    assert iters > 300
    assert np.array_equal(dist.view(np.uint32), o_dist.view(np.uint32))


def _oracle8(oracle_lib, edges12, V, P, S):
    """The oracle's unweighted graph of the same (src, dst) list: its BFS,
    like the reference's, runs on EdgeType<Empty>."""
    e8 = np.zeros(edges12.shape[0], np.dtype([("src", "<u4"), ("dst", "<u4")]))
    e8["src"], e8["dst"] = edges12["src"], edges12["dst"]
    return oracle_lib.OracleGraph(e8, V, P=P, S=S, eta=1, m=1, symmetrize=False, threads=8)


def test_traversals_edge_cases(eg, oracle_lib):
    """Device-driven BFS and SSSP on degenerate graphs: fewer edges than
    ranks with isolated and last-id roots (V not a multiple of 64), V = 1
    with self loops only, self loops spread over pages; every policy."""
    rng = np.random.default_rng(6)
    cases = []
    e = np.zeros(3, eg.REC12); e["src"] = [0, 1, 2]; e["dst"] = [2, 2, 0]; e["w"] = [0.5, 0.25, 1.0]
    cases.append((e, 5000, 4, 2, [0, 1, 2, 4999]))
    e = np.zeros(7, eg.REC12); e["w"] = 0.5
    cases.append((e, 1, 1, 1, [0]))
    e = np.zeros(400, eg.REC12); e["src"] = rng.integers(0, 50, 400); e["dst"] = e["src"]; e["w"] = 0.125
    cases.append((e, 9 * 4096 + 5, 3, 3, [int(e["src"][0]), 9 * 4096 + 4]))
    for edges, V, P, S, roots in cases:
        g, og, _ = _pair(eg, oracle_lib, edges, V, P, S)
        og8 = _oracle8(oracle_lib, edges, V, P, S)
        for root in roots:
            o_parent, o_depth, o_levels = og8.bfs(root, 0)
            o_dist, _ = og.sssp(root)
            for policy in (0, 1, 2, 3):
                print(f"edge case V={V} E={edges.shape[0]} root={root} policy={policy}", flush=True)
                parent, depth, levels, _ = g.bfs(root, policy)
                assert levels == o_levels
                assert np.array_equal(depth, o_depth) and np.array_equal(parent, o_parent)
                dist, _ = g.sssp(root, policy)
                assert np.array_equal(dist.view(np.uint32), o_dist.view(np.uint32))


def test_programs_interleaved_on_one_graph(eg, oracle_lib):
    """BFS, SSSP and the batched programs called in turn on one graph (each
    keeps or re-captures its step graph) stay equal to the oracle."""
    edges, V = _rmat(eg, 13, rec=12, seed=9)
    g, og, _ = _pair(eg, oracle_lib, edges, V, 2, 1)
    og8 = _oracle8(oracle_lib, edges, V, 2, 1)
    roots = [int(np.argmax(og.out_degree())), 7, 123, 4000]
    for i, root in enumerate(roots):
        print(f"interleaved root={root}", flush=True)
        parent, depth, _, _ = g.bfs(root, 1)
        o_parent, o_depth, _ = og8.bfs(root, 0)
        assert np.array_equal(parent, o_parent) and np.array_equal(depth, o_depth)
        dist, _ = g.sssp(root, 1)
        o_dist, _ = og.sssp(root)
        assert np.array_equal(dist.view(np.uint32), o_dist.view(np.uint32))
        mpar, mdep, _, _, _ = g.bfs_multi(np.array(roots[: i + 1], np.uint32), 1)
        assert np.array_equal(mpar[:, i], o_parent) and np.array_equal(mdep[:, i], o_depth)


def test_bfs_pull_only_matches(eg, oracle_lib):
    edges, V = _rmat(eg, 13, seed=4)
    g, og, _ = _pair(eg, oracle_lib, edges, V, 2, 1)
    root = int(np.argmax(og.out_degree()))
    parent, depth, _, _ = g.bfs(root, eg.POLICY_PULL)
    o_parent, o_depth, _ = og.bfs(root, 0)
    assert np.array_equal(depth, o_depth) and np.array_equal(parent, o_parent)


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (14, 4, 1), (13, 2, 2)])
@pytest.mark.parametrize("policy", [0, 2, 3])
def test_cc_bit_exact(eg, oracle_lib, scale, P, S, policy):
    print(f"cc scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, ef=2, seed=7)  # sparse: many components
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, S, sym=True)
    label, _ = g.cc(policy)
    o_label, _ = og.cc()
    assert np.array_equal(label, o_label)


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (14, 1, 1), (13, 4, 1), (13, 2, 2)])
@pytest.mark.parametrize("policy", [0, 1, 3])
def test_sssp_bit_exact(eg, oracle_lib, scale, P, S, policy):
    print(f"sssp scale={scale} P={P} S={S} policy={policy}", flush=True)
    edges, V = _rmat(eg, scale, rec=12)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, S)
    root = int(np.argmax(og.out_degree()))
    dist, _ = g.sssp(root, policy)
    o_dist, _ = og.sssp(root)
    assert np.array_equal(dist.view(np.uint32), o_dist.view(np.uint32))


@pytest.mark.parametrize("scale,P,S", [(12, 1, 1), (15, 1, 1), (14, 8, 1), (13, 2, 2)])
def test_pagerank_fp64_and_fp32c(eg, oracle_lib, scale, P, S):
    print(f"pagerank scale={scale} P={P} S={S}", flush=True)
    edges, V = _rmat(eg, scale)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, S, m=20)
    ref, ref_d = og.pagerank(20)
    runs = {
        "fp64 plain": (eg.PR_FP64, PR_FP64_TOL),
        "fp64 ordered": (eg.PR_FP64 | eg.PR_ORDERED, PR_FP64_TOL),
        "fp32c ordered": (eg.PR_FP32C, PR_FP32C_TOL),
        "fp32c plain": (eg.PR_FP32C | eg.PR_PLAIN, PR_FP32C_TOL),
        "fp32c sorted rows": (eg.PR_FP32C | eg.PR_SORTED, PR_FP32C_TOL),
        "fp64 sorted rows": (eg.PR_FP64 | eg.PR_SORTED, PR_FP64_TOL),
        "fp32c ordered again": (eg.PR_FP32C, PR_FP32C_TOL),
    }
    for name, (prec, tol) in runs.items():
        x, d = g.pagerank(20, prec)
        l1 = np.abs(x - ref).sum() / np.abs(ref).sum()
        print(f"  relative L1 {name}: {l1:.3e}", flush=True)
        assert l1 <= tol, name
        if prec & 0xF == eg.PR_FP64:
            assert np.allclose(d, ref_d, rtol=1e-9, atol=1e-15), name


@pytest.mark.parametrize("P,S", [(1, 1), (3, 1), (2, 3)])
@pytest.mark.parametrize("tv_first", [False, True])
def test_pagerank_ordered_view(eg, oracle_lib, P, S, tv_first):
    """The ordered view (the built CSC's rows relabelled at P = S = 1, rows
    sorted by new source id otherwise), derived before or after a traversal
    view, against the FP64 oracle, with several sub-partitions per partition
    too."""
    print(f"pagerank ordered view P={P} S={S} tv_first={tv_first}", flush=True)
    edges, V = _rmat(eg, 13)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, S, m=20)
    if tv_first:
        g.prepare_traversal_view()
    ref, ref_d = og.pagerank(20)
    for prec, tol in ((eg.PR_FP32C, PR_FP32C_TOL), (eg.PR_FP64 | eg.PR_ORDERED, PR_FP64_TOL)):
        x, d = g.pagerank(20, prec)
        assert np.abs(x - ref).sum() / np.abs(ref).sum() <= tol
    assert np.allclose(d, ref_d, rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("kind,P", [("rmat17", 1), ("rmat17", 4), ("powerlaw", 1), ("powerlaw", 2)])
def test_pagerank_fp32c_hot_and_cold_sources(eg, oracle_lib, kind, P):
    """Graphs with more sources than the FP32C pull stages in LDS (40,960 on
    gfx950), so that a wave's gathers mix LDS reads of the hottest sources
    with memory reads of the others, on both view forms; edge counts not a
    multiple of a wave's range."""
    print(f"pagerank fp32c hot/cold kind={kind} P={P}", flush=True)
    if kind == "rmat17":
        edges, V = _rmat(eg, 17)
        edges = edges[: edges.shape[0] - 1000 - 37]
    else:
        V, E = 300_000, 4_000_000 - 77
        buf = eg.DeviceBuffer(E * 8)
        eg.gen_powerlaw(buf.ptr, E, V, max_out=60_000.0, max_in=20_000.0, seed=5)
        edges = buf.download(eg.REC8, E)
    g, og, _ = _pair(eg, oracle_lib, edges, V, P, 1, m=20)
    ref, _ = og.pagerank(20)
    for prec in (eg.PR_FP32C, eg.PR_FP32C | eg.PR_SORTED):
        x, _ = g.pagerank(20, prec)
        l1 = np.abs(x - ref).sum() / np.abs(ref).sum()
        print(f"  relative L1 {prec:#x}: {l1:.3e}", flush=True)
        assert l1 <= PR_FP32C_TOL


@pytest.mark.parametrize("E", [1, 3, 4099, 200_003])
def test_pagerank_fp32c_run_that_derives_the_view(eg, oracle_lib, E):
    """An FP32C run that derives the ordered view, a later run on the view it
    left, and a run on a view derived by an FP64 run after the view was
    released agree (rows cut by wave ranges add their partials in any order:
    1e-14) and meet the FP64 oracle; edge counts not a multiple of a lane's
    four edges."""
    print(f"pagerank fp32c deriving run E={E}", flush=True)
    rng = np.random.default_rng(E)
    V = max(2, E // 7)
    edges = H.random_edges(rng, V, E, skew=True)
    g, og, _ = _pair(eg, oracle_lib, edges, V, 1, 1, m=20)
    ref, _ = og.pagerank(20)
    n1 = np.abs(ref).sum()
    x1, _ = g.pagerank(20, eg.PR_FP32C)   # derives the view
    x2, _ = g.pagerank(20, eg.PR_FP32C)   # on the view it left
    g.release_pagerank_view()
    g.pagerank(3, eg.PR_FP64 | eg.PR_ORDERED)  # derives the view with the relabelling pass
    x3, _ = g.pagerank(20, eg.PR_FP32C)
    for x in (x1, x2, x3):
        assert np.abs(x - ref).sum() / n1 <= PR_FP32C_TOL
    assert np.abs(x1 - x2).sum() / n1 <= 1e-14
    assert np.abs(x2 - x3).sum() / n1 <= 1e-14


def test_pagerank_layout_flags_rejected(eg):
    edges = np.zeros(4, dtype=eg.REC8)
    edges["src"] = [0, 1, 2, 3]
    edges["dst"] = [1, 2, 3, 0]
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(4, 1, 1, 1, 1, 8)
    g.build(buf.ptr, 4)
    with pytest.raises(eg.EndGraphError):
        g.pagerank(3, eg.PR_FP64 | eg.PR_ORDERED | eg.PR_PLAIN)
    with pytest.raises(eg.EndGraphError):
        g.pagerank(3, eg.PR_FP32C | eg.PR_SORTED | eg.PR_PLAIN)
    with pytest.raises(eg.EndGraphError):
        g.pagerank(3, 5)


def test_pagerank_ordered_view_rebuilt(eg, oracle_lib):
    """The degree-ordered view follows a rebuild of the same handle."""
    rng = np.random.default_rng(11)
    V = 5000
    g = eg.Graph(V, 1, 1, 1, 20, 8)
    for k in range(2):
        edges = H.random_edges(rng, V, 40000 + 7000 * k, skew=True)
        buf = eg.DeviceBuffer.from_array(edges)
        g.build(buf.ptr, edges.shape[0])
        og = oracle_lib.OracleGraph(edges, V, P=1, S=1, eta=1, m=20, threads=8)
        ref, _ = og.pagerank(10)
        x, _ = g.pagerank(10, eg.PR_FP32C)
        assert np.abs(x - ref).sum() / np.abs(ref).sum() <= PR_FP32C_TOL


def test_pagerank_dense_small_closed_form(eg, oracle_lib):
    """Tiny graph: GPU FP64 PageRank against a dense matrix power iteration."""
    rng = np.random.default_rng(3)
    V = 300
    edges = H.random_edges(rng, V, 3000)
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, 1, 1, 0, 1, 8)
    g.build(buf.ptr, edges.shape[0])
    x, _ = g.pagerank(15, eg.PR_FP64)
    exp = H.pagerank_dense(V, edges["src"].astype(np.int64), edges["dst"].astype(np.int64), 15)
    assert np.abs(x - exp).sum() / np.abs(exp).sum() < 1e-13


@pytest.mark.parametrize("V", [1 << 24, (1 << 24) + 1])
@pytest.mark.parametrize("policy", [1, 3])
def test_sssp_dense_summary_at_its_limit(eg, oracle_lib, V, policy):
    """The single-root SSSP's dense steps skip sources in empty words of the
    active bitmap through a summary staged in LDS: 4096 words cover exactly
    2^24 vertices (used), one vertex more runs the dense step without it."""
    print(f"sssp summary limit V={V} policy={policy}", flush=True)
    E = 1 << 22
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, 24, seed=51, rec_bytes=12, w_lo=0.0, w_hi=1.0)
    edges = buf.download(eg.REC12, E)
    g = eg.Graph(V, 1, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    og = oracle_lib.OracleGraph(edges, V, P=1, S=1, eta=1, m=1, threads=8)
    root = int(np.argmax(og.out_degree()))
    dist, _ = g.sssp(root, policy)
    o_dist, _ = og.sssp(root)
    assert np.array_equal(dist.view(np.uint32), o_dist.view(np.uint32))
