"""Per-graph tuning of the device-driven traversals (eg_autotune_engine,
eg_set_engine_tune): the tuned values come from the candidate sets, and no
setting, extreme ones included, changes a result: single-root and batched
BFS and SSSP stay bit-exact against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _weighted_graph(eg, oracle_lib, scale=14, P=2):
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=41, rec_bytes=12, w_lo=0.0, w_hi=1.0)
    e12 = buf.download(eg.REC12, E)
    g = eg.Graph(V, P, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    e8 = np.zeros(E, eg.REC8)
    e8["src"], e8["dst"] = e12["src"], e12["dst"]
    og8 = oracle_lib.OracleGraph(e8, V, P=P, S=1, eta=1, m=1, threads=8)
    og12 = oracle_lib.OracleGraph(e12, V, P=P, S=1, eta=1, m=1, threads=8)
    return g, og8, og12, buf, V


def _check(eg, g, og8, og12, roots):
    for r in roots[:3]:
        p, d, _, _ = g.bfs(int(r), eg.POLICY_AUTO)
        op, od, _ = og8.bfs(int(r), 0)
        assert np.array_equal(p, op) and np.array_equal(d, od), f"bfs root {r}"
        dist, _ = g.sssp(int(r), eg.POLICY_AUTO)
        assert np.array_equal(dist.view(np.uint32), og12.sssp(int(r))[0].view(np.uint32)), f"sssp root {r}"
    par, dep, _, _, _ = g.bfs_multi(roots, eg.POLICY_AUTO)
    mdist, _ = g.sssp_multi(roots, eg.POLICY_AUTO)
    for c in (0, len(roots) - 1):
        op, od, _ = og8.bfs(int(roots[c]), 0)
        assert np.array_equal(par[:, c], op) and np.array_equal(dep[:, c], od), f"batched bfs column {c}"
        od = og12.sssp(int(roots[c]))[0]
        assert np.array_equal(mdist[:, c].view(np.uint32), od.view(np.uint32)), f"batched sssp column {c}"


def test_autotune_engine_keeps_results(eg, oracle_lib):
    g, og8, og12, _, V = _weighted_graph(eg, oracle_lib)
    assert g.engine_tune() == {"bfs_dense_div": 20, "sssp_dense_div": 80, "bfs_multi_dense_div": 20,
                               "sssp_multi_dense_div": 2, "hub_degree": 4096, "hub_chunk": 1024,
                               "multi_item_edges": 4096}
    tune, ms = g.autotune_engine(n_roots=12, seed=3)
    print(f"tuned {tune}, best ms {ms}", flush=True)
    assert tune["bfs_dense_div"] in (5, 10, 20, 40, 80)
    assert tune["sssp_dense_div"] in (10, 20, 40, 80, 160)
    assert tune["bfs_multi_dense_div"] in (5, 10, 20, 40)
    assert tune["sssp_multi_dense_div"] in (1, 2, 5, 20)
    assert (tune["hub_degree"], tune["hub_chunk"]) in ((1024, 256), (1024, 1024), (4096, 1024), (4096, 4096),
                                                       (16384, 4096))
    assert tune["multi_item_edges"] in (1024, 4096, 16384)
    assert g.engine_tune() == tune and all(t > 0 for t in ms)
    roots = np.random.default_rng(42).integers(0, V, 20).astype(np.uint32)
    _check(eg, g, og8, og12, roots)


@pytest.mark.parametrize("setting", [
    dict(bfs_dense_div=1, sssp_dense_div=1, bfs_multi_dense_div=1, sssp_multi_dense_div=1),
    dict(bfs_dense_div=100000, sssp_dense_div=100000, bfs_multi_dense_div=100000, sssp_multi_dense_div=100000),
    dict(hub_degree=16, hub_chunk=16, multi_item_edges=64),
    dict(hub_degree=1 << 30, hub_chunk=1 << 20, multi_item_edges=1 << 20),
])
def test_extreme_tunes_keep_results(eg, oracle_lib, setting):
    g, og8, og12, _, V = _weighted_graph(eg, oracle_lib, scale=13, P=1)
    g.set_engine_tune(**setting)
    assert all(g.engine_tune()[k] == v for k, v in setting.items())
    roots = np.random.default_rng(43).integers(0, V, 9).astype(np.uint32)
    _check(eg, g, og8, og12, roots)


def test_invalid_tunes_rejected(eg, oracle_lib):
    g, _, _, _, _ = _weighted_graph(eg, oracle_lib, scale=10, P=1)
    for bad in (dict(hub_chunk=0), dict(bfs_dense_div=0), dict(hub_degree=15), dict(multi_item_edges=0)):
        with pytest.raises(eg.EndGraphError):
            g.set_engine_tune(**bad)
    assert g.engine_tune()["hub_degree"] == 4096
