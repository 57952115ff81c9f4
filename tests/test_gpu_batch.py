"""Batched CSR builds: every graph of a batch equals its own single build."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


@pytest.mark.parametrize("rec", [8, 12])
def test_batch_matches_single_builds(eg, oracle_lib, rec):
    rng = np.random.default_rng(31 + rec)
    n = 40
    V = np.exp(rng.uniform(np.log(50), np.log(20000), n)).astype(np.uint32)
    V[3] = 1           # single-vertex graph
    Ecnt = (V.astype(np.float64) * rng.uniform(8, 32, n)).astype(np.uint64)
    Ecnt[5] = 0        # graph without edges
    total = int(Ecnt.sum())
    print(f"batch of {n} graphs, {total} records, rec={rec}", flush=True)
    buf = eg.DeviceBuffer(max(total * rec, 1))
    eg.gen_er_batch(buf.ptr, V, Ecnt, seed=5, rec_bytes=rec)
    allrec = buf.download(eg.REC8 if rec == 8 else eg.REC12, total)
    b = eg.BatchCSR(buf.ptr, V, Ecnt, rec)
    off = np.concatenate([[0], np.cumsum(Ecnt)]).astype(np.int64)
    for g in range(n):
        e = allrec[off[g]:off[g + 1]]
        assert e.shape[0] == 0 or (e["src"].max() < V[g] and e["dst"].max() < V[g])
        lst, rp = b.get(g)
        order = np.argsort(e["src"], kind="stable")
        assert np.array_equal(lst.view(np.uint8), e[order].view(np.uint8)), g
        exp_rp = np.searchsorted(e["src"][order].astype(np.int64), np.arange(int(V[g]) + 1))
        assert np.array_equal(rp, exp_rp.astype(np.uint64)), g
        if g % 10 == 0 and e.shape[0]:
            og = oracle_lib.OracleGraph(e, int(V[g]), P=1, S=1)
            o_list, o_idx, o_bm = og.csr(0, 0)
            assert np.array_equal(lst.view(np.uint8), o_list.view(np.uint8))


def test_batch_rejects_out_of_range(eg):
    e = np.zeros(4, H.REC8)
    e["src"] = [0, 1, 2, 3]
    e["dst"] = [0, 1, 2, 5]  # graph 1 has 3 vertices
    buf = eg.DeviceBuffer.from_array(e)
    with pytest.raises(eg.EndGraphError) as ei:
        eg.BatchCSR(buf.ptr, [4, 3], [2, 2])
    assert ei.value.status == -2


def test_batch_rebuild_reuses_handle(eg):
    """A rebuild from new records of the same sizes equals a fresh build."""
    rng = np.random.default_rng(8)
    n = 12
    V = np.exp(rng.uniform(np.log(100), np.log(5000), n)).astype(np.uint32)
    Ecnt = (V.astype(np.float64) * rng.uniform(8, 32, n)).astype(np.uint64)
    total = int(Ecnt.sum())
    a = eg.DeviceBuffer(total * 8)
    c = eg.DeviceBuffer(total * 8)
    eg.gen_er_batch(a.ptr, V, Ecnt, seed=1)
    eg.gen_er_batch(c.ptr, V, Ecnt, seed=2)
    b = eg.BatchCSR(a.ptr, V, Ecnt)
    b.rebuild(c.ptr)
    fresh = eg.BatchCSR(c.ptr, V, Ecnt)
    for g in range(n):
        l1, r1 = b.get(g)
        l2, r2 = fresh.get(g)
        assert np.array_equal(l1.view(np.uint8), l2.view(np.uint8)) and np.array_equal(r1, r2), g
