"""The allocation-free batched CSR build (eg_batch_csr_build_noalloc) against
the handle-owning build (eg_batch_csr_build) on the same records: two calls
in a row into the same output buffers, the second with a smaller batch,
every list and row pointer bit for bit; a batch over the scratch limit is
rejected with EG_ERR_ARG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _batch(eg, rng, n, vmax, rec):
    nv = np.exp(rng.uniform(np.log(20), np.log(vmax), n)).astype(np.uint32)
    nv[n // 2] = 1
    ne = (nv.astype(np.float64) * rng.uniform(4, 24, n)).astype(np.uint64)
    ne[1] = 0
    total = int(ne.sum())
    buf = eg.DeviceBuffer(max(total * rec, 1))
    eg.gen_er_batch(buf.ptr, nv, ne, seed=int(rng.integers(1 << 30)), rec_bytes=rec)
    return buf, nv, ne, total


@pytest.mark.parametrize("rec", [8, 12])
def test_noalloc_equals_batch_build_twice(eg, rec):
    rng = np.random.default_rng(70 + rec)
    dt = eg.REC8 if rec == 8 else eg.REC12
    big = _batch(eg, rng, 300, 2000, rec)   # ~1.9M records: 15-23 MiB of scratch
    small = _batch(eg, rng, 17, 4000, rec)
    cap_e = max(big[3], small[3])
    cap_p = int(big[1].astype(np.int64).sum()) + big[1].size
    out_list = eg.DeviceBuffer(cap_e * rec)
    out_rp = eg.DeviceBuffer(cap_p * 8)
    for call, (buf, nv, ne, total) in enumerate((big, small)):
        print(f"noalloc call {call}: rec={rec} graphs={nv.size} records={total}", flush=True)
        eg.batch_csr_build_noalloc(buf.ptr, nv, ne, out_list.ptr, out_rp.ptr, rec)
        got_list = out_list.download(dt, total)
        nptr = int(nv.astype(np.int64).sum()) + nv.size
        got_rp = out_rp.download(np.uint64, nptr)
        ref = eg.BatchCSR(buf.ptr, nv, ne, rec)
        eoff = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        voff = np.concatenate([[0], np.cumsum(nv.astype(np.int64) + 1)])
        for g in range(nv.size):
            lst, rp = ref.get(g)
            assert np.array_equal(got_list[eoff[g]:eoff[g + 1]].view(np.uint8), lst.view(np.uint8)), (call, g)
            assert np.array_equal(got_rp[voff[g]:voff[g + 1]], rp), (call, g)
        ref.close()


def test_noalloc_rejects_batches_over_the_scratch_limit(eg):
    rng = np.random.default_rng(71)
    nv = np.full(4, 1_000_000, np.uint32)
    ne = np.full(4, 2_000_000, np.uint64)  # 64 MB of 8-byte records, over the 32 MiB scratch
    buf = eg.DeviceBuffer(8)
    with pytest.raises(eg.EndGraphError) as ei:
        eg.batch_csr_build_noalloc(buf.ptr, nv, ne, buf.ptr, buf.ptr, 8)
    assert ei.value.status == -1 and "scratch" in str(ei.value)
    _ = rng
