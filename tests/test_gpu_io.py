"""Files around the path: the pipelined edge-list reader (preload_graph_file's
read) and the vertex-array dump / restore, byte for byte."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


@pytest.mark.parametrize("weighted,E", [(False, 0), (False, 1), (False, 12_345_678), (True, 9_000_001),
                                        (False, 50_000_003)])
def test_edge_file_round_trip(eg, tmp_path, weighted, E):
    """Sizes below one 32 MiB chunk, across the four reader threads' first
    chunks (odd tails), and over 12 chunks (every thread reuses its buffers)."""
    rng = np.random.default_rng(E)
    edges = H.random_edges(rng, 1 << 20, E, weighted=weighted, skew=False) if E else np.zeros(0, H.REC8)
    path = tmp_path / "g.bin"
    edges.tofile(path)
    rec = edges.dtype.itemsize
    buf, n = eg.load_edge_file(path, rec)
    assert n == E
    got = buf.download(edges.dtype, E)
    assert got.tobytes() == edges.tobytes()


def test_edge_file_errors(eg, tmp_path):
    path = tmp_path / "bad.bin"
    path.write_bytes(bytes(13))
    with pytest.raises(eg.EndGraphError):
        eg.load_edge_file(path, 8)
    with pytest.raises(eg.EndGraphError):
        eg.load_edge_file(tmp_path / "missing.bin", 8)


def test_build_from_file_equals_build_from_memory(eg, oracle_lib, tmp_path):
    rng = np.random.default_rng(3)
    V = 20000
    edges = H.random_edges(rng, V, 200000, skew=True)
    path = tmp_path / "g.bin"
    edges.tofile(path)
    buf, n = eg.load_edge_file(path, 8)
    g = eg.Graph(V, 4, 2, 1, 20, 8)
    g.build(buf.ptr, n)
    og = oracle_lib.OracleGraph(edges, V, P=4, S=2, eta=1, m=20, threads=8)
    from parity import assert_build_equal
    assert_build_equal(g, og)


@pytest.mark.parametrize("P,S,sym,rec", [(1, 1, False, 8), (4, 2, False, 8), (8, 1, True, 8), (3, 1, True, 12),
                                         (5, 1, False, 12)])
def test_build_file_overlapped_equals_oracle(eg, oracle_lib, tmp_path, P, S, sym, rec):
    """eg_graph_build_file (each rank slice's first sorts while the next is
    read): bit-exact against the oracle, rank slices that are not whole
    reader chunks (and a list of 9.3M records: several chunks per slice)."""
    rng = np.random.default_rng(P * 7 + rec + sym)
    V = 70000
    E = 9_300_001 if P == 8 else 250_003
    edges = H.random_edges(rng, V, E, weighted=rec == 12, skew=True)
    path = tmp_path / "g.bin"
    edges.tofile(path)
    buf = eg.DeviceBuffer(E * rec)
    g = eg.Graph(V, P, S, 1, 20, rec)
    assert g.build_file(path, buf.ptr, E, sym) == E
    assert buf.download(edges.dtype, E).tobytes() == edges.tobytes()
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=20, symmetrize=sym, threads=8)
    from parity import assert_build_equal
    assert_build_equal(g, og)
    with pytest.raises(eg.EndGraphError):
        g.build_file(path, buf.ptr, E - 1, sym)  # capacity too small


@pytest.mark.parametrize("dtype", [np.float64, np.uint32, np.float32])
def test_vertex_array_dump_restore(eg, tmp_path, dtype):
    V = 9_876_543
    x = np.random.default_rng(1).random(V).astype(dtype) if dtype != np.uint32 else \
        np.random.default_rng(1).integers(0, 1 << 31, V, dtype=np.uint32)
    d = eg.DeviceBuffer.from_array(x)
    path = tmp_path / "a.bin"
    eg.dump_vertex_array(d.ptr, x.itemsize, V, path)
    assert path.read_bytes() == x.tobytes()  # the reference's format: raw V * sizeof(T)
    back = eg.DeviceBuffer(x.nbytes)
    eg.restore_vertex_array(back.ptr, x.itemsize, V, path)
    assert back.download(dtype, V).tobytes() == x.tobytes()
    with pytest.raises(eg.EndGraphError):
        eg.restore_vertex_array(back.ptr, x.itemsize, V + 1, path)


@pytest.mark.parametrize("nbytes,src_off,dst_off", [(0, 0, 0), (1, 0, 0), (16, 0, 0), (1 << 20, 0, 0),
                                                    ((1 << 24) + 13, 0, 0), (4099, 4, 8), (1000, 1, 3)])
def test_memcpy_d2d(eg, nbytes, src_off, dst_off):
    """Device copies: the 16-byte streaming kernel when both ends are aligned
    (with a byte tail), the runtime copy otherwise."""
    print(f"memcpy_d2d {nbytes} bytes, offsets {src_off}/{dst_off}", flush=True)
    rng = np.random.default_rng(nbytes)
    x = rng.integers(0, 256, nbytes + src_off, dtype=np.uint8)
    a = eg.DeviceBuffer.from_array(x) if x.size else eg.DeviceBuffer(1)
    b = eg.DeviceBuffer(nbytes + dst_off + 1)
    eg.memcpy_d2d(b.ptr + dst_off, a.ptr + src_off, nbytes)
    eg.synchronize()
    got = b.download(np.uint8, nbytes + dst_off)[dst_off:]
    assert got.tobytes() == x[src_off:].tobytes()
