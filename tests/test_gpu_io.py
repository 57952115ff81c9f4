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


@pytest.mark.parametrize("weighted,E", [(False, 0), (False, 1), (False, 12_345_678), (True, 9_000_001)])
def test_edge_file_round_trip(eg, tmp_path, weighted, E):
    """Sizes below, at and across several 64 MiB chunks (odd chunk tails)."""
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
