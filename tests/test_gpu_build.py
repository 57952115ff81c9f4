"""GPU parity of the build: every array bit-exact against the oracle."""
import numpy as np
import pytest

import helpers as H
from parity import assert_build_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _gpu_build(eg, edges, V, P, S, eta, m, sym):
    rec = edges.dtype.itemsize
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, eta, m, rec)
    g.build(buf.ptr, edges.shape[0], sym)
    return g, buf


@pytest.mark.parametrize("P,S,sym,weighted", [
    (1, 1, False, False), (1, 2, False, False), (2, 1, False, True), (3, 1, True, False),
    (8, 1, False, False), (3, 2, True, True), (8, 3, False, False), (5, 1, False, False),
])
def test_build_random_skewed(eg, oracle_lib, P, S, sym, weighted):
    print(f"build random P={P} S={S} sym={sym} weighted={weighted}", flush=True)
    rng = np.random.default_rng(1000 + P * 10 + S)
    V = 3 * 4096 + 1000 if P < 8 else 12 * 4096 + 77
    edges = H.random_edges(rng, V, 50000 + 13, weighted=weighted)
    g, _ = _gpu_build(eg, edges, V, P, S, 3, 20, sym)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=3, m=20, symmetrize=sym, threads=8)
    assert_build_equal(g, og)


@pytest.mark.parametrize("bits", [8, 9])
@pytest.mark.parametrize("weighted", [False, True])
def test_build_four_digit_passes(eg, oracle_lib, weighted, bits):
    """More than 2^24 vertices at P = 3: 25-bit keys take four passes of at
    most 8 bits, so the partition-local sorts pass their field arrays through
    both temporaries twice before writing the records; with 9-bit digits
    they take three passes (9, 9, 7 bits)."""
    print(f"build V > 2^24 P=3 weighted={weighted} bits={bits}", flush=True)
    rng = np.random.default_rng(2024)
    V = (1 << 24) + 5003
    edges = H.random_edges(rng, V, 300_007, weighted=weighted)
    keep = eg.get_sort_variant(edges.dtype.itemsize)
    eg.set_sort_variant(edges.dtype.itemsize, 2, 512, bits)
    try:
        g, _ = _gpu_build(eg, edges, V, 3, 1, 1, 1, False)
        og = oracle_lib.OracleGraph(edges, V, P=3, S=1, eta=1, m=1, threads=8)
        assert_build_equal(g, og)
    finally:
        eg.set_sort_variant(edges.dtype.itemsize, *keep)


@pytest.mark.parametrize("scale,P,sym", [(14, 1, False), (16, 1, False), (16, 8, False), (15, 4, True)])
def test_build_rmat(eg, oracle_lib, scale, P, sym):
    print(f"build rmat scale={scale} P={P} sym={sym}", flush=True)
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * 8)
    eg.gen_rmat(buf.ptr, E, scale, seed=1)
    edges = buf.download(eg.REC8, E)
    g = eg.Graph(V, P, 1, 1, 20, 8)
    g.build(buf.ptr, E, sym)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=1, eta=1, m=20, symmetrize=sym, threads=8)
    assert_build_equal(g, og)


def test_build_edge_cases(eg, oracle_lib):
    rng = np.random.default_rng(5)
    # fewer edges than ranks, V = 1, self loops only, all edges in one page
    cases = []
    e = np.zeros(3, H.REC8); e["src"] = [0, 1, 2]; e["dst"] = [2, 2, 0]
    cases.append((e, 5000, 4, 2, False))
    e = np.zeros(7, H.REC8)
    cases.append((e, 1, 1, 1, True))
    e = np.zeros(400, H.REC8); e["src"] = rng.integers(0, 50, 400); e["dst"] = e["src"]
    cases.append((e, 9 * 4096 + 5, 3, 3, False))
    e = H.random_edges(rng, 70000, 4000, skew=False)
    cases.append((e, 70000, 8, 1, True))
    for edges, V, P, S, sym in cases:
        print(f"edge case V={V} E={edges.shape[0]} P={P} S={S} sym={sym}", flush=True)
        g, _ = _gpu_build(eg, edges, V, P, S, 2, 1, sym)
        og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=2, m=1, symmetrize=sym, threads=4)
        assert_build_equal(g, og)


@pytest.mark.parametrize("E", [40 * 512, 40 * 512 - 64, 40 * 512 + 2, 512])
@pytest.mark.parametrize("P", [1, 2, 8])
def test_build_runs_ending_at_a_chunk_end(eg, oracle_lib, E, P):
    """The lists sorted by source and by destination end in a run of several
    records of the top vertex, and the list ends on the last lane of a
    512-record wave chunk (E a multiple of 512), on the last lane of an
    earlier row, or two records into a chunk: each record of the last run
    counts once in the degrees, and the run ends only at the list's end."""
    print(f"build runs at a chunk end E={E} P={P}", flush=True)
    rng = np.random.default_rng(E + P)
    V = 5000
    edges = H.random_edges(rng, V - 1, E, skew=False)  # the top vertex only where set below
    k = min(E // 4, 37)
    edges["src"][-k:] = V - 1
    edges["dst"][:k] = V - 1
    g, _ = _gpu_build(eg, edges, V, P, 1, 1, 1, False)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=1, eta=1, m=1, threads=4)
    assert g.out_degree()[V - 1] == k and g.in_degree()[V - 1] == k
    assert_build_equal(g, og)


def test_build_rejects_out_of_range_ids(eg):
    e = np.zeros(10, H.REC8)
    e["src"] = np.arange(10)
    e["dst"] = 10  # == V
    buf = eg.DeviceBuffer.from_array(e)
    g = eg.Graph(10, 1, 1, 0, 1, 8)
    with pytest.raises(eg.EndGraphError) as ei:
        g.build(buf.ptr, 10)
    assert ei.value.status == -2


def test_rebuild_reuses_handle(eg, oracle_lib):
    rng = np.random.default_rng(9)
    V = 20000
    g = eg.Graph(V, 2, 1, 1, 1, 8)
    for E in (30000, 10000):  # second build smaller than the first
        edges = H.random_edges(rng, V, E)
        buf = eg.DeviceBuffer.from_array(edges)
        g.build(buf.ptr, E)
        og = oracle_lib.OracleGraph(edges, V, P=2, S=1, eta=1, m=1, threads=4)
        assert_build_equal(g, og)


VARIANTS = [(256, 8), (512, 8), (512, 9), (1024, 8)]


@pytest.mark.parametrize("scheme", [0, 1, 2])
@pytest.mark.parametrize("threads,bits", VARIANTS)
@pytest.mark.parametrize("rec", [8, 12])
def test_build_every_sort_variant(eg, oracle_lib, scheme, threads, bits, rec):
    if rec == 12 and threads == 1024:
        pytest.skip("12-byte records use at most 512 threads (LDS)")
    print(f"sort variant scheme={scheme} threads={threads} bits={bits} rec={rec}", flush=True)
    rng = np.random.default_rng(77)
    # 9-bit digits are taken only where they save a pass: 18-bit keys (2
    # passes instead of 3)
    V = 9 * 4096 + 100 if bits == 8 else 200_000 + 17
    edges = H.random_edges(rng, V, 70001, weighted=(rec == 12))
    keep = eg.get_sort_variant(rec)
    eg.set_sort_variant(rec, scheme, threads, bits)
    try:
        for P, S, sym in ((1, 1, False), (4, 2, True)):
            g, _ = _gpu_build(eg, edges, V, P, S, 2, 5, sym)
            og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=2, m=5, symmetrize=sym, threads=8)
            assert_build_equal(g, og)
    finally:
        eg.set_sort_variant(rec, *keep)


def test_autotune_picks_a_variant_and_stays_exact(eg, oracle_lib):
    rng = np.random.default_rng(78)
    V = 50000
    edges = H.random_edges(rng, V, 200000)
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, 3, 1, 1, 1, 8)
    keep = eg.get_sort_variant(8)
    try:
        scheme, threads, bits, ms = g.autotune_sort(buf.ptr, edges.shape[0])
        assert scheme in (0, 1, 2) and threads in (256, 512, 1024) and bits in (8, 9) and ms > 0
        og = oracle_lib.OracleGraph(edges, V, P=3, S=1, eta=1, m=1, threads=8)
        assert_build_equal(g, og)
    finally:
        eg.set_sort_variant(8, *keep)
