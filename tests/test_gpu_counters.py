"""The reference's byte counters (Graph::prep_data / exe_data,
core/graph.hpp:128-129, printed by the toolkits as "algo data / pre data"):
eg_prep_data against the records each rank slice would send in the two
shuffles (core/graph.hpp:906-908), eg_dense_messages against the dense
signals of toolkits/*.cpp evaluated row by row on the oracle's CSC, and
eg_mode_sparse against the mode rule (core/graph.hpp:1483-1498)."""
import os
import re
import subprocess

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def numpy_prep_bytes(edges, V, P, sym, rec, part_off):
    """2 degree allreduces of V x 4 B per rank, then per shuffle (by source,
    by destination) the records of rank slice r owned by another partition;
    slice r is records [r * (E0 // P), ...) of the input (the last takes the
    rest), followed by its reversed copy when symmetrised."""
    total = 2 * P * V * 4
    if P == 1:
        return total
    E0 = edges.shape[0]
    src = edges["src"].astype(np.int64)
    dst = edges["dst"].astype(np.int64)
    owner = lambda v: np.searchsorted(part_off, v, side="right") - 1
    per = E0 // P
    for r in range(P):
        lo, hi = r * per, (E0 if r == P - 1 else (r + 1) * per)
        s, d = src[lo:hi], dst[lo:hi]
        if sym:
            s, d = np.concatenate([s, d]), np.concatenate([d, s])
        total += rec * (int(np.count_nonzero(owner(s) != r)) + int(np.count_nonzero(owner(d) != r)))
    return total


def numpy_dense_messages(og, P, S, part_off, program, active=None, visited=None, label=None, dist=None):
    """Rows (sub-partition, destination) of the oracle's CSC whose destination
    another partition owns and whose dense signal emits."""
    bit = lambda bm, v: (bm[v >> 6] >> (v & 63).astype(np.uint64)) & np.uint64(1)
    n = 0
    for p in range(P):
        for s in range(S):
            recs, ci = og.csc(p, s)
            if ci.shape[0] < 2:
                continue
            starts = ci["index"][:-1].astype(np.int64)
            dsts = ci["vertex"][:-1].astype(np.int64)
            cross = (np.searchsorted(part_off, dsts, side="right") - 1) != p
            src = recs["src"].astype(np.int64)
            if program == "pagerank":
                emits = np.ones(dsts.shape, bool)
            elif program == "bfs":
                any_active = np.maximum.reduceat(bit(active, src).astype(np.int64), starts) > 0
                emits = any_active & (bit(visited, dsts) == 0)
            elif program == "cc":
                emits = np.minimum.reduceat(label[src].astype(np.int64), starts) < dsts
            else:
                relax = (dist[src] + recs["w"]).astype(np.float32)
                emits = np.minimum.reduceat(relax, starts) < np.float32(1e9)
            n += int(np.count_nonzero(cross & emits))
    return n


def _bitmap(rng, V, p):
    bits = rng.random(V) < p
    words = np.zeros((V >> 6) + 1, np.uint64)
    idx = np.flatnonzero(bits)
    np.bitwise_or.at(words, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    return words


@pytest.mark.parametrize("P,S,sym,weighted", [(1, 1, False, False), (3, 1, False, True), (4, 2, True, False),
                                              (8, 1, False, False)])
def test_prep_data(eg, oracle_lib, P, S, sym, weighted):
    print(f"prep_data P={P} S={S} sym={sym} weighted={weighted}", flush=True)
    rng = np.random.default_rng(40 + P)
    V = 6000
    edges = H.random_edges(rng, V, 50000, weighted=weighted)
    buf = eg.DeviceBuffer.from_array(edges)
    rec = edges.dtype.itemsize
    g = eg.Graph(V, P, S, 1, 1, rec)
    g.build(buf.ptr, edges.shape[0], sym)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=1, symmetrize=sym, threads=8)
    assert g.prep_data() == numpy_prep_bytes(edges, V, P, sym, rec, og.partition_offset().astype(np.int64))


@pytest.mark.parametrize("P,S", [(1, 1), (3, 1), (2, 2)])
@pytest.mark.parametrize("program", ["pagerank", "bfs", "cc", "sssp"])
def test_dense_messages(eg, oracle_lib, P, S, program):
    print(f"dense messages {program} P={P} S={S}", flush=True)
    rng = np.random.default_rng(7 * P + S)
    V = 5000
    weighted = program == "sssp"
    edges = H.random_edges(rng, V, 40000, weighted=weighted)
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, 1, 1, edges.dtype.itemsize)
    g.build(buf.ptr, edges.shape[0])
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=1, threads=8)
    part_off = og.partition_offset().astype(np.int64)
    kw, dev = {}, {}
    if program == "bfs":
        kw = {"active": _bitmap(rng, V, 0.05), "visited": _bitmap(rng, V, 0.3)}
        dev = {"active": eg.DeviceBuffer.from_array(kw["active"]), "visited": eg.DeviceBuffer.from_array(kw["visited"])}
    elif program == "cc":
        kw = {"label": rng.integers(0, V, V).astype(np.uint32)}
        dev = {"values": eg.DeviceBuffer.from_array(kw["label"])}
    elif program == "sssp":
        dist = rng.random(V, dtype=np.float32) * 4
        dist[rng.random(V) < 0.6] = np.float32(1e9)
        kw = {"dist": dist}
        dev = {"values": eg.DeviceBuffer.from_array(dist)}
    code = {"pagerank": eg.PROGRAM_PAGERANK, "bfs": eg.PROGRAM_BFS, "cc": eg.PROGRAM_CC, "sssp": eg.PROGRAM_SSSP}
    got = g.dense_messages(code[program], **dev)
    exp = numpy_dense_messages(og, P, S, part_off, program, **kw)
    assert got == exp
    if P == 1:
        assert got == 0
    else:
        assert got > 0, "the case should send messages"


def test_mode_sparse_rule(eg):
    edges = H.random_edges(np.random.default_rng(2), 1000, 20000)
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(1000, 2, 1, 1, 1, 8)
    g.build(buf.ptr, edges.shape[0])
    E = g.edges_num
    for ae in (0, E // 20 - 1, E // 20, E):
        assert not g.mode_sparse(eg.POLICY_REFERENCE, eg.PROGRAM_PAGERANK, ae)
        for prog in (eg.PROGRAM_BFS, eg.PROGRAM_SSSP):
            assert g.mode_sparse(eg.POLICY_REFERENCE, prog, ae)
            assert g.mode_sparse(eg.POLICY_AUTO, prog, ae) == (ae < E // 20)
        assert g.mode_sparse(eg.POLICY_REFERENCE, eg.PROGRAM_CC, ae) == (ae < E // 20)
        assert g.mode_sparse(eg.POLICY_PUSH, eg.PROGRAM_CC, ae) and not g.mode_sparse(eg.POLICY_PULL,
                                                                                       eg.PROGRAM_BFS, ae)


def test_toolkit_prints_algo_and_pre_data(eg, tmp_path):
    """pagerank.cpp:127's printout; at P = 2 the PageRank run's algo data is
    the iterations times 12 B per cross-partition row."""
    rng = np.random.default_rng(9)
    V = 200000
    edges = H.random_edges(rng, V, 2_000_000)
    f = tmp_path / "g.bin"
    edges.tofile(f)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "endgraph_amd", "toolkits",
                       "_build", "pagerank")
    iters = 10
    p = subprocess.run([exe, str(f), str(V), str(iters), "1", "2", "1"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    got = [(int(a), int(b)) for a, b in re.findall(r"algo data: (\d+) MB, pre data: (\d+) MB", p.stdout)]
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, 2, 1, 1, iters, 8)
    g.build(buf.ptr, edges.shape[0])
    cross = g.dense_messages(eg.PROGRAM_PAGERANK)
    assert got and all(a == (iters * 12 * cross) >> 20 and b == g.prep_data() >> 20 for a, b in got)
    assert got[0][0] > 0 and got[0][1] > 0
