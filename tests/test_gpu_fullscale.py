"""Full-size PageRank precision check on the twitter-shaped graph
(41,652,230 vertices, 1,468,365,182 edges, seed 3): compensated FP32 against
the FP64 oracle over the whole graph, 20 iterations, d = 0.85.

It needs ~80 GB of host memory and a few minutes of host CPU for the oracle,
so it runs only when EG_FULL_SCALE=1 (run it with -s: it prints each step).
"""
import os
import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("EG_FULL_SCALE") != "1", reason="set EG_FULL_SCALE=1")]

V, E = 41_652_230, 1_468_365_182
PR_FP32C_TOL = 1e-6   # relative L1 against the FP64 oracle
PR_FP64_TOL = 1e-12


def test_twitter_pagerank_l1_vs_fp64_oracle(oracle_lib):
    import endgraph_amd as eg
    t0 = time.time()
    say = lambda m: print(f"[fullscale {time.time() - t0:6.1f}s] {m}", flush=True)
    buf = eg.DeviceBuffer(E * 8)
    eg.gen_powerlaw(buf.ptr, E, V, seed=3)
    g = eg.Graph(V, 1, 1, 1, 20, 8)
    g.build(buf.ptr, E)
    say("GPU build done")
    x64, _ = g.pagerank(20, eg.PR_FP64)
    x32, _ = g.pagerank(20, eg.PR_FP32C)
    say("GPU PageRank FP64 and FP32C done")
    edges = buf.download(eg.REC8, E)
    del g, buf
    say("edge list on the host")
    threads = min(16, os.cpu_count() or 1)
    og = oracle_lib.OracleGraph(edges, V, P=1, S=1, eta=1, m=20, threads=threads)
    del edges
    say(f"oracle build done ({og.build_seconds:.1f} s on {threads} threads)")
    ref, _ = og.pagerank(20)
    say("oracle PageRank done")
    n1 = np.abs(ref).sum()
    l1_32 = float(np.abs(x32 - ref).sum() / n1)
    l1_64 = float(np.abs(x64 - ref).sum() / n1)
    say(f"relative L1 vs FP64 oracle: fp32c {l1_32:.3e}  fp64 {l1_64:.3e}")
    assert l1_64 <= PR_FP64_TOL
    assert l1_32 <= PR_FP32C_TOL
