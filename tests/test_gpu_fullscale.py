"""Full-size checks, run only when EG_FULL_SCALE=1 (run with -s: each step
prints):
  * the twitter-shaped graph (41,652,230 vertices, 1,468,365,182 edges,
    seed 3) at P = 1: the GPU build bit-exact against the oracle's build
    (degrees, CSC and CSR lists, compressed CSC index, CSR index and bitmap),
    and PageRank precision: compensated FP32 against the FP64 oracle over the
    whole graph, 20 iterations, d = 0.85 (~110 GB of host memory, a few
    minutes of host CPU for the oracle);
  * personalised PageRank gradients on R-MAT scale 22 (weights U[0.1,1]
    row-normalised, p Dirichlet(1) on 1024 seeds, c ~ N(0,1)) against FP64
    central differences on 64 sampled edges (fourth-order stencil, step 1% of
    the weight) and 64 sampled vertices (L is linear in p: step 1e-6).
"""
import os
import time

import numpy as np
import pytest

from parity import assert_build_equal

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
    del buf
    say("edge list on the host")
    threads = min(16, os.cpu_count() or 1)
    og = oracle_lib.OracleGraph(edges, V, P=1, S=1, eta=1, m=20, threads=threads)
    del edges
    say(f"oracle build done ({og.build_seconds:.1f} s on {threads} threads)")
    assert_build_equal(g, og)
    del g
    say("GPU build bit-exact against the oracle build")
    ref, _ = og.pagerank(20)
    say("oracle PageRank done")
    n1 = np.abs(ref).sum()
    l1_32 = float(np.abs(x32 - ref).sum() / n1)
    l1_64 = float(np.abs(x64 - ref).sum() / n1)
    say(f"relative L1 vs FP64 oracle: fp32c {l1_32:.3e}  fp64 {l1_64:.3e}")
    assert l1_64 <= PR_FP64_TOL
    assert l1_32 <= PR_FP32C_TOL


def test_rmat22_ppr_gradients_vs_central_differences():
    import torch
    import endgraph_amd as eg
    from endgraph_amd import ppr
    t0 = time.time()
    say = lambda m: print(f"[fullscale {time.time() - t0:6.1f}s] {m}", flush=True)
    rng = np.random.default_rng(22)
    scale, D, K = 22, 0.85, 250
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=22, rec_bytes=12, w_lo=0.1, w_hi=1.0)
    g = eg.Graph(V, 1, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    src, dst, w0 = ppr.slots(g)
    w = ppr.row_normalized(src, w0, V).requires_grad_(True)
    p_np = np.zeros(V)
    p_np[rng.choice(V, 1024, replace=False)] = rng.dirichlet(np.ones(1024))
    p = torch.tensor(p_np, device="cuda", requires_grad=True)
    c = torch.tensor(rng.standard_normal(V), device="cuda")
    (c * ppr.personalized_pagerank(g, p, w, D, K)).sum().backward()
    say("forward + adjoint done")

    def L(pp, ww):
        with torch.no_grad():
            return float((c * ppr.personalized_pagerank(g, pp, ww, D, K)).sum())

    ei = rng.choice(w.numel(), 64, replace=False)
    vi = rng.choice(V, 64, replace=False)
    fd_w, fd_p = [], []
    with torch.no_grad():
        w_, p_ = w.detach().clone(), p.detach().clone()
        # L is not linear in w: a fourth-order central stencil with h = 1% of
        # the weight keeps the truncation error far below the tolerance,
        # while a much smaller h lets FP64 rounding noise of the 250-step
        # forward (about 1e-19 per rank) dominate, divided by h
        for e in ei:
            h = 1e-2 * float(w_[e])
            f = {}
            for k in (-2, -1, 1, 2):
                a = w_.clone()
                a[e] += k * h
                f[k] = L(p_, a)
            fd_w.append((-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h))
        say("edge differences done")
        for v in vi:
            a, b = p_.clone(), p_.clone()
            a[v] += 1e-6
            b[v] -= 1e-6
            fd_p.append((L(a, w_) - L(b, w_)) / 2e-6)
    gw = w.grad.cpu().numpy()[ei]
    gp = p.grad.cpu().numpy()[vi]
    err_w = np.abs(np.array(fd_w) - gw).max() / np.abs(gw).max()
    err_p = np.abs(np.array(fd_p) - gp).max() / np.abs(gp).max()
    say(f"max rel error vs central differences: edges {err_w:.2e}, vertices {err_p:.2e}")
    assert err_w < 1e-6 and err_p < 1e-6
