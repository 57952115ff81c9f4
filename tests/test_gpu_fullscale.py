"""Every measured configuration at its full size, against the oracle (run with
-s: each step prints its elapsed time):

  C1  R-MAT scale 16 (65,536 vertices, 1,048,576 edges, seed 1) at P = 1:
      build bit-exact and 20 PageRank iterations (FP64 plain and
      degree-ordered, compensated FP32) with every delta(i);
  C2  R-MAT scale 24 (16,777,216 vertices, 268,435,456 edges, seed 2) at
      P = 8: every build array bit-exact, BFS from the max-out-degree vertex
      (reference and auto policies) bit-exact, and the symmetrised
      (load_undirected_from_directed) build and its CC labels bit-exact;
  C3  the twitter-shaped graph (41,652,230 vertices, 1,468,365,182 edges,
      seed 3) at P = 1: the build bit-exact and PageRank precision,
      compensated FP32 against the FP64 oracle over the whole graph,
      20 iterations, d = 0.85 (~110 GB of host memory);
  C4  R-MAT scale 22 with float weights U[0,1) (seed 4), 64 seeded roots:
      every column of the batched BFS and SSSP equal to the single-root
      program, and 4 BFS and 4 SSSP columns equal to the oracle; 256
      Erdos-Renyi graphs (|V| log-uniform in 1e3..1e6, degree U[8,32]) built
      in one batch, each equal to the oracle's CSR;
  C5  personalised PageRank gradients on R-MAT scale 22 (weights U[0.1,1]
      row-normalised, p Dirichlet(1) on 1024 seeds, c ~ N(0,1)) against FP64
      central differences on 64 sampled edges (fourth-order stencil, step 1%
      of the weight) and 64 sampled vertices (L is linear in p: step 1e-6).
"""
import os
import time

import numpy as np
import pytest

from parity import assert_build_equal

pytestmark = pytest.mark.gpu

PR_FP32C_TOL = 1e-6   # relative L1 against the FP64 oracle
PR_FP64_TOL = 1e-12


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def stopwatch(tag):
    t0 = time.time()
    return lambda m: print(f"[{tag} {time.time() - t0:6.1f}s] {m}", flush=True)


def test_c1_rmat16_pagerank(oracle_lib):
    import endgraph_amd as eg
    say = stopwatch("C1")
    V, E, iters = 1 << 16, 1 << 20, 20
    buf = eg.DeviceBuffer(E * 8)
    eg.gen_rmat(buf.ptr, E, 16, seed=1)
    g = eg.Graph(V, 1, 1, 1, iters, 8)
    g.build(buf.ptr, E)
    og = oracle_lib.OracleGraph(buf.download(eg.REC8, E), V, P=1, S=1, eta=1, m=iters, threads=host_threads())
    assert_build_equal(g, og)
    say("build bit-exact")
    ref, ref_d = og.pagerank(iters)
    n1 = np.abs(ref).sum()
    for prec, tol in ((eg.PR_FP64, PR_FP64_TOL), (eg.PR_FP64 | eg.PR_ORDERED, PR_FP64_TOL),
                      (eg.PR_FP32C, PR_FP32C_TOL)):
        x, d = g.pagerank(iters, prec)
        l1 = float(np.abs(x - ref).sum() / n1)
        say(f"pagerank precision {prec:#x}: relative L1 {l1:.3e}")
        assert l1 <= tol
        assert d[-1] == 0.0 and d[0] == 1.0
        np.testing.assert_allclose(d[:-1], ref_d[:-1], rtol=1e-9 if tol == PR_FP64_TOL else 1e-4, atol=1e-15)


def test_c2_rmat24_p8_build_bfs_cc(oracle_lib):
    import endgraph_amd as eg
    say = stopwatch("C2")
    V, E, P = 1 << 24, 1 << 28, 8
    T = host_threads()
    buf = eg.DeviceBuffer(E * 8)
    eg.gen_rmat(buf.ptr, E, 24, seed=2)
    g = eg.Graph(V, P, 1, 1, 1, 8)
    g.build(buf.ptr, E)
    root = int(np.argmax(g.out_degree()))
    bfs = {pol: g.bfs(root, pol) for pol in (eg.POLICY_REFERENCE, eg.POLICY_AUTO)}
    say(f"GPU build and BFS from {root} done")
    edges = buf.download(eg.REC8, E)
    og = oracle_lib.OracleGraph(edges, V, P=P, S=1, eta=1, m=1, threads=T)
    say(f"oracle build done ({og.build_seconds:.1f} s on {T} threads)")
    assert_build_equal(g, og)
    say("build bit-exact")
    o_parent, o_depth, o_levels = og.bfs(root, 1)
    for pol, (parent, depth, levels, _) in bfs.items():
        assert levels == o_levels
        assert np.array_equal(depth, o_depth), f"depth, policy {pol}"
        assert np.array_equal(parent, o_parent), f"parent, policy {pol}"
    say(f"BFS bit-exact ({o_levels} levels, {int((o_parent < V).sum())} reached)")
    del g, og, bfs, o_parent, o_depth
    gs = eg.Graph(V, P, 1, 1, 1, 8)
    gs.build(buf.ptr, E, True)
    label, it = gs.cc(eg.POLICY_REFERENCE)
    say(f"GPU symmetrised build and CC done ({it} iterations)")
    ogs = oracle_lib.OracleGraph(edges, V, P=P, S=1, eta=1, m=1, symmetrize=True, threads=T)
    del edges
    say(f"oracle symmetrised build done ({ogs.build_seconds:.1f} s)")
    assert_build_equal(gs, ogs)
    say("symmetrised build bit-exact")
    o_label, o_it = ogs.cc()
    assert np.array_equal(label, o_label)
    say(f"CC bit-exact ({np.unique(o_label).size} components; oracle {o_it} iterations)")


def test_c3_twitter_build_and_pagerank_l1_vs_fp64_oracle(oracle_lib):
    import endgraph_amd as eg
    say = stopwatch("C3")
    V, E = 41_652_230, 1_468_365_182
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
    threads = host_threads()
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


def _c4_roots(V):
    return np.random.default_rng(4).integers(0, V, 64).astype(np.uint32)


def test_c4_rmat22_batched_64_roots(oracle_lib):
    import endgraph_amd as eg
    say = stopwatch("C4")
    scale = 22
    V, E = 1 << scale, 16 << scale
    T = host_threads()
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=4, rec_bytes=12, w_lo=0.0, w_hi=1.0)
    g = eg.Graph(V, 1, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    roots = _c4_roots(V)
    par, dep, lev, edg, steps = g.bfs_multi(roots, eg.POLICY_AUTO)
    say(f"batched BFS done ({steps} steps)")
    for r, root in enumerate(roots):
        p1, d1, l1, e1 = g.bfs(int(root), eg.POLICY_AUTO)
        assert np.array_equal(par[:, r], p1) and np.array_equal(dep[:, r], d1), f"BFS column {r} (root {root})"
        assert lev[r] == l1 and edg[r] == e1
    say("every batched BFS column equals the single-root program")
    dist, _ = g.sssp_multi(roots, eg.POLICY_AUTO)
    for r, root in enumerate(roots):
        d1, _ = g.sssp(int(root), eg.POLICY_AUTO)
        assert np.array_equal(dist[:, r].view(np.uint32), d1.view(np.uint32)), f"SSSP column {r} (root {root})"
    say("every batched SSSP column equals the single-root program")
    e12 = buf.download(eg.REC12, E)
    del buf, g
    e8 = np.zeros(E, eg.REC8)
    e8["src"], e8["dst"] = e12["src"], e12["dst"]
    og8 = oracle_lib.OracleGraph(e8, V, P=1, S=1, eta=1, m=1, threads=T)
    del e8
    og12 = oracle_lib.OracleGraph(e12, V, P=1, S=1, eta=1, m=1, threads=T)
    del e12
    say("oracle builds done")
    for r in (0, 21, 42, 63):
        o_parent, o_depth, _ = og8.bfs(int(roots[r]), 1)
        assert np.array_equal(par[:, r], o_parent) and np.array_equal(dep[:, r], o_depth), f"BFS column {r}"
        o_dist, _ = og12.sssp(int(roots[r]))
        assert np.array_equal(dist[:, r].view(np.uint32), o_dist.view(np.uint32)), f"SSSP column {r}"
    say("BFS and SSSP columns 0, 21, 42, 63 equal the oracle")


def test_c4_er_256_batched_csr(oracle_lib):
    import endgraph_amd as eg
    say = stopwatch("C4-ER")
    rng = np.random.default_rng(256)
    n = 256
    nv = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n)).astype(np.uint32)
    ne = (nv.astype(np.float64) * rng.uniform(8, 32, n)).astype(np.uint64)
    total = int(ne.sum())
    buf = eg.DeviceBuffer(total * 8)
    eg.gen_er_batch(buf.ptr, nv, ne, seed=5)
    b = eg.BatchCSR(buf.ptr, nv, ne)
    say(f"batched build of {n} graphs ({int(nv.sum())} vertices, {total} edges) done")
    edges = buf.download(eg.REC8, total)
    del buf
    exp_list, exp_rp = oracle_lib.csr_batch(edges, nv, ne, threads=host_threads())
    del edges
    say("oracle CSRs done")
    eoff = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
    voff = np.concatenate([[0], np.cumsum(nv.astype(np.int64) + 1)])
    for i in range(n):
        lst, rp = b.get(i)
        assert np.array_equal(lst.view(np.uint8), exp_list[eoff[i]:eoff[i + 1]].view(np.uint8)), f"graph {i} list"
        assert np.array_equal(rp, exp_rp[voff[i]:voff[i + 1]]), f"graph {i} row pointer"
    say("every graph's CSR list and row pointer equal the oracle")


def test_c5_rmat22_ppr_gradients_vs_central_differences():
    import torch
    import endgraph_amd as eg
    from endgraph_amd import ppr
    say = stopwatch("C5")
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
