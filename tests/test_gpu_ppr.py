"""Personalised PageRank and its adjoint: the forward fixed point against a
dense FP64 solve, the adjoint gradients against the closed forms and against
FP64 central differences of the forward op, and the torch.autograd.Function
against torch's own gradcheck."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

D = 0.85
K = 250  # 0.85**250 ~ 2e-18: the K-step map equals the fixed point to FP64


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def eg():
    import endgraph_amd
    endgraph_amd.lib()
    return endgraph_amd


def _graph(eg, edges, V, P=1, S=1):
    buf = eg.DeviceBuffer.from_array(edges)
    g = eg.Graph(V, P, S, 1, 1, edges.dtype.itemsize)
    g.build(buf.ptr, edges.shape[0])
    return g, buf


def _setup(eg, torch, edges, V, seeds, rng, P=1, S=1):
    from endgraph_amd import ppr
    g, buf = _graph(eg, edges, V, P, S)
    src, dst, w0 = ppr.slots(g)
    w = ppr.row_normalized(src, w0, V)
    p = np.zeros(V)
    p[rng.choice(V, seeds, replace=False)] = rng.dirichlet(np.ones(seeds))
    return g, buf, src, dst, w, torch.tensor(p, device="cuda")


def _dense(V, src, dst, w):
    A = np.zeros((V, V))
    np.add.at(A, (dst, src), w)
    return A


@pytest.mark.parametrize("P,S", [(1, 1), (4, 2)])
def test_ppr_forward_and_adjoint_closed_form(eg, torch, P, S):
    from endgraph_amd import ppr
    rng = np.random.default_rng(7 + P)
    V = 300
    edges = H.random_edges(rng, V, 3000, weighted=True, skew=True)
    edges["w"] = 0.1 + 0.9 * edges["w"]
    g, _, src, dst, w, p = _setup(eg, torch, edges, V, 30, rng, P, S)
    x = ppr.personalized_pagerank(g, p, w, D, K).cpu().numpy()
    s, d_, wn = src.cpu().numpy(), dst.cpu().numpy(), w.cpu().numpy()
    A = _dense(V, s, d_, wn)
    x_ref = np.linalg.solve(np.eye(V) - D * A, (1 - D) * p.cpu().numpy())
    assert np.abs(x - x_ref).sum() / np.abs(x_ref).sum() < 1e-12
    c = rng.standard_normal(V)
    lam = np.linalg.solve(np.eye(V) - D * A.T, c)
    gp = torch.empty(V, dtype=torch.float64, device="cuda")
    gw = torch.empty(w.numel(), dtype=torch.float64, device="cuda")
    xt = torch.tensor(x, device="cuda")
    ct = torch.tensor(c, device="cuda")
    g.ppr_adjoint(w.data_ptr(), xt.data_ptr(), ct.data_ptr(), D, K, gp.data_ptr(), gw.data_ptr())
    np.testing.assert_allclose(gp.cpu().numpy(), (1 - D) * lam, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(gw.cpu().numpy(), D * lam[d_] * x[s], rtol=1e-10, atol=1e-15)


def test_ppr_gradients_vs_central_differences(eg, torch):
    """64 sampled edge weights and 64 sampled personalisation entries of an
    R-MAT scale-14 graph (weights U[0.1,1] row-normalised, p Dirichlet(1) on
    1024 seeds, loss sum c_i x_i with c ~ N(0,1)) against FP64 central
    differences of the forward op."""
    from endgraph_amd import ppr
    rng = np.random.default_rng(5)
    scale = 14
    V, E = 1 << scale, 16 << scale
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=5, rec_bytes=12, w_lo=0.1, w_hi=1.0)
    edges = buf.download(eg.REC12, E)
    g, _, src, dst, w, p = _setup(eg, torch, edges, V, 1024, rng)
    c = torch.tensor(rng.standard_normal(V), device="cuda")
    w.requires_grad_(True)
    p.requires_grad_(True)
    loss = (c * ppr.personalized_pagerank(g, p, w, D, K)).sum()
    loss.backward()

    def L(pp, ww):
        with torch.no_grad():
            return float((c * ppr.personalized_pagerank(g, pp, ww, D, K)).sum())

    ei = rng.choice(w.numel(), 64, replace=False)
    vi = rng.choice(V, 64, replace=False)
    fd_w, fd_p = [], []
    with torch.no_grad():
        w0, p0 = w.detach().clone(), p.detach().clone()
        for e in ei:
            h = 1e-4 * float(w0[e])  # x is rational in w: O(h^2) truncation stays below the FP64 noise
            wp, wm = w0.clone(), w0.clone()
            wp[e] += h
            wm[e] -= h
            fd_w.append((L(p0, wp) - L(p0, wm)) / (2 * h))
        for v in vi:
            h = 1e-6  # x is linear in p
            pp, pm = p0.clone(), p0.clone()
            pp[v] += h
            pm[v] -= h
            fd_p.append((L(pp, w0) - L(pm, w0)) / (2 * h))
    gw = w.grad.cpu().numpy()[ei]
    gp = p.grad.cpu().numpy()[vi]
    err_w = np.abs(np.array(fd_w) - gw).max() / np.abs(gw).max()
    err_p = np.abs(np.array(fd_p) - gp).max() / np.abs(gp).max()
    print(f"max rel error vs central differences: edges {err_w:.2e}, vertices {err_p:.2e}", flush=True)
    assert err_w < 1e-6 and err_p < 1e-6


def test_ppr_torch_gradcheck(eg, torch):
    from endgraph_amd import ppr
    rng = np.random.default_rng(9)
    V = 40
    edges = H.random_edges(rng, V, 160, weighted=True, skew=False)
    g, _, src, dst, w, p = _setup(eg, torch, edges, V, 8, rng)
    w = w.detach().requires_grad_(True)
    p = p.detach().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda pp, ww: ppr.personalized_pagerank(g, pp, ww, D, 120), (p, w),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)


def test_ppr_argument_errors(eg, torch):
    from endgraph_amd import ppr
    rng = np.random.default_rng(2)
    V = 50
    edges = H.random_edges(rng, V, 200, weighted=True)
    g, _, src, dst, w, p = _setup(eg, torch, edges, V, 5, rng)
    with pytest.raises(TypeError):
        ppr.personalized_pagerank(g, p.float(), w)
    with pytest.raises(ValueError):
        ppr.personalized_pagerank(g, p[:-1], w)
    with pytest.raises(eg.EndGraphError):
        g.ppr_forward(w.data_ptr(), p.data_ptr(), 1.5, 10, p.data_ptr())
