"""Differentiable personalised PageRank on a built Graph.

``personalized_pagerank(graph, p, w)`` returns x with
x = d * A x + (1 - d) * p (``iterations`` fixed-point steps from x = p), where
A[v][u] sums the weights ``w`` of the edge slots u -> v.  Gradients with
respect to ``p`` and ``w`` come from the hand-written adjoint kernel
(eg_ppr_adjoint): lambda = c + d * A^T lambda for dL/dx = c, then
dL/dp = (1 - d) * lambda and dL/dw_e = d * lambda[dst_e] * x[src_e].

Tensors live on the GPU in float64; ``w`` is in slot order (``slots``).  The
library runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import DeviceBuffer, Graph, check, lib
from . import _lib as _loaded

if _loaded is not None:
    from . import lib_shares_torch_runtime as _shared

    if not _shared:
        raise ImportError("endgraph_amd's native library was loaded before torch, so the process holds two HIP "
                          "runtimes and torch sees no GPU: import torch (or endgraph_amd.ppr) before the first "
                          "endgraph_amd call")

__all__ = ["slots", "row_normalized", "personalized_pagerank", "PersonalizedPageRank"]


def _stream():
    return C.c_void_p(int(torch.cuda.current_stream().cuda_stream))


def _to_torch(buf: DeviceBuffer, n: int, dtype, np_dtype) -> torch.Tensor:
    out = torch.empty(n, dtype=dtype, device="cuda")
    if n:
        check(lib().eg_memcpy_d2d(C.c_void_p(out.data_ptr()), C.c_void_p(buf.ptr), n * np.dtype(np_dtype).itemsize,
                                  _stream()))
        torch.cuda.current_stream().synchronize()
    return out


def slots(graph: Graph):
    """(src, dst, record weight) of every edge slot as int64 / float64 CUDA tensors."""
    E = graph.edges_num
    s, d, w = graph.ppr_slots(stream=int(torch.cuda.current_stream().cuda_stream))
    src = _to_torch(s, E, torch.int32, np.uint32).to(torch.int64)
    dst = _to_torch(d, E, torch.int32, np.uint32).to(torch.int64)
    wt = _to_torch(w, E, torch.float32, np.float32).to(torch.float64)
    return src, dst, wt


def row_normalized(src: torch.Tensor, w: torch.Tensor, V: int) -> torch.Tensor:
    """w divided by the weight sum of its source's out-slots (sources without slots keep none)."""
    tot = torch.zeros(V, dtype=w.dtype, device=w.device).index_add_(0, src, w)
    return w / tot[src]


class PersonalizedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, w, graph: Graph, damping: float, iterations: int):
        if p.dtype != torch.float64 or w.dtype != torch.float64 or not p.is_cuda or not w.is_cuda:
            raise TypeError("p and w must be float64 CUDA tensors")
        if p.numel() != graph.V or w.numel() != graph.edges_num:
            raise ValueError(f"need p of {graph.V} and w of {graph.edges_num} elements")
        p = p.contiguous()
        w = w.contiguous()
        x = torch.empty_like(p)
        check(lib().eg_ppr_forward(graph._h, C.c_void_p(w.data_ptr()), C.c_void_p(p.data_ptr()), float(damping),
                                   int(iterations), C.c_void_p(x.data_ptr()), _stream()))
        ctx.save_for_backward(w, x)
        ctx.graph, ctx.damping, ctx.iterations = graph, float(damping), int(iterations)
        return x

    @staticmethod
    def backward(ctx, gx):
        w, x = ctx.saved_tensors
        c = gx.contiguous().to(torch.float64)
        gp = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        gw = torch.empty_like(w) if ctx.needs_input_grad[1] else None
        check(lib().eg_ppr_adjoint(ctx.graph._h, C.c_void_p(w.data_ptr()), C.c_void_p(x.data_ptr()),
                                   C.c_void_p(c.data_ptr()), ctx.damping, ctx.iterations,
                                   C.c_void_p(gp.data_ptr() if gp is not None else None),
                                   C.c_void_p(gw.data_ptr() if gw is not None else None), _stream()))
        return gp, gw, None, None, None


def personalized_pagerank(graph: Graph, p: torch.Tensor, w: torch.Tensor, damping: float = 0.85,
                          iterations: int = 250) -> torch.Tensor:
    return PersonalizedPageRank.apply(p, w, graph, damping, iterations)
