// Personalized PageRank with per-edge weights and its adjoint (no reference
// counterpart: toolkits/pagerank.cpp is unweighted and not differentiable).
//
//   forward   x_{k+1} = d * A x_k + (1 - d) * p,  x_0 = p, K iterations, where
//             A[v][u] = sum of w_e over the edge slots e = (u -> v).  The
//             product A x is the engine's pull (rows of the CSC already built,
//             rowreduce.hpp), in FP64.
//   adjoint   for a loss L with dL/dx = c:  lambda_{k+1} = c + d * A^T lambda_k,
//             lambda_0 = c, K iterations; then
//               dL/dp   = (1 - d) * lambda
//               dL/dw_e = d * lambda[dst_e] * x[src_e]
//             (differentiating x = d A x + (1 - d) p at the fixed point).  The
//             product A^T lambda is a pull over the out-edges of each source:
//             a transposed view of the same CSC slots (slots sorted by source,
//             each row gathering lambda at the slots' destinations), so no
//             atomics are needed and the sums are deterministic.
//
// Edge slots are the engine's CSC order: the sub-partitions' incoming lists
// back to back; eg_ppr_slots exposes (src, dst, record weight) per slot.
#include <algorithm>
#include <vector>

#include "graph_state.hpp"
#include "rowreduce.hpp"

namespace eg {
namespace {

constexpr int kNT = 256;

unsigned grid_for(uint64_t n, uint64_t per = kNT, uint64_t cap = 65536) {
  uint64_t b = ceil_div(n, per);
  return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

template <class T>
__global__ void k_fill_v(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// (src, dst, record weight) of every CSC slot
template <class R>
__global__ void k_slots(const R* __restrict__ lst, uint64_t E, uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                        float* __restrict__ w) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    const R r = lst[e];
    if (src) src[e] = r.src;
    if (dst) dst[e] = r.dst;
    if (w) {
      if constexpr (sizeof(R) == 12)
        w[e] = reinterpret_cast<const Rec12&>(r).w;
      else
        w[e] = 1.0f;
    }
  }
}

// ------------------------------------------------------- transposed view --
// record (slot e, source of e): a stable sort by source groups the slots of
// every source, in slot order
template <class R>
__global__ void k_ad_records(const R* __restrict__ lst, uint64_t E, Rec8* __restrict__ out) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x)
    out[e] = Rec8{uint32_t(e), lst[e].src};
}

struct KeyDst {
  __device__ __forceinline__ uint64_t operator()(const Rec8& r) const { return r.dst; }
};

// ad_slot[k] = slot, ad_dst[k] = destination of that slot, row-start bit
// when the source changes
__global__ __launch_bounds__(kNT) void k_ad_soa(const Rec8* __restrict__ s, uint64_t E,
                                                const uint32_t* __restrict__ slot_dst, uint32_t* __restrict__ ad_slot,
                                                uint32_t* __restrict__ ad_dst, uint64_t* __restrict__ rowflag) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t wd = gw; wd < nwords; wd += nw) {
    const uint64_t k = wd * 64 + lane;
    bool head = false;
    if (k < E) {
      const Rec8 r = s[k];
      ad_slot[k] = r.src;
      ad_dst[k] = slot_dst[r.src];
      head = k == 0 || s[k - 1].dst != r.dst;
    }
    const uint64_t m = __ballot(head);
    if (lane == 0) rowflag[wd] = m;
  }
}

__global__ void k_popc(const uint64_t* __restrict__ w, uint64_t n, uint32_t* __restrict__ c) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    c[i] = uint32_t(__popcll(w[i]));
}

__global__ __launch_bounds__(kNT) void k_ad_rows(const Rec8* __restrict__ s, uint64_t E,
                                                 const uint64_t* __restrict__ rowflag,
                                                 const uint32_t* __restrict__ wordrank, uint32_t* __restrict__ row_src) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t nwords = ceil_div(E, 64);
  const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / kWave;
  for (uint64_t wd = gw; wd < nwords; wd += nw) {
    const uint64_t f = rowflag[wd];
    if ((f >> lane) & 1ull) row_src[wordrank[wd] + uint32_t(__popcll(f & ((1ull << lane) - 1)))] = s[wd * 64 + lane].dst;
  }
}

template <class R>
int adjoint_impl(eg_graph* g, hipStream_t s) {
  const uint64_t E = g->E;
  const uint64_t nwords = ceil_div(E, 64);
  EG_TRY(g->ad_slot_dst.resize(std::max<uint64_t>(E, 1) * 4));
  EG_TRY(g->ad_slot.resize(std::max<uint64_t>(E, 1) * 4));
  EG_TRY(g->ad_dst.resize(std::max<uint64_t>(E, 1) * 4));
  EG_TRY(g->ad_rowflag.resize((nwords + 1) * 8));
  EG_TRY(g->ad_wordrank.resize((nwords + 2) * 4));
  const R* lst = g->csc_list.as<R>();
  uint32_t* wordrank = g->ad_wordrank.as<uint32_t>();
  uint64_t* rowflag = g->ad_rowflag.as<uint64_t>();
  DevBuf a, b, c;
  if (E) {
    EG_TRY(a.resize(E * 8));
    EG_TRY(b.resize(E * 8));
    EG_TRY(c.resize(E * 8));
    hipLaunchKernelGGL((k_slots<R>), dim3(grid_for(E)), dim3(kNT), 0, s, lst, E, (uint32_t*)nullptr,
                       g->ad_slot_dst.as<uint32_t>(), (float*)nullptr);
    hipLaunchKernelGGL((k_ad_records<R>), dim3(grid_for(E)), dim3(kNT), 0, s, lst, E, a.as<Rec8>());
    EG_HIP(hipGetLastError());
    RxWork w;
    EG_TRY(w.resize_segs(1));
    EG_HIP(hipMemsetAsync(w.err.ptr, 0, 4, s));
    EG_TRY(rx_sort((const Rec8*)a.as<Rec8>(), b.as<Rec8>(), c.as<Rec8>(), std::vector<uint64_t>{0, E}, KeyDst{},
                   std::max(1, counting_bits(g->V - 1)), w, s));
    hipLaunchKernelGGL(k_ad_soa, dim3(grid_for(nwords, kNT / kWave)), dim3(kNT), 0, s, (const Rec8*)b.as<Rec8>(), E,
                       (const uint32_t*)g->ad_slot_dst.as<uint32_t>(), g->ad_slot.as<uint32_t>(),
                       g->ad_dst.as<uint32_t>(), rowflag);
    hipLaunchKernelGGL(k_popc, dim3(grid_for(nwords)), dim3(kNT), 0, s, (const uint64_t*)rowflag, nwords, wordrank);
    EG_HIP(hipGetLastError());
    uint32_t err = 0;
    EG_HIP(hipMemcpyAsync(&err, w.err.ptr, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    if (err) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  }
  hipLaunchKernelGGL((k_fill_v<uint32_t>), dim3(1), dim3(64), 0, s, wordrank + nwords, uint64_t(1), 0u);
  EG_TRY(exclusive_scan_u32(wordrank, wordrank, nwords + 1, g->scan_tmp.as<uint32_t>(), s));
  uint32_t R_h = 0;
  EG_HIP(hipMemcpyAsync(&R_h, wordrank + nwords, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  g->ad_R = R_h;
  EG_TRY(g->ad_row_src.resize((uint64_t(R_h) + 1) * 4));
  if (E) {
    hipLaunchKernelGGL(k_ad_rows, dim3(grid_for(nwords, kNT / kWave)), dim3(kNT), 0, s, (const Rec8*)b.as<Rec8>(), E,
                       (const uint64_t*)rowflag, (const uint32_t*)wordrank, g->ad_row_src.as<uint32_t>());
    EG_HIP(hipGetLastError());
  }
  EG_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
  return EG_OK;
}

int ensure_adjoint(eg_graph* g, hipStream_t s) {
  EG_TRY(engine_ensure(g, s));
  if (g->ad_ready) return EG_OK;
  EG_TRY(g->rec == 8 ? adjoint_impl<Rec8>(g, s) : adjoint_impl<Rec12>(g, s));
  g->ad_ready = true;
  return EG_OK;
}

// --------------------------------------------------------------- kernels --
// y[row target] (+)= sum over the row's slots of w * x[gather index]; rows
// of one target split over sub-partitions or wave ranges meet in atomics.
__global__ __launch_bounds__(kNT) void k_wpull(uint64_t E, const uint32_t* __restrict__ idx,
                                               const uint64_t* __restrict__ rowflag,
                                               const uint32_t* __restrict__ wordrank,
                                               const uint32_t* __restrict__ row_target,
                                               const double* __restrict__ w, const double* __restrict__ x,
                                               double* __restrict__ y, int all_atomic) {
  row_reduce_wave<double>(
      first_word(), E, idx, rowflag, wordrank, 0.0, OpSum{},
      [&](uint64_t, uint32_t s, double we) { return __dmul_rn(we, x[s]); },
      [&](uint32_t row, double v, bool shared) {
        double* p = &y[row_target[row]];
        if (shared || all_atomic)
          atomicAdd(p, v);
        else
          *p = v;
      },
      w);
}

// out = a * y + b * base
__global__ void k_axpby(uint32_t V, double a, const double* __restrict__ y, double b, const double* __restrict__ base,
                        double* __restrict__ out) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    out[v] = __dadd_rn(__dmul_rn(a, y[v]), __dmul_rn(b, base[v]));
}

__global__ void k_gather_w(const uint32_t* __restrict__ slot, uint64_t E, const double* __restrict__ w,
                           double* __restrict__ out) {
  for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < E; k += uint64_t(gridDim.x) * blockDim.x)
    out[k] = w[slot[k]];
}

// dL/dw_e = d * lambda[dst_e] * x[src_e]
__global__ void k_grad_w(uint64_t E, const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, double d,
                         const double* __restrict__ lam, const double* __restrict__ x, double* __restrict__ gw) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x)
    gw[e] = __dmul_rn(__dmul_rn(d, lam[dst[e]]), x[src[e]]);
}

int check_ppr(const eg_graph* g, double d, int iterations) {
  if (!(d >= 0.0 && d < 1.0)) return fail(EG_ERR_ARG, "damping must be in [0, 1)");
  if (iterations < 0) return fail(EG_ERR_ARG, "iterations must be >= 0");
  return EG_OK;
}

// K iterations of out = a * (M v) + b * base from v = base; M = A (forward,
// CSC rows) or A^T (adjoint, transposed rows).  Work vectors in g->e_ppr.
int fixed_point(eg_graph* g, bool transposed, const double* w_rows, const double* base, double a, double b,
                int iterations, double* out, hipStream_t s) {
  const uint32_t V = g->V;
  const uint64_t E = g->E;
  EG_TRY(g->e_ppr.resize(uint64_t(V) * 8 * 2 + 256));
  double* y = g->e_ppr.as<double>();
  double* t = y + V;
  const unsigned vg = grid_for(V);
  double* cur = out;  // ping-pong between out and t
  double* nxt = t;
  EG_HIP(hipMemcpyAsync(out, base, uint64_t(V) * 8, hipMemcpyDeviceToDevice, s));
  for (int it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL((k_fill_v<double>), dim3(vg), dim3(kNT), 0, s, y, uint64_t(V), 0.0);
    if (E) {
      if (!transposed)
        hipLaunchKernelGGL(k_wpull, dim3(row_grid(E)), dim3(kNT), 0, s, E, g->e_csc_src.as<uint32_t>(),
                           g->e_rowflag.as<uint64_t>(), g->e_wordrank.as<uint32_t>(), g->e_row_dst.as<uint32_t>(),
                           w_rows, (const double*)cur, y, g->Q > 1 ? 1 : 0);
      else
        hipLaunchKernelGGL(k_wpull, dim3(row_grid(E)), dim3(kNT), 0, s, E, g->ad_dst.as<uint32_t>(),
                           g->ad_rowflag.as<uint64_t>(), g->ad_wordrank.as<uint32_t>(),
                           g->ad_row_src.as<uint32_t>(), w_rows, (const double*)cur, y, 0);
    }
    hipLaunchKernelGGL(k_axpby, dim3(vg), dim3(kNT), 0, s, V, a, (const double*)y, b, base, nxt);
    EG_HIP(hipGetLastError());
    std::swap(cur, nxt);
  }
  if (cur != out) EG_HIP(hipMemcpyAsync(out, cur, uint64_t(V) * 8, hipMemcpyDeviceToDevice, s));
  return EG_OK;
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_ppr_slots(eg_graph* g, uint32_t* d_src, uint32_t* d_dst, float* d_w, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  if (!g->built) return fail(EG_ERR_STATE, "graph has not been built");
  if (g->E && g->rec == 8)
    hipLaunchKernelGGL((k_slots<Rec8>), dim3(grid_for(g->E)), dim3(kNT), 0, s, g->csc_list.as<Rec8>(), g->E, d_src,
                       d_dst, d_w);
  else if (g->E)
    hipLaunchKernelGGL((k_slots<Rec12>), dim3(grid_for(g->E)), dim3(kNT), 0, s, g->csc_list.as<Rec12>(), g->E, d_src,
                       d_dst, d_w);
  EG_HIP(hipGetLastError());
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

int eg_ppr_forward(eg_graph* g, const double* d_w, const double* d_p, double damping, int iterations, double* d_x,
                   void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  EG_TRY(check_ppr(g, damping, iterations));
  if ((!d_w && g->E) || !d_p || !d_x) return fail(EG_ERR_ARG, "d_w, d_p and d_x are required");
  EG_TRY(engine_ensure(g, s));
  EG_TRY(fixed_point(g, false, d_w, d_p, damping, 1.0 - damping, iterations, d_x, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

int eg_ppr_adjoint(eg_graph* g, const double* d_w, const double* d_x, const double* d_c, double damping, int iterations,
                   double* d_grad_p, double* d_grad_w, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!g) return fail(EG_ERR_ARG, "graph is NULL");
  EG_TRY(check_ppr(g, damping, iterations));
  if ((!d_w && g->E) || !d_x || !d_c) return fail(EG_ERR_ARG, "d_w, d_x and d_c are required");
  EG_TRY(ensure_adjoint(g, s));
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  EG_TRY(g->e_ppr_adj.resize(std::max<uint64_t>(E, 1) * 8 + uint64_t(V) * 8 + 256));
  double* w_rows = g->e_ppr_adj.as<double>();            // weights in transposed-row order
  double* lam = w_rows + std::max<uint64_t>(E, 1);        // adjoint vector
  if (E) {
    hipLaunchKernelGGL(k_gather_w, dim3(grid_for(E)), dim3(kNT), 0, s, (const uint32_t*)g->ad_slot.as<uint32_t>(), E,
                       d_w, w_rows);
    EG_HIP(hipGetLastError());
  }
  // lambda = c + d A^T lambda
  EG_TRY(fixed_point(g, true, w_rows, d_c, damping, 1.0, iterations, lam, s));
  if (d_grad_p)
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(V)), dim3(kNT), 0, s, V, 1.0 - damping, (const double*)lam, 0.0,
                       (const double*)lam, d_grad_p);
  if (d_grad_w && E)
    hipLaunchKernelGGL(k_grad_w, dim3(grid_for(E)), dim3(kNT), 0, s, E, (const uint32_t*)g->e_csc_src.as<uint32_t>(),
                       (const uint32_t*)g->ad_slot_dst.as<uint32_t>(), damping, (const double*)lam, d_x, d_grad_w);
  EG_HIP(hipGetLastError());
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

}  // extern "C"
