// Batched CSR construction of many independent graphs in one pass over the
// GPU (no reference counterpart: the reference builds one graph per process;
// each graph here gets exactly the CSR that prepare_push_mode builds for it at
// P = 1, S = 1, core/graph.hpp:1248-1355: records ordered by (src, input
// position), dense row pointer).
//
// The graphs stay in place as the segments of one segmented stable LSD sort
// by local source id (a tile never spans two graphs; the digit count is set by
// the largest graph).  Row pointers come from run boundaries (first/last
// record of every source) written as global positions, one inclusive
// max-scan over all pointer arrays (an empty leading row of graph g reads the
// previous graph's end, ebase[g]) and a final shift to local positions.
#include <algorithm>
#include <vector>

#include "radix.hpp"

struct eg_batch {
  int n = 0;
  int rec = 8;
  uint64_t E = 0, Vtot = 0;
  uint32_t max_v = 0;
  std::vector<uint64_t> vbase, ebase;  // n+1 each (host)
  eg::DevBuf d_vbase, d_ebase;         // n+1 each (device)
  eg::DevBuf t2, out;                  // records
  eg::DevBuf rowptr;                   // u64[Vtot + n]: graph g's pointers at vbase[g] + g
  eg::DevBuf scan_tmp, err;
  eg::RxWork sortw;
};

namespace eg {
namespace {

constexpr int kNT = 256;
constexpr int kLdsGraphs = 1024;  // bases (graphs + 1) that a block copies to LDS

unsigned grid_for(uint64_t n, uint64_t per = kNT, uint64_t cap = 65536) {
  uint64_t b = ceil_div(n, per);
  return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

// largest g in [0, n) with base[g] <= x (base non-decreasing, base[0] = 0)
__device__ __forceinline__ int graph_of(const uint64_t* base, int n, uint64_t x) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (base[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// The n+1 bases in LDS when they fit (a block reads them once), else global.
__device__ __forceinline__ const uint64_t* stage_bases(const uint64_t* base, int n, uint64_t* lds) {
  if (n + 1 > kLdsGraphs) return base;
  for (int i = threadIdx.x; i <= n; i += blockDim.x) lds[i] = base[i];
  __syncthreads();
  return lds;
}

struct KeyLocalSrc {
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.src; }
};

// Run boundaries of the sorted graphs (graph g = records [ebase[g],
// ebase[g+1])); ids outside their graph flag err and write nothing.  Graph
// g's pointer array starts at slot vbase[g] + g; entries are global record
// positions until the localise step.
template <class R>
__global__ __launch_bounds__(kNT) void k_batch_runs(const R* __restrict__ a, uint64_t E,
                                                    const uint64_t* __restrict__ ebase_g,
                                                    const uint64_t* __restrict__ vbase_g, int n,
                                                    uint64_t* __restrict__ rowptr, uint32_t* __restrict__ err) {
  __shared__ uint64_t s_e[kLdsGraphs], s_v[kLdsGraphs];
  const uint64_t* ebase = stage_bases(ebase_g, n, s_e);
  const uint64_t* vbase = stage_bases(vbase_g, n, s_v);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += uint64_t(gridDim.x) * blockDim.x) {
    const int g = graph_of(ebase, n, i);
    const uint64_t nv = vbase[g + 1] - vbase[g];
    const R r = a[i];
    if (r.src >= nv || r.dst >= nv) {
      atomicOr(err, 1u);
      continue;
    }
    const uint64_t slot = vbase[g] + g + r.src;
    if (i == ebase[g] || a[i - 1].src != r.src) rowptr[slot] = i;
    if (i + 1 == ebase[g + 1] || a[i + 1].src != r.src) rowptr[slot + 1] = i + 1;
  }
}

// Global positions -> local: pointer slot i of graph g (vbase[g] + g <= i
// <= vbase[g+1] + g) loses ebase[g].
__global__ __launch_bounds__(kNT) void k_batch_localize(uint64_t* __restrict__ rowptr, uint64_t nptr,
                                                        const uint64_t* __restrict__ vbase_g,
                                                        const uint64_t* __restrict__ ebase_g, int n) {
  __shared__ uint64_t s_e[kLdsGraphs], s_v[kLdsGraphs];
  const uint64_t* ebase = stage_bases(ebase_g, n, s_e);
  const uint64_t* vbase = stage_bases(vbase_g, n, s_v);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nptr; i += uint64_t(gridDim.x) * blockDim.x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (vbase[mid] + mid <= i)
        lo = mid;
      else
        hi = mid;
    }
    rowptr[i] -= ebase[lo];
  }
}

template <class R>
int batch_build(eg_batch* b, const R* d_in, hipStream_t s) {
  const uint64_t E = b->E, n = uint64_t(b->n);
  const uint64_t nptr = b->Vtot + n;
  EG_TRY(b->t2.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->out.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->rowptr.resize(nptr * 8));
  EG_TRY(b->scan_tmp.resize(scan_tmp_elems(nptr) * 8));
  EG_TRY(b->err.resize(4));
  uint32_t* err = b->err.as<uint32_t>();
  EG_HIP(hipMemsetAsync(err, 0, 4, s));
  const uint64_t* vb = b->d_vbase.as<uint64_t>();
  const uint64_t* eb = b->d_ebase.as<uint64_t>();
  const int bits = std::max(1, counting_bits(uint64_t(b->max_v ? b->max_v - 1 : 0)));
  EG_TRY(b->sortw.resize_segs(b->n));
  EG_HIP(hipMemsetAsync(b->sortw.err.ptr, 0, 4, s));
  EG_TRY(rx_sort(d_in, b->out.as<R>(), b->t2.as<R>(), b->ebase, KeyLocalSrc{}, bits, b->sortw, s));
  uint64_t* rp = b->rowptr.as<uint64_t>();
  EG_HIP(hipMemsetAsync(rp, 0, nptr * 8, s));
  if (E) {
    hipLaunchKernelGGL((k_batch_runs<R>), dim3(grid_for(E, kNT, 8192)), dim3(kNT), 0, s, (const R*)b->out.as<R>(), E,
                       eb, vb, b->n, rp, err);
    EG_HIP(hipGetLastError());
  }
  EG_TRY(inclusive_max_scan_u64(rp, rp, nptr, b->scan_tmp.as<uint64_t>(), s));
  hipLaunchKernelGGL(k_batch_localize, dim3(grid_for(nptr, kNT, 8192)), dim3(kNT), 0, s, rp, nptr, vb, eb, b->n);
  EG_HIP(hipGetLastError());
  uint32_t h_err[2] = {0, 0};
  EG_HIP(hipMemcpyAsync(&h_err[0], err, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipMemcpyAsync(&h_err[1], b->sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (h_err[0]) return fail(EG_ERR_RANGE, "a record's vertex id is outside its graph");
  if (h_err[1]) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  return EG_OK;
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_batch_csr_build(const void* d_edges, int rec_bytes, int n_graphs, const uint32_t* h_vertices,
                       const uint64_t* h_edges, eg_batch** out, void* stream) {
  if (!out) return fail(EG_ERR_ARG, "eg_batch_csr_build: out is NULL");
  *out = nullptr;
  if (n_graphs < 1 || !h_vertices || !h_edges) return fail(EG_ERR_ARG, "need n_graphs >= 1 and the size arrays");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  eg_batch* b = new (std::nothrow) eg_batch();
  if (!b) return fail(EG_ERR_NOMEM, "host allocation failed");
  b->n = n_graphs;
  b->rec = rec_bytes;
  b->vbase.assign(n_graphs + 1, 0);
  b->ebase.assign(n_graphs + 1, 0);
  for (int g = 0; g < n_graphs; ++g) {
    b->vbase[g + 1] = b->vbase[g] + h_vertices[g];
    b->ebase[g + 1] = b->ebase[g] + h_edges[g];
    b->max_v = std::max(b->max_v, h_vertices[g]);
  }
  b->Vtot = b->vbase[n_graphs];
  b->E = b->ebase[n_graphs];
  if (b->Vtot >= (uint64_t(1) << 32) || b->E >= (uint64_t(1) << 32) || b->Vtot == 0) {
    delete b;
    return fail(EG_ERR_ARG, "batch totals must satisfy 0 < vertices < 2^32 and edges < 2^32");
  }
  if (!d_edges && b->E) {
    delete b;
    return fail(EG_ERR_ARG, "edge list is NULL");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = b->d_vbase.resize((n_graphs + 1) * 8);
  if (rc == EG_OK) rc = b->d_ebase.resize((n_graphs + 1) * 8);
  if (rc == EG_OK && hipMemcpyAsync(b->d_vbase.ptr, b->vbase.data(), (n_graphs + 1) * 8, hipMemcpyHostToDevice, s) !=
                         hipSuccess)
    rc = fail(EG_ERR_HIP, "copy of vertex bases failed");
  if (rc == EG_OK && hipMemcpyAsync(b->d_ebase.ptr, b->ebase.data(), (n_graphs + 1) * 8, hipMemcpyHostToDevice, s) !=
                         hipSuccess)
    rc = fail(EG_ERR_HIP, "copy of edge bases failed");
  if (rc == EG_OK)
    rc = rec_bytes == 8 ? batch_build<Rec8>(b, static_cast<const Rec8*>(d_edges), s)
                        : batch_build<Rec12>(b, static_cast<const Rec12*>(d_edges), s);
  if (rc != EG_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return EG_OK;
}

int eg_batch_csr_rebuild(eg_batch* b, const void* d_edges, void* stream) {
  if (!b) return fail(EG_ERR_ARG, "batch is NULL");
  if (!d_edges && b->E) return fail(EG_ERR_ARG, "edge list is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->rec == 8 ? batch_build<Rec8>(b, static_cast<const Rec8*>(d_edges), s)
                     : batch_build<Rec12>(b, static_cast<const Rec12*>(d_edges), s);
}

int eg_batch_csr_get(const eg_batch* b, int g, const void** d_list, uint64_t* n_edges, const uint64_t** d_rowptr,
                     uint32_t* n_vertices) {
  if (!b || g < 0 || g >= b->n) return fail(EG_ERR_ARG, "bad batch or graph index");
  if (d_list) *d_list = static_cast<const char*>(b->out.ptr) + b->ebase[g] * b->rec;
  if (n_edges) *n_edges = b->ebase[g + 1] - b->ebase[g];
  if (d_rowptr) *d_rowptr = b->rowptr.as<uint64_t>() + b->vbase[g] + g;
  if (n_vertices) *n_vertices = uint32_t(b->vbase[g + 1] - b->vbase[g]);
  return EG_OK;
}

int eg_batch_destroy(eg_batch* b) {
  delete b;
  return EG_OK;
}

}  // extern "C"
