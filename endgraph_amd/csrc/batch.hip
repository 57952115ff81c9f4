// Batched CSR construction of many independent graphs in one pass over the
// GPU (no reference counterpart: the reference builds one graph per process;
// each graph here gets exactly the CSR that prepare_push_mode builds for it at
// P = 1, S = 1, core/graph.hpp:1248-1355: records ordered by (src, input
// position), dense row pointer).
//
// The batch is treated as the disjoint union of its graphs: sources are
// relabelled to global ids (graph base + local id), one stable LSD sort by the
// global source orders every graph at once, and the sources are mapped back.
// Row pointers come from run boundaries (first/last record of every source)
// and an inclusive max-scan, as in the single-graph build.
#include <algorithm>
#include <vector>

#include "radix.hpp"

struct eg_batch {
  int n = 0;
  int rec = 8;
  uint64_t E = 0, Vtot = 0;
  std::vector<uint64_t> vbase, ebase;  // n+1 each (host)
  eg::DevBuf d_vbase, d_ebase;         // n+1 each (device)
  eg::DevBuf relabel, t2, out;         // records
  eg::DevBuf rowptr;                   // u64[Vtot + n]: graph g's pointers at vbase[g] + g
  eg::DevBuf hist, scan_tmp, err;
  eg::RxWork sortw;
  uint32_t max_v = 0;
};

namespace eg {
namespace {

constexpr int kNT = 256;

unsigned grid_for(uint64_t n, uint64_t per = kNT, uint64_t cap = 65536) {
  uint64_t b = ceil_div(n, per);
  return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

__device__ __forceinline__ int graph_of(const uint64_t* base, int n, uint64_t x) {
  int lo = 0, hi = n;  // base[lo] <= x < base[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (base[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// src -> vbase[g] + src; records whose ids fall outside their graph flag err.
template <class R>
__global__ void k_relabel(const R* __restrict__ in, R* __restrict__ out, uint64_t E, const uint64_t* __restrict__ ebase,
                          const uint64_t* __restrict__ vbase, int n, uint32_t* __restrict__ err) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    const int g = graph_of(ebase, n, e);
    R r = in[e];
    const uint64_t nv = vbase[g + 1] - vbase[g];
    if (r.src >= nv || r.dst >= nv) {
      atomicOr(err, 1u);
      r.src = 0;
      r.dst = 0;
    }
    r.src = uint32_t(vbase[g] + r.src);
    out[e] = r;
  }
}

// Run boundaries of the globally sorted list: graph g's pointer array starts
// at vbase[g] + g and holds vbase[g+1]-vbase[g]+1 entries; entries are global
// record positions until the localise step.
template <class R>
__global__ void k_batch_runs(const R* __restrict__ a, uint64_t E, const uint64_t* __restrict__ vbase, int n,
                             uint64_t* __restrict__ rowptr) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t s = a[i].src;
    const int g = graph_of(vbase, n, s);
    const uint64_t slot = uint64_t(s) + g;  // = vbase[g] + g + local id
    if (i == 0 || a[i - 1].src != s) rowptr[slot] = i;
    if (i == E - 1 || a[i + 1].src != s) rowptr[slot + 1] = i + 1;
  }
}

// Back to local ids: src -= vbase[g]; pointers -= ebase[g].
template <class R>
__global__ void k_batch_localize(R* __restrict__ a, uint64_t E, uint64_t* __restrict__ rowptr, uint64_t nptr,
                                 const uint64_t* __restrict__ vbase, const uint64_t* __restrict__ ebase, int n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < E; i += stride) {
    const uint32_t s = a[i].src;
    a[i].src = uint32_t(s - vbase[graph_of(vbase, n, s)]);
  }
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nptr; i += stride) {
    // pointer slot i belongs to graph g with vbase[g] + g <= i <= vbase[g+1] + g
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

struct KeyGlobalSrc {
  template <class R>
  __device__ __forceinline__ uint64_t operator()(const R& r) const { return r.src; }
};

template <class R>
int batch_build(eg_batch* b, const R* d_in, hipStream_t s) {
  const uint64_t E = b->E, n = uint64_t(b->n);
  const uint64_t nptr = b->Vtot + n;
  EG_TRY(b->relabel.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->t2.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->out.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->rowptr.resize(nptr * 8));
  EG_TRY(b->scan_tmp.resize(scan_tmp_elems(nptr) * 8));
  DevBuf& err = b->err;
  EG_TRY(err.resize(4));
  EG_HIP(hipMemsetAsync(err.ptr, 0, 4, s));
  const uint64_t* vb = b->d_vbase.as<uint64_t>();
  const uint64_t* eb = b->d_ebase.as<uint64_t>();
  if (E) {
    hipLaunchKernelGGL((k_relabel<R>), dim3(grid_for(E)), dim3(kNT), 0, s, d_in, b->relabel.as<R>(), E, eb, vb, b->n,
                       err.as<uint32_t>());
    EG_HIP(hipGetLastError());
  }
  uint32_t h_err = 0;
  EG_HIP(hipMemcpyAsync(&h_err, err.ptr, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (h_err) return fail(EG_ERR_RANGE, "a record's vertex id is outside its graph");
  const int bits = std::max(1, counting_bits(b->Vtot - 1));
  EG_TRY(b->sortw.resize_segs(1));
  EG_HIP(hipMemsetAsync(b->sortw.err.ptr, 0, 4, s));
  EG_TRY(rx_sort(static_cast<const R*>(b->relabel.as<R>()), b->out.as<R>(), b->t2.as<R>(), std::vector<uint64_t>{0, E},
                 KeyGlobalSrc{}, bits, b->sortw, s));
  uint64_t* rp = b->rowptr.as<uint64_t>();
  EG_HIP(hipMemsetAsync(rp, 0, nptr * 8, s));
  if (E) {
    hipLaunchKernelGGL((k_batch_runs<R>), dim3(grid_for(E)), dim3(kNT), 0, s, (const R*)b->out.as<R>(), E, vb, b->n,
                       rp);
    EG_HIP(hipGetLastError());
  }
  // Empty leading rows of graph g read the previous graph's end = ebase[g]:
  // positions are global and non-decreasing across graphs.
  EG_TRY(inclusive_max_scan_u64(rp, rp, nptr, b->scan_tmp.as<uint64_t>(), s));
  hipLaunchKernelGGL((k_batch_localize<R>), dim3(grid_for(std::max(E, nptr))), dim3(kNT), 0, s, b->out.as<R>(), E, rp,
                     nptr, vb, eb, b->n);
  EG_HIP(hipGetLastError());
  uint32_t sort_err = 0;
  EG_HIP(hipMemcpyAsync(&sort_err, b->sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (sort_err) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
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
