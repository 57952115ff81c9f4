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
#include <mutex>
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

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
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

__global__ void k_batch_zero(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

// Host-side sizes of a batch: per-graph vertex and record bases.
struct BatchDims {
  int n = 0;
  uint64_t E = 0, Vtot = 0;
  uint32_t max_v = 0;
  std::vector<uint64_t> vbase, ebase;  // n+1 each
};

int batch_dims(int n_graphs, const uint32_t* h_vertices, const uint64_t* h_edges, const void* d_edges,
               BatchDims* d) {
  if (n_graphs < 1 || !h_vertices || !h_edges) return fail(EG_ERR_ARG, "need n_graphs >= 1 and the size arrays");
  d->n = n_graphs;
  d->vbase.assign(n_graphs + 1, 0);
  d->ebase.assign(n_graphs + 1, 0);
  for (int g = 0; g < n_graphs; ++g) {
    d->vbase[g + 1] = d->vbase[g] + h_vertices[g];
    d->ebase[g + 1] = d->ebase[g] + h_edges[g];
    d->max_v = std::max(d->max_v, h_vertices[g]);
  }
  d->Vtot = d->vbase[n_graphs];
  d->E = d->ebase[n_graphs];
  if (d->Vtot >= (uint64_t(1) << 32) || d->E >= (uint64_t(1) << 32) || d->Vtot == 0)
    return fail(EG_ERR_ARG, "batch totals must satisfy 0 < vertices < 2^32 and edges < 2^32");
  if (!d_edges && d->E) return fail(EG_ERR_ARG, "edge list is NULL");
  return EG_OK;
}

// Device buffers of one batched build.  err[0]: an id outside its graph;
// sortw.err: the sort's look-back gave up.
template <class R>
struct BatchIo {
  const R* in;
  R* out;            // E records
  uint64_t* rowptr;  // Vtot + n
  R* tmp;            // E records (sort ping-pong)
  uint64_t* scan_tmp;
  uint32_t* err;
  const uint64_t *d_vbase, *d_ebase;
};

// The build itself, shared by the handle-owning and the allocation-free
// entry points: a segmented stable sort of every graph by local source, the
// run boundaries, a max-scan and the shift to local positions.  Counters are
// cleared with a kernel (no memset node in a stream a caller captures).
template <class R>
int batch_build_into(const BatchDims& d, const BatchIo<R>& io, const SortVariant& v, RxWork& sortw, hipStream_t s) {
  const uint64_t E = d.E, n = uint64_t(d.n);
  const uint64_t nptr = d.Vtot + n;
  EG_TRY(sortw.resize_segs(d.n));
  hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<uint64_t*>(io.err), uint64_t(1));
  hipLaunchKernelGGL((k_fill_u32), dim3(1), dim3(64), 0, s, sortw.err.as<uint32_t>(), uint64_t(1), 0u);
  EG_HIP(hipGetLastError());
  const int bits = std::max(1, counting_bits(uint64_t(d.max_v ? d.max_v - 1 : 0)));
  EG_TRY(rx_sort_with(v, RxIn<R>{io.in, nullptr, nullptr, nullptr, 0}, io.out, io.tmp, d.ebase, KeyLocalSrc{}, bits,
                      sortw, s));
  hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(nptr, kNT, 8192)), dim3(kNT), 0, s, io.rowptr, nptr);
  EG_HIP(hipGetLastError());
  if (E) {
    hipLaunchKernelGGL((k_batch_runs<R>), dim3(grid_for(E, kNT, 8192)), dim3(kNT), 0, s, (const R*)io.out, E,
                       io.d_ebase, io.d_vbase, d.n, io.rowptr, io.err);
    EG_HIP(hipGetLastError());
  }
  EG_TRY(inclusive_max_scan_u64(io.rowptr, io.rowptr, nptr, io.scan_tmp, s));
  hipLaunchKernelGGL(k_batch_localize, dim3(grid_for(nptr, kNT, 8192)), dim3(kNT), 0, s, io.rowptr, nptr, io.d_vbase,
                     io.d_ebase, d.n);
  EG_HIP(hipGetLastError());
  uint32_t h_err[2] = {0, 0};
  EG_HIP(hipMemcpyAsync(&h_err[0], io.err, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipMemcpyAsync(&h_err[1], sortw.err.ptr, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (h_err[0]) return fail(EG_ERR_RANGE, "a record's vertex id is outside its graph");
  if (h_err[1]) return fail(EG_ERR_HIP, "radix sort look-back exceeded its spin limit");
  return EG_OK;
}

template <class R>
int batch_build(eg_batch* b, const R* d_in, hipStream_t s) {
  const uint64_t E = b->E, n = uint64_t(b->n);
  const uint64_t nptr = b->Vtot + n;
  EG_TRY(b->t2.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->out.resize(std::max<uint64_t>(E, 1) * sizeof(R)));
  EG_TRY(b->rowptr.resize(nptr * 8));
  EG_TRY(b->scan_tmp.resize(scan_tmp_elems(nptr) * 8));
  EG_TRY(b->err.resize(8));
  BatchDims d;
  d.n = b->n;
  d.E = b->E;
  d.Vtot = b->Vtot;
  d.max_v = b->max_v;
  d.vbase = b->vbase;
  d.ebase = b->ebase;
  const BatchIo<R> io{d_in, b->out.as<R>(), b->rowptr.as<uint64_t>(), b->t2.as<R>(), b->scan_tmp.as<uint64_t>(),
                      b->err.as<uint32_t>(), b->d_vbase.as<uint64_t>(), b->d_ebase.as<uint64_t>()};
  return batch_build_into(d, io, sort_variant(int(sizeof(R))), b->sortw, s);
}

// --------------------------------------------- allocation-free variant --
// For callers that cannot allocate device memory: every temporary of the
// build lives in this array, one per GPU for the life of the process, whose
// address each device resolves once (hipGetSymbolAddress) on its first call.
// The host never uses the array's name as a device address.
constexpr uint64_t kScratchBytes = 32ull << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(void** out) {
  int dev = 0;
  EG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(EG_ERR_ARG, "device %d beyond the %d this library tracks", dev,
                                                 kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(EG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return EG_OK;
}

// Carves 256-byte aligned pieces off the scratch; fails when it runs out.
struct Carve {
  char* base;
  uint64_t used = 0, cap;
  void* take(uint64_t bytes) {
    const uint64_t at = (used + 255) & ~uint64_t(255);
    if (at + bytes > cap) {
      used = cap + 1;
      return nullptr;
    }
    used = at + bytes;
    return base + at;
  }
  uint64_t need(uint64_t bytes) const { return ((used + 255) & ~uint64_t(255)) + bytes; }
};

// The variant every allocation-free build uses: three launches, tiles of
// 256 x 16 records in XCD order, 8-bit digits (no look-back status words,
// the smallest digit table).
const SortVariant kNoAllocSort{kThreeLaunchXcd, 256, 8};

template <class R>
int batch_build_noalloc(const BatchDims& d, const R* in, R* out, uint64_t* rowptr, hipStream_t s) {
  const uint64_t E = d.E, n = uint64_t(d.n), nptr = d.Vtot + n;
  const uint64_t tile = uint64_t(kNoAllocSort.nt) * kRxIPT;
  // the pieces, in carving order
  const uint64_t sizes[] = {
      std::max<uint64_t>(E, 1) * sizeof(R),                      // sort ping-pong
      std::max<uint64_t>(rx_three_launch_table_bytes(d.ebase, tile), 4),  // digit table
      rx_three_launch_scan_bytes(d.ebase, tile),                 // its scan temporaries
      (n + 1) * 12 + 16,                                         // sort segment tables
      uint64_t(kRxMaxPasses) * 4,                                // sort ticket words
      8,                                                         // sort error word
      scan_tmp_elems(nptr) * 8,                                  // row-pointer scan temporaries
      8,                                                         // id error word
      (n + 1) * 8, (n + 1) * 8};                                 // vertex and record bases
  uint64_t total = 0;
  for (uint64_t b : sizes) total = ((total + 255) & ~uint64_t(255)) + b;
  if (total > kScratchBytes)
    return fail(EG_ERR_ARG,
                "eg_batch_csr_build_noalloc: a batch of %llu records, %llu vertices and %d graphs needs %llu bytes of "
                "scratch; the limit is %llu (use eg_batch_csr_build)",
                (unsigned long long)E, (unsigned long long)d.Vtot, d.n, (unsigned long long)total,
                (unsigned long long)kScratchBytes);
  void* base = nullptr;
  EG_TRY(scratch_of_current_device(&base));
  Carve c{static_cast<char*>(base), 0, kScratchBytes};
  R* tmp = static_cast<R*>(c.take(sizes[0]));
  RxWork w;
  w.table.borrow(c.take(sizes[1]), sizes[1]);
  w.scan.borrow(c.take(sizes[2]), sizes[2]);
  w.segs.borrow(c.take(sizes[3]), sizes[3]);
  w.ticket.borrow(c.take(sizes[4]), sizes[4]);
  w.err.borrow(c.take(sizes[5]), sizes[5]);
  uint64_t* scan_tmp = static_cast<uint64_t*>(c.take(sizes[6]));
  uint32_t* err = static_cast<uint32_t*>(c.take(sizes[7]));
  uint64_t* d_vbase = static_cast<uint64_t*>(c.take(sizes[8]));
  uint64_t* d_ebase = static_cast<uint64_t*>(c.take(sizes[9]));
  if (c.used > kScratchBytes) return fail(EG_ERR_ARG, "internal: scratch carving overran its sizing");
  EG_HIP(hipMemcpyAsync(d_vbase, d.vbase.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(d_ebase, d.ebase.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  const BatchIo<R> io{in, out, rowptr, tmp, scan_tmp, err, d_vbase, d_ebase};
  return batch_build_into(d, io, kNoAllocSort, w, s);
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_batch_csr_build(const void* d_edges, int rec_bytes, int n_graphs, const uint32_t* h_vertices,
                       const uint64_t* h_edges, eg_batch** out, void* stream) {
  if (!out) return fail(EG_ERR_ARG, "eg_batch_csr_build: out is NULL");
  *out = nullptr;
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  BatchDims d;
  EG_TRY(batch_dims(n_graphs, h_vertices, h_edges, d_edges, &d));
  eg_batch* b = new (std::nothrow) eg_batch();
  if (!b) return fail(EG_ERR_NOMEM, "host allocation failed");
  b->n = n_graphs;
  b->rec = rec_bytes;
  b->vbase = d.vbase;
  b->ebase = d.ebase;
  b->max_v = d.max_v;
  b->Vtot = d.Vtot;
  b->E = d.E;
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

int eg_batch_csr_build_noalloc(const void* d_edges, int rec_bytes, int n_graphs, const uint32_t* h_vertices,
                               const uint64_t* h_edges, void* d_out_list, uint64_t* d_out_rowptr, void* stream) {
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  BatchDims d;
  EG_TRY(batch_dims(n_graphs, h_vertices, h_edges, d_edges, &d));
  if ((!d_out_list && d.E) || !d_out_rowptr) return fail(EG_ERR_ARG, "eg_batch_csr_build_noalloc: NULL output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return rec_bytes == 8 ? batch_build_noalloc<Rec8>(d, static_cast<const Rec8*>(d_edges),
                                                    static_cast<Rec8*>(d_out_list), d_out_rowptr, s)
                        : batch_build_noalloc<Rec12>(d, static_cast<const Rec12*>(d_edges),
                                                     static_cast<Rec12*>(d_out_list), d_out_rowptr, s);
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
