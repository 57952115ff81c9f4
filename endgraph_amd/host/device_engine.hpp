// Generic process_vertices / process_edges over device functors: the
// reference's signal/slot interface (core/graph.hpp:1377-1441, :1473-1817)
// for toolkits compiled with hipcc for gfx950.  Include it after graph.hpp
// in a .hip source; the lambdas of a reference toolkit become __device__
// functors with these shapes (M = message type, R = reducer type, E =
// EdgeData):
//
//   process_vertices:  R   process(UINT v)
//   sparse_signal:     bool signal(UINT v, M& msg)              emit(v, msg) <=> return true
//   sparse_slot:       R   slot(UINT src, M msg, AdjacentListType<E> out_edges)
//   dense_signal:      bool signal(UINT dst, AdjacentListType<E> in_edges, M& msg)
//   dense_slot:        R   slot(UINT dst, M msg)
//
// As in the reference, a vertex's adjacency reaches its slot or signal one
// sub-partition (partition, socket) at a time, as a run of the built
// EdgeType records, so a destination can receive several dense messages in
// one step; a step's signals all run before its slots (the reference buffers
// the messages between them); slot results are summed into the reducer.
// The mode follows the graph's policy and use_* flags with the rule of
// core/graph.hpp:1483-1498 (R3/R4).  Every call runs on the default stream
// and returns when the step is done.  One thread serves a (vertex,
// sub-partition) run or a compressed-index row: a general path for new
// programs; the library's own programs (programs::*) are the fast one.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "graph.hpp"

namespace endgraph {

// core/type.hpp:66-74
template <typename EdgeData>
struct AdjacentListType {
  const EdgeType<EdgeData>* begin;
  const EdgeType<EdgeData>* end;
};

namespace device_detail {

constexpr int kThreads = 256;

inline unsigned grid_of(uint64_t n) {
  uint64_t b = (n + kThreads - 1) / kThreads;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return unsigned(b);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(EG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ bool get_bit(const uint64_t* bm, UINT v) { return (bm[v >> 6] >> (v & 63)) & 1ull; }

// Block sum of one reducer value per thread, added to *out by thread 0.
template <typename R>
__device__ __forceinline__ void reduce_to(R v, R* out) {
  __shared__ R part[kThreads / 64];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    R t = R(0);
    for (int i = 0; i < kThreads / 64; ++i) t += part[i];
    if (t != R(0)) atomicAdd(out, t);
  }
}

template <typename R, typename F>
__global__ __launch_bounds__(kThreads) void k_vertices(UINT V, const uint64_t* active, F f, R* out) {
  R acc = R(0);
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    if (get_bit(active, UINT(v))) acc += f(UINT(v));
  reduce_to(acc, out);
}

// sparse signal: messages of the active vertices that emit, and their bits
template <typename M, typename F>
__global__ __launch_bounds__(kThreads) void k_sparse_signal(UINT V, const uint64_t* active, F f, M* msg,
                                                            uint64_t* emitted) {
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x) {
    if (!get_bit(active, UINT(v))) continue;
    M m;
    if (f(UINT(v), m)) {
      msg[v] = m;
      atomicOr(reinterpret_cast<unsigned long long*>(&emitted[v >> 6]), 1ull << (v & 63));
    }
  }
}

// one sub-partition's CSR: its records, dense index and bitmap
template <typename E>
struct CsrPart {
  const EdgeType<E>* list;
  const uint64_t* index;
  const uint64_t* bitmap;
};

template <typename R, typename M, typename E, typename F>
__global__ __launch_bounds__(kThreads) void k_sparse_slot(UINT V, CsrPart<E> part, const uint64_t* emitted,
                                                          const M* msg, F f, R* out) {
  R acc = R(0);
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V; v += uint64_t(gridDim.x) * blockDim.x)
    if (get_bit(emitted, UINT(v)) && get_bit(part.bitmap, UINT(v)))
      acc += f(UINT(v), msg[v], AdjacentListType<E>{part.list + part.index[v], part.list + part.index[v + 1]});
  reduce_to(acc, out);
}

// one sub-partition's CSC: its records and compressed index (rows + sentinel)
template <typename E>
struct CscPart {
  const EdgeType<E>* list;
  const uint8_t* cindex;  // packed 12-byte {u64 index, u32 vertex} entries
  uint64_t rows;
};

__device__ __forceinline__ uint64_t cidx_index(const uint8_t* c, uint64_t k) {
  uint64_t x;
  __builtin_memcpy(&x, c + k * 12, 8);
  return x;
}
__device__ __forceinline__ UINT cidx_vertex(const uint8_t* c, uint64_t k) {
  UINT x;
  __builtin_memcpy(&x, c + k * 12 + 8, 4);
  return x;
}

template <typename M, typename E, typename F>
__global__ __launch_bounds__(kThreads) void k_dense_signal(CscPart<E> part, F f, M* msg, uint8_t* has) {
  for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < part.rows;
       k += uint64_t(gridDim.x) * blockDim.x) {
    const UINT dst = cidx_vertex(part.cindex, k);
    const AdjacentListType<E> adj{part.list + cidx_index(part.cindex, k), part.list + cidx_index(part.cindex, k + 1)};
    M m;
    const bool e = f(dst, adj, m);
    if (e) msg[k] = m;
    has[k] = e ? 1 : 0;
  }
}

template <typename R, typename M, typename E, typename F>
__global__ __launch_bounds__(kThreads) void k_dense_slot(CscPart<E> part, const M* msg, const uint8_t* has, F f,
                                                         R* out) {
  R acc = R(0);
  for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < part.rows;
       k += uint64_t(gridDim.x) * blockDim.x)
    if (has[k]) acc += f(cidx_vertex(part.cindex, k), msg[k]);
  reduce_to(acc, out);
}

// device scratch of one size, grown on demand and kept by the engine
struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
  void* need(size_t n) {
    if (n > bytes) {
      if (p) eg_device_free(p);
      p = nullptr;
      check(eg_device_alloc(&p, n));
      bytes = n;
    }
    return p;
  }
  ~Scratch() {
    if (p) eg_device_free(p);
  }
};

}  // namespace device_detail

// The generic engine of one built Graph<EdgeData>; its buffers are kept
// between calls.
template <typename EdgeData>
class DeviceEngine {
 public:
  explicit DeviceEngine(Graph<EdgeData>& g) : g_(g) {}

  // process_vertices (core/graph.hpp:1377-1441)
  template <typename R, typename F>
  R process_vertices(F process, VertexSubset* active) {
    R* out = reducer<R>();
    hipLaunchKernelGGL((device_detail::k_vertices<R, F>), dim3(device_detail::grid_of(g_.vertices_num)),
                       dim3(device_detail::kThreads), 0, nullptr, g_.vertices_num, (const uint64_t*)active->data,
                       process, out);
    return finish<R>(out);
  }

  // process_edges (core/graph.hpp:1473-1817)
  template <typename R, typename M, typename SS, typename SL, typename DS, typename DL>
  R process_edges(SS sparse_signal, SL sparse_slot, DS dense_signal, DL dense_slot, VertexSubset* active) {
    const UINT V = g_.vertices_num;
    R* out = reducer<R>();
    if (sparse(active)) {
      const uint64_t words = (uint64_t(V) >> 6) + 1;
      M* msg = static_cast<M*>(msg_.need(sizeof(M) * (V ? V : 1)));
      uint64_t* emitted = static_cast<uint64_t*>(emitted_.need(words * 8));
      check(eg_bitmap_clear(emitted, V, nullptr));
      hipLaunchKernelGGL((device_detail::k_sparse_signal<M, SS>), dim3(device_detail::grid_of(V)),
                         dim3(device_detail::kThreads), 0, nullptr, V, (const uint64_t*)active->data, sparse_signal,
                         msg, emitted);
      for (int p = 0; p < g_.partitions_num; ++p)
        for (int s = 0; s < g_.sockets_num; ++s) {
          device_detail::CsrPart<EdgeData> part{};
          const void* lst = nullptr;
          uint64_t n = 0;
          check(eg_graph_csr(g_.handle(), p, s, &lst, &n, &part.index, &part.bitmap));
          if (!n) continue;
          part.list = static_cast<const EdgeType<EdgeData>*>(lst);
          hipLaunchKernelGGL((device_detail::k_sparse_slot<R, M, EdgeData, SL>), dim3(device_detail::grid_of(V)),
                             dim3(device_detail::kThreads), 0, nullptr, V, part, (const uint64_t*)emitted,
                             (const M*)msg, sparse_slot, out);
        }
    } else {
      // every row of every sub-partition signals (dense mode ignores the
      // active set, core/graph.hpp:1686-1700), then the slots apply
      std::vector<device_detail::CscPart<EdgeData>> parts;
      uint64_t rows = 0;
      for (int p = 0; p < g_.partitions_num; ++p)
        for (int s = 0; s < g_.sockets_num; ++s) {
          const void *lst = nullptr, *ci = nullptr;
          uint64_t n = 0, nci = 0;
          check(eg_graph_csc(g_.handle(), p, s, &lst, &n, &ci, &nci));
          if (nci < 2) continue;
          parts.push_back({static_cast<const EdgeType<EdgeData>*>(lst), static_cast<const uint8_t*>(ci), nci - 1});
          rows += nci - 1;
        }
      M* msg = static_cast<M*>(msg_.need(sizeof(M) * (rows ? rows : 1)));
      uint8_t* has = static_cast<uint8_t*>(emitted_.need(rows ? rows : 1));
      uint64_t at = 0;
      for (const auto& part : parts) {
        hipLaunchKernelGGL((device_detail::k_dense_signal<M, EdgeData, DS>), dim3(device_detail::grid_of(part.rows)),
                           dim3(device_detail::kThreads), 0, nullptr, part, dense_signal, msg + at, has + at);
        at += part.rows;
      }
      at = 0;
      for (const auto& part : parts) {
        hipLaunchKernelGGL((device_detail::k_dense_slot<R, M, EdgeData, DL>), dim3(device_detail::grid_of(part.rows)),
                           dim3(device_detail::kThreads), 0, nullptr, part, (const M*)(msg + at),
                           (const uint8_t*)(has + at), dense_slot, out);
        at += part.rows;
      }
    }
    return finish<R>(out);
  }

  // the mode the last process_edges call ran in
  bool last_was_sparse() const { return last_sparse_; }

 private:
  Graph<EdgeData>& g_;
  device_detail::Scratch msg_, emitted_, red_;
  bool last_sparse_ = false;

  // core/graph.hpp:1483-1498 with R3/R4, under the graph's policy
  bool sparse(VertexSubset* active) {
    bool sp;
    if (g_.policy == EG_POLICY_PUSH) {
      sp = true;
    } else if (g_.policy == EG_POLICY_PULL) {
      sp = false;
    } else if (g_.policy == EG_POLICY_REFERENCE && !g_.use_undirected) {
      sp = !g_.use_iteration;
    } else {
      uint64_t active_edges = 0;
      check(eg_bitmap_degree_sum(g_.handle(), active->data, &active_edges, nullptr));
      sp = active_edges < g_.edges_num / 20;
    }
    last_sparse_ = sp;
    return sp;
  }

  template <typename R>
  R* reducer() {
    R* out = static_cast<R*>(red_.need(sizeof(R) < 8 ? 8 : sizeof(R)));
    const R zero = R(0);
    check(eg_memcpy_h2d(out, &zero, sizeof(R), nullptr));
    return out;
  }

  template <typename R>
  R finish(R* out) {
    device_detail::hip_check(hipGetLastError(), "generic engine launch");
    R r = R(0);
    check(eg_memcpy_d2h(&r, out, sizeof(R), nullptr));  // synchronises
    return r;
  }
};

}  // namespace endgraph
