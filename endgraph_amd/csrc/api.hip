// C-ABI glue: error reporting, owned device buffers, memory helpers.
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "radix.hpp"

namespace eg {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return status;
}

int DevBuf::resize(uint64_t want) {
  if (want <= bytes && ptr) return EG_OK;
  if (!owned)
    return fail(EG_ERR_ARG, "internal: a scratch view of %llu bytes cannot grow to %llu", (unsigned long long)bytes,
                (unsigned long long)want);
  release();
  if (want == 0) want = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error of a failed allocation
    return fail(EG_ERR_NOMEM, "hipMalloc(%llu bytes) failed: %s", (unsigned long long)want, hipGetErrorString(e));
  }
  ptr = p;
  bytes = want;
  return EG_OK;
}

void DevBuf::release() {
  if (ptr && owned) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  owned = true;
}

namespace {
struct ProfRec {
  hipEvent_t a, b;
  int kernel;
  uint64_t bytes;
};
bool g_prof = false;
hipEvent_t g_pending = nullptr;
std::vector<ProfRec> g_recs;
std::vector<hipEvent_t> g_pool;
struct ProfAgg {
  double ms = 0;
  uint64_t launches = 0, bytes = 0;
};
ProfAgg g_agg[PK_COUNT];
const char* kProfNames[PK_COUNT] = {"radix_upsweep", "radix_downsweep", "scan", "degree",
                                    "cindex", "csr_index", "pr_pull", "bfs"};

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void drain() {
  for (auto& r : g_recs) {
    float ms = 0;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      g_agg[r.kernel].ms += ms;
      g_agg[r.kernel].launches += 1;
      g_agg[r.kernel].bytes += r.bytes;
    }
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
}

constexpr int kCopyNT = 256, kCopyUnroll = 4;

// 16-byte streaming copy: each thread moves kCopyUnroll vectors per step,
// all loads issued before the stores.
__global__ __launch_bounds__(kCopyNT) void k_copy16(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                    uint64_t n) {
  const uint64_t step = uint64_t(gridDim.x) * kCopyNT * kCopyUnroll;
  for (uint64_t i = uint64_t(blockIdx.x) * kCopyNT * kCopyUnroll + threadIdx.x; i < n; i += step) {
    uint4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const uint64_t j = i + uint64_t(u) * kCopyNT;
      if (j < n) v[u] = in[j];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const uint64_t j = i + uint64_t(u) * kCopyNT;
      if (j < n) out[j] = v[u];
    }
  }
}
}  // namespace

bool prof_on() { return g_prof; }

// Sort variant per record size (8 or 12 bytes): three launches in XCD
// order with 512 threads, the autotuner's choice on RMAT-24 P=8 and on the
// twitter-shaped graph; 8-byte records take 9-bit digits where they save a
// pass (the twitter-shaped graph's 26-bit keys: 3 passes, build 65.0 ->
// 60.6 ms).  Replaced by eg_set_sort_variant / eg_autotune_sort.
static SortVariant g_variant8{kThreeLaunchXcd, 512, 9}, g_variant12{kThreeLaunchXcd, 512, 8};
SortVariant& sort_variant(int rec_bytes) { return rec_bytes <= 8 ? g_variant8 : g_variant12; }

void prof_begin(hipStream_t s) {
  if (!g_prof) return;
  g_pending = take_event();
  if (g_pending) (void)hipEventRecord(g_pending, s);
}

void prof_end(int kernel, uint64_t bytes, hipStream_t s) {
  if (!g_prof || !g_pending) return;
  hipEvent_t b = take_event();
  if (!b) return;
  (void)hipEventRecord(b, s);
  g_recs.push_back(ProfRec{g_pending, b, kernel, bytes});
  g_pending = nullptr;
  if (g_recs.size() > 4096) drain();
}

}  // namespace eg

using namespace eg;

extern "C" {

const char* eg_last_error(void) { return g_last_error; }
int eg_version(void) { return 100; }

int eg_device_alloc(void** d_ptr, uint64_t bytes) {
  if (!d_ptr) return fail(EG_ERR_ARG, "eg_device_alloc: d_ptr is NULL");
  *d_ptr = nullptr;
  hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(EG_ERR_NOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
  }
  return EG_OK;
}

int eg_device_free(void* d_ptr) {
  if (d_ptr) EG_HIP(hipFree(d_ptr));
  return EG_OK;
}

int eg_memcpy_h2d(void* d_dst, const void* h_src, uint64_t bytes, void* stream) {
  if (!bytes) return EG_OK;
  EG_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  EG_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return EG_OK;
}

int eg_memcpy_d2h(void* h_dst, const void* d_src, uint64_t bytes, void* stream) {
  if (!bytes) return EG_OK;
  EG_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
  EG_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return EG_OK;
}

int eg_memcpy_d2d(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
  if (!bytes) return EG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t n = bytes / 16;
  if ((reinterpret_cast<uintptr_t>(d_dst) | reinterpret_cast<uintptr_t>(d_src)) % 16 == 0 && n) {
    // aligned bulk: a plain streaming kernel (also the known-size copy
    // against which HBM byte counters are calibrated)
    const unsigned grid = unsigned(std::min<uint64_t>(ceil_div(n, uint64_t(kCopyNT) * kCopyUnroll), 16384));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kCopyNT), 0, s, static_cast<const uint4*>(d_src),
                       static_cast<uint4*>(d_dst), n);
    EG_HIP(hipGetLastError());
    if (bytes % 16)
      EG_HIP(hipMemcpyAsync(static_cast<char*>(d_dst) + n * 16, static_cast<const char*>(d_src) + n * 16, bytes % 16,
                            hipMemcpyDeviceToDevice, s));
    return EG_OK;
  }
  EG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, s));
  return EG_OK;
}

int eg_profile_enable(int on) {
  if (!on) drain();
  g_prof = on != 0;
  return EG_OK;
}

int eg_profile_reset(void) {
  drain();
  for (auto& a : g_agg) a = ProfAgg{};
  return EG_OK;
}

int eg_profile_read(const char* kernel, double* total_ms, uint64_t* launches, uint64_t* bytes) {
  if (!kernel) return fail(EG_ERR_ARG, "eg_profile_read: kernel name is NULL");
  drain();
  for (int k = 0; k < PK_COUNT; ++k) {
    if (std::strcmp(kernel, kProfNames[k]) == 0) {
      if (total_ms) *total_ms = g_agg[k].ms;
      if (launches) *launches = g_agg[k].launches;
      if (bytes) *bytes = g_agg[k].bytes;
      return EG_OK;
    }
  }
  return fail(EG_ERR_ARG, "unknown kernel class '%s'", kernel);
}

int eg_set_sort_variant(int rec_bytes, int scheme, int threads, int digit_bits) {
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  if (scheme != kThreeLaunch && scheme != kOnePass && scheme != kThreeLaunchXcd)
    return fail(EG_ERR_ARG, "scheme must be 0, 1 or 2");
  if (threads != 256 && threads != 512 && threads != 1024) return fail(EG_ERR_ARG, "threads must be 256, 512 or 1024");
  if (rec_bytes == 12 && threads == 1024) return fail(EG_ERR_ARG, "12-byte records need threads <= 512 (LDS)");
  if (digit_bits != 8 && digit_bits != 9) return fail(EG_ERR_ARG, "digit_bits must be 8 or 9");
  if (digit_bits == 9 && threads != 512) return fail(EG_ERR_ARG, "9-bit digits need 512 threads");
  sort_variant(rec_bytes) = SortVariant{scheme, threads, digit_bits};
  return EG_OK;
}

int eg_get_sort_variant(int rec_bytes, int* scheme, int* threads, int* digit_bits) {
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  const SortVariant v = sort_variant(rec_bytes);
  if (scheme) *scheme = v.scheme;
  if (threads) *threads = v.nt;
  if (digit_bits) *digit_bits = v.bits;
  return EG_OK;
}

int eg_stream_synchronize(void* stream) {
  EG_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return EG_OK;
}

}  // extern "C"
