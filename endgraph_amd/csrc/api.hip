// C-ABI glue: error reporting, owned device buffers, memory helpers.
#include <cstdarg>
#include <cstring>

#include "common.hpp"

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
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
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
  EG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
  return EG_OK;
}

int eg_stream_synchronize(void* stream) {
  EG_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return EG_OK;
}

}  // extern "C"
