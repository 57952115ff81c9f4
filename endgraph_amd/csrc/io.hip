// File I/O around the path: the binary edge-list reader in front of the
// build (preload_graph_file, core/graph.hpp:1011-1084; format README.md:40-41)
// and the vertex-array dump / restore after it (core/graph.hpp:324-375).
//
// Edge files are read in 64 MiB chunks (pread) into two pinned host buffers;
// the H2D copy of one chunk runs while the next chunk is read, so the upload
// hides behind the file read.  Vertex arrays use the same double buffering
// in the other direction.  The file formats are the reference's: packed
// EdgeType records in file order; a vertex array is V * sizeof(T) raw bytes.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>

#include "common.hpp"

namespace eg {
namespace {

constexpr uint64_t kIoChunk = 64ull << 20;

struct Pinned {
  void* p[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Pinned() {
    for (int i = 0; i < 2; ++i) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (p[i]) (void)hipHostFree(p[i]);
    }
  }
  int init(uint64_t bytes) {
    for (int i = 0; i < 2; ++i) {
      if (hipHostMalloc(&p[i], bytes, hipHostMallocDefault) != hipSuccess) {
        p[i] = nullptr;
        return fail(EG_ERR_NOMEM, "pinned host allocation of %llu bytes failed", (unsigned long long)bytes);
      }
      EG_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    return EG_OK;
  }
};

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) ::close(fd);
  }
};

int full_pread(int fd, char* dst, uint64_t n, uint64_t off) {
  while (n) {
    const ssize_t r = ::pread(fd, dst, std::min<uint64_t>(n, 1ull << 30), off_t(off));
    if (r <= 0) return fail(EG_ERR_ARG, "short read at offset %llu", (unsigned long long)off);
    dst += r;
    off += uint64_t(r);
    n -= uint64_t(r);
  }
  return EG_OK;
}

int full_pwrite(int fd, const char* src, uint64_t n, uint64_t off) {
  while (n) {
    const ssize_t r = ::pwrite(fd, src, std::min<uint64_t>(n, 1ull << 30), off_t(off));
    if (r <= 0) return fail(EG_ERR_ARG, "short write at offset %llu", (unsigned long long)off);
    src += r;
    off += uint64_t(r);
    n -= uint64_t(r);
  }
  return EG_OK;
}

// host file [0, n) -> device, chunk k's copy overlapping chunk k+1's read
int upload_file(int fd, uint64_t n, char* d, hipStream_t s) {
  if (!n) return EG_OK;
  Pinned pin;
  EG_TRY(pin.init(std::min(n, kIoChunk)));
  for (uint64_t off = 0, k = 0; off < n; off += kIoChunk, ++k) {
    const int b = int(k & 1);
    const uint64_t len = std::min(kIoChunk, n - off);
    EG_HIP(hipEventSynchronize(pin.ev[b]));  // the copy that last used buffer b is done
    EG_TRY(full_pread(fd, static_cast<char*>(pin.p[b]), len, off));
    EG_HIP(hipMemcpyAsync(d + off, pin.p[b], len, hipMemcpyHostToDevice, s));
    EG_HIP(hipEventRecord(pin.ev[b], s));
  }
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

// device [0, n) -> host file, chunk k+1's copy overlapping chunk k's write
int download_file(int fd, uint64_t n, const char* d, hipStream_t s) {
  if (!n) return EG_OK;
  Pinned pin;
  EG_TRY(pin.init(std::min(n, kIoChunk)));
  const uint64_t chunks = ceil_div(n, kIoChunk);
  auto issue = [&](uint64_t k) -> int {
    const int b = int(k & 1);
    const uint64_t off = k * kIoChunk, len = std::min(kIoChunk, n - off);
    EG_HIP(hipMemcpyAsync(pin.p[b], d + off, len, hipMemcpyDeviceToHost, s));
    EG_HIP(hipEventRecord(pin.ev[b], s));
    return EG_OK;
  };
  EG_TRY(issue(0));
  for (uint64_t k = 0; k < chunks; ++k) {
    const int b = int(k & 1);
    EG_HIP(hipEventSynchronize(pin.ev[b]));
    if (k + 1 < chunks) EG_TRY(issue(k + 1));
    const uint64_t off = k * kIoChunk, len = std::min(kIoChunk, n - off);
    EG_TRY(full_pwrite(fd, static_cast<const char*>(pin.p[b]), len, off));
  }
  return EG_OK;
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_edge_file_records(const char* path, int rec_bytes, uint64_t* h_records) {
  if (!path || !h_records) return fail(EG_ERR_ARG, "path and h_records are required");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  struct stat st;
  if (::stat(path, &st) != 0) return fail(EG_ERR_ARG, "cannot stat %s", path);
  if (uint64_t(st.st_size) % uint64_t(rec_bytes))
    return fail(EG_ERR_ARG, "%s: %llu bytes is not a whole number of %d-byte records", path,
                (unsigned long long)st.st_size, rec_bytes);
  *h_records = uint64_t(st.st_size) / uint64_t(rec_bytes);
  return EG_OK;
}

int eg_load_edge_file(const char* path, int rec_bytes, void* d_edges, uint64_t capacity_records, uint64_t* h_records,
                      void* stream) {
  uint64_t n = 0;
  EG_TRY(eg_edge_file_records(path, rec_bytes, &n));
  if (n > capacity_records) return fail(EG_ERR_ARG, "%s holds %llu records, the buffer %llu", path,
                                        (unsigned long long)n, (unsigned long long)capacity_records);
  if (!d_edges && n) return fail(EG_ERR_ARG, "d_edges is NULL");
  Fd f;
  f.fd = ::open(path, O_RDONLY);
  if (f.fd < 0) return fail(EG_ERR_ARG, "cannot open %s", path);
  EG_TRY(upload_file(f.fd, n * uint64_t(rec_bytes), static_cast<char*>(d_edges), static_cast<hipStream_t>(stream)));
  if (h_records) *h_records = n;
  return EG_OK;
}

int eg_dump_vertex_array(const void* d_array, uint64_t elem_bytes, uint32_t V, const char* path, void* stream) {
  if (!path || (!d_array && V) || !elem_bytes) return fail(EG_ERR_ARG, "d_array, elem_bytes and path are required");
  Fd f;
  f.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (f.fd < 0) return fail(EG_ERR_ARG, "cannot create %s", path);
  return download_file(f.fd, elem_bytes * V, static_cast<const char*>(d_array), static_cast<hipStream_t>(stream));
}

int eg_restore_vertex_array(void* d_array, uint64_t elem_bytes, uint32_t V, const char* path, void* stream) {
  if (!path || (!d_array && V) || !elem_bytes) return fail(EG_ERR_ARG, "d_array, elem_bytes and path are required");
  struct stat st;
  if (::stat(path, &st) != 0) return fail(EG_ERR_ARG, "cannot stat %s", path);
  if (uint64_t(st.st_size) != elem_bytes * V)
    return fail(EG_ERR_ARG, "%s holds %llu bytes, not vertices_num * %llu", path, (unsigned long long)st.st_size,
                (unsigned long long)elem_bytes);
  Fd f;
  f.fd = ::open(path, O_RDONLY);
  if (f.fd < 0) return fail(EG_ERR_ARG, "cannot open %s", path);
  return upload_file(f.fd, elem_bytes * V, static_cast<char*>(d_array), static_cast<hipStream_t>(stream));
}

}  // extern "C"
