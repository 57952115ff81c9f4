// File I/O around the path: the binary edge-list reader in front of the
// build (preload_graph_file, core/graph.hpp:1011-1084; format README.md:40-41)
// and the vertex-array dump / restore after it (core/graph.hpp:324-375).
//
// Files are read by kIoThreads host threads, each taking every kIoThreads-th
// 32 MiB chunk: pread into one of its two pinned buffers, then an H2D copy
// on its own stream, so reads from the page cache run in parallel and every
// copy overlaps the next read.  The pinned buffers are kept for the
// process's later loads (kIoThreads x 2 x 32 MiB).  Vertex arrays are written
// back with double buffering in the other direction.  The file formats are
// the reference's: packed EdgeType records in file order; a vertex array is
// V * sizeof(T) raw bytes.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

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

constexpr int kIoThreads = 4;
constexpr uint64_t kUpChunk = 32ull << 20;

// Staging of the parallel reader: per thread two pinned buffers, two events
// and a stream, created on first use and kept (one load at a time).
struct UploadPool {
  std::mutex busy;
  bool ready = false;
  void* buf[kIoThreads][2] = {};
  hipEvent_t ev[kIoThreads][2] = {};
  hipEvent_t done[kIoThreads] = {};  // a range's last copy on each stream
  hipStream_t st[kIoThreads] = {};
  // creates what is still missing (a failed attempt keeps what it made)
  int init() {
    if (ready) return EG_OK;
    for (int t = 0; t < kIoThreads; ++t) {
      if (!st[t] && hipStreamCreateWithFlags(&st[t], hipStreamNonBlocking) != hipSuccess) {
        st[t] = nullptr;
        return fail(EG_ERR_HIP, "stream creation for the file reader failed");
      }
      for (int b = 0; b < 2; ++b) {
        if (!buf[t][b] && hipHostMalloc(&buf[t][b], kUpChunk, hipHostMallocDefault) != hipSuccess) {
          buf[t][b] = nullptr;
          return fail(EG_ERR_NOMEM, "pinned host allocation of %llu bytes failed", (unsigned long long)kUpChunk);
        }
        if (!ev[t][b] && hipEventCreateWithFlags(&ev[t][b], hipEventDisableTiming) != hipSuccess) {
          ev[t][b] = nullptr;
          return fail(EG_ERR_HIP, "event creation for the file reader failed");
        }
      }
      if (!done[t] && hipEventCreateWithFlags(&done[t], hipEventDisableTiming) != hipSuccess) {
        done[t] = nullptr;
        return fail(EG_ERR_HIP, "event creation for the file reader failed");
      }
    }
    ready = true;
    return EG_OK;
  }
};

UploadPool& upload_pool() {
  static UploadPool* p = new UploadPool;  // kept to process end
  return *p;
}

}  // namespace

// Host file -> device, range by range: the bytes [bounds[k], bounds[k+1])
// are read by kIoThreads readers (chunk i of the range to reader i mod
// kIoThreads), then stream s is made to wait for their copies and
// on_range(k) runs (it queues work on s that reads range k) while range k+1
// is being read.  Returns with every copy queued; s is ordered after them.
int upload_file_ranges(int fd, const std::vector<uint64_t>& bounds, char* d, hipStream_t s,
                       const std::function<int(int)>& on_range) {
  UploadPool& pool = upload_pool();
  std::lock_guard<std::mutex> lock(pool.busy);
  EG_TRY(pool.init());
  // the caller's earlier work on s (it may still use d) before the copies
  hipEvent_t after_caller = nullptr;
  EG_HIP(hipEventCreateWithFlags(&after_caller, hipEventDisableTiming));
  EG_HIP(hipEventRecord(after_caller, s));
  for (int t = 0; t < kIoThreads; ++t) EG_HIP(hipStreamWaitEvent(pool.st[t], after_caller, 0));
  (void)hipEventDestroy(after_caller);
  uint64_t used[kIoThreads] = {};  // copies issued per reader: its buffers alternate across ranges
  for (size_t k = 0; k + 1 < bounds.size(); ++k) {
    const uint64_t r0 = bounds[k], n = bounds[k + 1] - r0;
    const uint64_t chunks = ceil_div(n, kUpChunk);
    int status[kIoThreads];
    std::string msg[kIoThreads];
    auto reader = [&](int t) {
      status[t] = EG_OK;
      for (uint64_t c = uint64_t(t); c < chunks; c += kIoThreads) {
        const int b = int(used[t]++ & 1);
        const uint64_t off = r0 + c * kUpChunk, len = std::min(kUpChunk, bounds[k + 1] - off);
        hipError_t e = hipEventSynchronize(pool.ev[t][b]);  // the copy that last used buffer b is done
        if (e == hipSuccess && full_pread(fd, static_cast<char*>(pool.buf[t][b]), len, off) != EG_OK) {
          status[t] = EG_ERR_ARG;
          msg[t] = eg_last_error();
          return;
        }
        if (e == hipSuccess) e = hipMemcpyAsync(d + off, pool.buf[t][b], len, hipMemcpyHostToDevice, pool.st[t]);
        if (e == hipSuccess) e = hipEventRecord(pool.ev[t][b], pool.st[t]);
        if (e != hipSuccess) {
          status[t] = EG_ERR_HIP;
          msg[t] = hipGetErrorString(e);
          return;
        }
      }
    };
    const int nt = int(std::min<uint64_t>(kIoThreads, chunks));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(reader, t);
    if (nt > 0) reader(0);
    for (auto& x : th) x.join();
    for (int t = 0; t < nt; ++t)
      if (status[t] != EG_OK) return fail(status[t], "file reader: %s", msg[t].c_str());
    for (int t = 0; t < nt; ++t) {
      EG_HIP(hipEventRecord(pool.done[t], pool.st[t]));
      EG_HIP(hipStreamWaitEvent(s, pool.done[t], 0));
    }
    EG_TRY(on_range(int(k)));
  }
  return EG_OK;
}

namespace {

// host file [0, n) -> device with kIoThreads readers
int upload_file(int fd, uint64_t n, char* d, hipStream_t s) {
  if (!n) return EG_OK;
  EG_TRY(upload_file_ranges(fd, std::vector<uint64_t>{0, n}, d, s, [](int) { return EG_OK; }));
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
