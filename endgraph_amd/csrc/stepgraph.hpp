// Programs whose steps are decided on the device (the frontier counters of
// one step select the next step's direction and tell when to stop) run as a
// captured HIP graph: kStepBatch steps, each a fixed list of launches
// reading the step number as *base + j, then one launch that advances the
// base and copies the next step's frontier flag, the step count and an
// optional counter into pinned host memory.  The graph is captured once on
// the graph handle's own stream and replayed until that flag is zero (the
// host reads it after each replay, with no copy call); a step after the last
// live one returns at once.  The captured graph is kept for the next call
// with the same arguments (same buffers) and launch list.
#pragma once

#include <vector>

#include "graph_state.hpp"

namespace eg {
namespace {

constexpr uint32_t kStepBatch = 4;
constexpr int kStepNT = 256;

// The graph's last node: advances the base and leaves the next step's
// frontier flag, the step count and the caller's counter in the mailbox.
struct StepRing {
  const unsigned long long* nv[3];
  const uint32_t* steps;
  const unsigned long long* extra;
};

__global__ void k_step_advance(uint32_t* base, uint32_t n, StepRing r, StepMail* mail) {
  const uint32_t b = *base + n;
  *base = b;
  mail->next_nv = *r.nv[b % 3];
  mail->steps = *r.steps;
  mail->extra = r.extra ? *r.extra : 0ull;
  __threadfence_system();
}

template <class Args>
struct StepLaunch {
  void (*fn)(Args, uint32_t);
  unsigned grid;
  unsigned lds = 0;  // dynamic LDS bytes per block
};

// Runs steps until the frontier count of the next step (nv_ring[step % 3])
// is zero or max_steps have run; *h_steps gets *d_steps at the end, and
// *h_extra (when given) *d_extra.
template <class Args, int N>
int run_step_graph(MultiGraph& mg, const Args& a, const StepLaunch<Args> (&ls)[N], uint32_t* d_base,
                   const unsigned long long* const (&nv_ring)[3], const uint32_t* d_steps, uint32_t max_steps,
                   uint32_t* h_steps, hipStream_t s, const unsigned long long* d_extra = nullptr,
                   unsigned long long* h_extra = nullptr) {
  EG_TRY(mg.init());
  StepRing ring{};
  for (int k = 0; k < 3; ++k) ring.nv[k] = nv_ring[k];
  ring.steps = d_steps;
  ring.extra = d_extra;
  std::vector<char> key(reinterpret_cast<const char*>(&a), reinterpret_cast<const char*>(&a) + sizeof(Args));
  for (int k = 0; k < 3; ++k)
    key.insert(key.end(), reinterpret_cast<const char*>(&ring.nv[k]),
               reinterpret_cast<const char*>(&ring.nv[k]) + sizeof(ring.nv[k]));
  key.insert(key.end(), reinterpret_cast<const char*>(&ring.steps),
             reinterpret_cast<const char*>(&ring.steps) + sizeof(ring.steps));
  key.insert(key.end(), reinterpret_cast<const char*>(&ring.extra),
             reinterpret_cast<const char*>(&ring.extra) + sizeof(ring.extra));
  key.insert(key.end(), reinterpret_cast<const char*>(&d_base), reinterpret_cast<const char*>(&d_base) + sizeof(d_base));
  for (int i = 0; i < N; ++i) {  // field by field: the struct has padding
    const void* fn = reinterpret_cast<const void*>(ls[i].fn);
    key.insert(key.end(), reinterpret_cast<const char*>(&fn), reinterpret_cast<const char*>(&fn) + sizeof(fn));
    key.insert(key.end(), reinterpret_cast<const char*>(&ls[i].grid),
               reinterpret_cast<const char*>(&ls[i].grid) + sizeof(ls[i].grid));
    key.insert(key.end(), reinterpret_cast<const char*>(&ls[i].lds),
               reinterpret_cast<const char*>(&ls[i].lds) + sizeof(ls[i].lds));
  }
  if (!mg.exec || mg.key != key) {
    mg.reset_exec();
    EG_HIP(hipStreamBeginCapture(mg.stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t j = 0; j < kStepBatch; ++j)
      for (int i = 0; i < N; ++i)
        hipLaunchKernelGGL(ls[i].fn, dim3(ls[i].grid), dim3(kStepNT), ls[i].lds, mg.stream, a, j);
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, mg.stream, d_base, kStepBatch, ring, mg.mail);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(mg.stream, &graph);
    if (ce != hipSuccess) return fail(EG_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&mg.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      mg.exec = nullptr;
      return fail(EG_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    mg.key = key;
  }
  // the caller's queued work (state set-up) before the graph, the graph
  // before the caller's later work
  EG_HIP(hipEventRecord(mg.ev_in, s));
  EG_HIP(hipStreamWaitEvent(mg.stream, mg.ev_in, 0));
  volatile StepMail* mail = mg.mail;
  uint32_t step = 0;
  for (;;) {
    EG_HIP(hipGraphLaunch(mg.exec, mg.stream));
    step += kStepBatch;
    EG_HIP(hipStreamSynchronize(mg.stream));
    if (!mail->next_nv || step >= max_steps) break;
  }
  *h_steps = mail->steps;
  if (h_extra) *h_extra = mail->extra;
  EG_HIP(hipEventRecord(mg.ev_out, mg.stream));
  EG_HIP(hipStreamWaitEvent(s, mg.ev_out, 0));
  return EG_OK;
}

}  // namespace
}  // namespace eg
