// Per-graph tuning of the device-driven traversals (no reference counterpart:
// the reference's schedule is fixed by its thread chunks).  The auto
// policy's mode thresholds of the four programs, the single-root push's hub
// threshold and chunk size, and the batched push's work-item size are timed
// on sample roots of the graph itself and the fastest kept.  None of them
// changes a result: directions and work splits only reorder the same
// relaxations, whose fixpoints (depths, min-id parents, fp32 distances) are
// unique.
#include <algorithm>
#include <chrono>
#include <vector>

#include "graph_state.hpp"

namespace eg {
namespace {

// SplitMix64: deterministic sample roots from the caller's seed
uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void to_api(const EngineTune& t, eg_engine_tune* o) {
  o->bfs_dense_div = t.sbfs_div;
  o->sssp_dense_div = t.dsssp_div;
  o->bfs_multi_dense_div = t.mbfs_div;
  o->sssp_multi_dense_div = t.msssp_div;
  o->hub_degree = t.hub;
  o->hub_chunk = t.hub_chunk;
  o->multi_item_edges = t.m_item;
}

int from_api(const eg_engine_tune& o, EngineTune* t) {
  if (!o.bfs_dense_div || !o.sssp_dense_div || !o.bfs_multi_dense_div || !o.sssp_multi_dense_div || !o.hub_chunk ||
      !o.multi_item_edges)
    return fail(EG_ERR_ARG, "engine tune: divisors and sizes must be positive");
  if (o.hub_degree < kSbfsLaneEdges)
    return fail(EG_ERR_ARG, "engine tune: hub_degree below %u (vertices up to it run in one lane)", kSbfsLaneEdges);
  t->sbfs_div = o.bfs_dense_div;
  t->dsssp_div = o.sssp_dense_div;
  t->mbfs_div = o.bfs_multi_dense_div;
  t->msssp_div = o.sssp_multi_dense_div;
  t->hub = o.hub_degree;
  t->hub_chunk = o.hub_chunk;
  t->m_item = o.multi_item_edges;
  return EG_OK;
}

// Wall time of `reps` calls after one untimed call (which also captures the
// step graph the candidate's arguments key).
template <class F>
int time_ms(F f, int reps, double* ms) {
  EG_TRY(f());
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; ++i) EG_TRY(f());
  *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
  return EG_OK;
}

// Tries each candidate value of one field (the others as tuned so far) and
// keeps the fastest.
template <class Set, class Run>
int pick(eg_graph* g, const std::vector<uint32_t>& cands, Set set, Run run, double* best_ms) {
  EngineTune best = g->tune;
  double best_t = 1e300;
  for (uint32_t c : cands) {
    set(g->tune, c);
    double t = 0;
    EG_TRY(time_ms(run, 2, &t));
    if (t < best_t) {
      best_t = t;
      best = g->tune;
    }
  }
  g->tune = best;
  if (best_ms) *best_ms = best_t;
  return EG_OK;
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_get_engine_tune(const eg_graph* g, eg_engine_tune* out) {
  if (!g || !out) return fail(EG_ERR_ARG, "eg_get_engine_tune: NULL argument");
  to_api(g->tune, out);
  return EG_OK;
}

int eg_set_engine_tune(eg_graph* g, const eg_engine_tune* t) {
  if (!g || !t) return fail(EG_ERR_ARG, "eg_set_engine_tune: NULL argument");
  return from_api(*t, &g->tune);
}

int eg_autotune_engine(eg_graph* g, int n_roots, uint64_t seed, eg_engine_tune* h_out, double* h_ms, void* stream) {
  if (!g) return fail(EG_ERR_ARG, "eg_autotune_engine: graph is NULL");
  if (n_roots < 1 || n_roots > 64) return fail(EG_ERR_ARG, "n_roots must be in 1..64");
  hipStream_t s = static_cast<hipStream_t>(stream);
  EG_TRY(eg_engine_prepare(g, nullptr, stream));
  EG_TRY(eg_traversal_view_prepare(g, nullptr, stream));
  const uint32_t V = g->V;
  // sample roots: the max-out-degree vertex, then seeded vertices with out-edges
  std::vector<uint32_t> deg(V);
  EG_HIP(hipMemcpyAsync(deg.data(), g->out_deg.ptr, uint64_t(V) * 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> roots{uint32_t(std::max_element(deg.begin(), deg.end()) - deg.begin())};
  uint64_t x = seed;
  for (int tries = 0; int(roots.size()) < n_roots && tries < 64 * n_roots + 1024; ++tries) {
    const uint32_t v = uint32_t(splitmix(x) % V);
    if (deg[v]) roots.push_back(v);
  }
  const int nr = int(roots.size());
  const bool weighted = g->rec == 12;
  DevBuf parent, depth, dist, mpar, mdep, mdist;
  EG_TRY(parent.resize(uint64_t(V) * 4));
  EG_TRY(depth.resize(uint64_t(V) * 4));
  if (weighted) EG_TRY(dist.resize(uint64_t(V) * 4));
  const int single = std::min(nr, 4);  // single-root programs time the first roots
  auto bfs = [&]() -> int {
    for (int i = 0; i < single; ++i)
      EG_TRY(eg_bfs(g, roots[i], EG_POLICY_AUTO, parent.as<uint32_t>(), depth.as<uint32_t>(), nullptr, nullptr,
                    stream));
    return EG_OK;
  };
  auto sssp = [&]() -> int {
    for (int i = 0; i < single; ++i)
      EG_TRY(eg_sssp(g, roots[i], EG_POLICY_AUTO, dist.as<float>(), nullptr, stream));
    return EG_OK;
  };
  auto single_all = [&]() -> int {
    EG_TRY(bfs());
    return weighted ? sssp() : EG_OK;
  };
  double ms[5] = {0, 0, 0, 0, 0};
  // single-root: the BFS switch, then the hub split (both programs), then SSSP's switch
  EG_TRY(pick(g, {5, 10, 20, 40, 80}, [](EngineTune& t, uint32_t c) { t.sbfs_div = c; }, bfs, &ms[0]));
  const std::vector<std::pair<uint32_t, uint32_t>> hubs{{1024, 256}, {1024, 1024}, {4096, 1024}, {4096, 4096},
                                                        {16384, 4096}};
  {
    EngineTune best = g->tune;
    double best_t = 1e300;
    for (const auto& h : hubs) {
      g->tune.hub = h.first;
      g->tune.hub_chunk = h.second;
      double t = 0;
      EG_TRY(time_ms(single_all, 2, &t));
      if (t < best_t) {
        best_t = t;
        best = g->tune;
      }
    }
    g->tune = best;
    ms[1] = best_t;
  }
  if (weighted)
    EG_TRY(pick(g, {10, 20, 40, 80, 160}, [](EngineTune& t, uint32_t c) { t.dsssp_div = c; }, sssp, &ms[2]));
  EG_TRY(ensure_hub_items(g, s));
  // batched: all sample roots at once
  EG_TRY(mpar.resize(uint64_t(V) * 64 * 4));
  EG_TRY(mdep.resize(uint64_t(V) * 64 * 4));
  auto mbfs = [&]() -> int {
    return eg_bfs_multi(g, roots.data(), nr, EG_POLICY_AUTO, mpar.as<uint32_t>(), mdep.as<uint32_t>(), nullptr,
                        nullptr, nullptr, stream);
  };
  EG_TRY(pick(g, {1024, 4096, 16384}, [](EngineTune& t, uint32_t c) { t.m_item = c; }, mbfs, nullptr));
  EG_TRY(pick(g, {5, 10, 20, 40}, [](EngineTune& t, uint32_t c) { t.mbfs_div = c; }, mbfs, &ms[3]));
  if (weighted) {
    mpar.release();
    mdep.release();
    EG_TRY(mdist.resize(uint64_t(V) * 64 * 4));
    auto msssp = [&]() -> int {
      return eg_sssp_multi(g, roots.data(), nr, EG_POLICY_AUTO, mdist.as<float>(), nullptr, stream);
    };
    EG_TRY(pick(g, {1, 2, 5, 20}, [](EngineTune& t, uint32_t c) { t.msssp_div = c; }, msssp, &ms[4]));
  }
  if (h_out) to_api(g->tune, h_out);
  if (h_ms)
    for (int i = 0; i < 5; ++i) h_ms[i] = ms[i];
  return EG_OK;
}

}  // extern "C"
