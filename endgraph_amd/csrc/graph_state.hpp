// Device-resident state of one built graph (the eg_graph handle).
#pragma once

#include <vector>

#include "common.hpp"
#include "radix.hpp"

// A captured HIP graph of the batched programs' steps, its stream and the
// events that order it after and before the caller's stream.
// What the last node of a replayed step graph leaves for the host, in
// pinned host memory (no copy call per replay).
struct StepMail {
  unsigned long long next_nv;  // frontier flag of the step after the replay
  unsigned long long extra;    // a program counter the caller names (or 0)
  uint32_t steps;              // steps that had a frontier
};

struct MultiGraph {
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t exec = nullptr;
  StepMail* mail = nullptr;  // pinned host memory, written by the graph
  std::vector<char> key;     // kernel arguments, grids and kernels it was captured with
  int init() {
    if (stream && mail) return EG_OK;
    if ((!stream && hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) ||
        (!ev_in && hipEventCreateWithFlags(&ev_in, hipEventDisableTiming) != hipSuccess) ||
        (!ev_out && hipEventCreateWithFlags(&ev_out, hipEventDisableTiming) != hipSuccess))
      return eg::fail(EG_ERR_HIP, "stream or event creation for the step graph failed");
    if (!mail && hipHostMalloc(reinterpret_cast<void**>(&mail), sizeof(StepMail), hipHostMallocDefault) != hipSuccess) {
      mail = nullptr;
      (void)hipGetLastError();
      return eg::fail(EG_ERR_NOMEM, "pinned host memory for the step graph failed");
    }
    return EG_OK;
  }
  void reset_exec() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
  ~MultiGraph() {
    reset_exec();
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
    if (mail) (void)hipHostFree(mail);
  }
};

// Sizes and auto-policy thresholds of the device-driven traversals, per
// graph (eg_autotune_engine; the defaults are the sweeps' picks on R-MAT and
// the twitter-shaped graph).  None changes a result, only the schedule.
struct EngineTune {
  uint32_t sbfs_div = 20;     // single-root BFS: dense once the frontier's out-edges reach E / sbfs_div
  uint32_t dsssp_div = 80;    // single-root SSSP: the same for the active set
  uint32_t mbfs_div = 20;     // batched BFS
  uint32_t msssp_div = 2;     // batched SSSP
  uint32_t hub = 4096;        // single-root push: out-degree above which a vertex is cut into chunks
  uint32_t hub_chunk = 1024;  // out-edges per such chunk
  uint32_t m_item = 4096;     // batched push: out-edges per work item
};

struct eg_graph {
  // configuration (Graph(int eta), core/graph.hpp:202; P ranks and S sockets)
  uint32_t V = 0;
  int P = 1, S = 1, Q = 1;  // Q = P*S sub-partitions
  uint64_t eta = 0;
  uint32_t m = 1;
  int rec = 8;  // 8 = EdgeType<Empty>, 12 = EdgeType<float>
  uint32_t pages = 0;

  // result of the last build
  bool built = false;
  int symmetrized = 0;
  uint64_t E0 = 0;  // records in the input list
  uint64_t E = 0;   // records after symmetrisation
  int vbits = 1, qbits = 0;
  // bytes the reference's ranks would exchange while preprocessing
  // (prep_data, core/graph.hpp:716, :908)
  uint64_t prep_data = 0;
  uint64_t pr_cross_rows = ~0ull;    // pull rows whose destination another partition owns (~0: not counted yet)

  eg::DevBuf out_deg, in_deg;        // u32[V]
  eg::DevBuf page_prefix;            // u64[pages+1]: page weights, then their exclusive prefix
  eg::DevBuf meta;                   // small device tables, see Meta
  eg::DevBuf qmap;                   // u16[pages]: sub-partition of each page
  eg::DevBuf csc_list, csr_list;     // records, sub-partitions back to back
  eg::DevBuf csc_cidx;               // CIdx, per sub-partition: entries + sentinel
  eg::DevBuf csr_index;              // u64[Q][V+1] (sub-partition-local positions)
  eg::DevBuf csr_bitmap;             // u64[Q][(V>>6)+1]
  eg::DevBuf blk;                    // block tables of the P > 1 regrouping
  eg::DevBuf idx_tmp;                // long-gap pieces of the dense CSR index build
  eg::DevBuf sym;                    // symmetrised logical list
  eg::DevBuf t1, t2, t3;             // first-sort results and sort ping-pong
  eg::DevBuf hist, scan_tmp, tile_cnt, err;
  eg::RxWork sortw;                  // segmented radix sort workspace

  // Engine view, derived from the built arrays on the first engine call
  // (eg_engine_prepare).  Edge positions are u32: a build holds < 2^32 edges.
  bool eng_ready = false;
  uint64_t R = 0;                    // pull rows: compressed CSC entries of all sub-partitions
  eg::DevBuf e_csc_src;              // u32[E]   source of every CSC edge (SoA of csc_list)
  eg::DevBuf e_csc_w;                // f32[E]   weights (12-byte records only)
  eg::DevBuf e_rowflag;              // u64[ceil(E/64)]: bit e set iff CSC edge e starts a row
  eg::DevBuf e_wordrank;             // u32[ceil(E/64)+1]: rows started before word w
  eg::DevBuf e_row_dst;              // u32[R]   destination of each row
  eg::DevBuf e_row_start;            // u32[R+1] first edge of each row
  eg::DevBuf e_out_ptr;              // u32[V+1] merged CSR over all sub-partitions
  eg::DevBuf e_out_dst;              // u32[E]
  eg::DevBuf e_out_w;                // f32[E]
  eg::DevBuf e_deg8;                 // u8[V]    out-degree saturated at 255 (BFS level ends)
  eg::DevBuf e_hub_items;            // u64 (v << 32 | chunk): out-edge chunks of the BFS push hubs
  uint64_t e_nhub_items = 0;
  uint32_t e_hub_for[2] = {0, 0};    // (hub, hub_chunk) the list was made for
  EngineTune tune;
  eg::DevBuf e_scratch;              // frontier list, counters
  eg::DevBuf e_msg;                  // per pull row values and a counter (eg_dense_messages)
  eg::DevBuf e_bm[3];                // three bitmaps of (V>>6)+1 words
  eg::DevBuf e_pr;                   // PageRank work vectors
  // degree-ordered PageRank view: vertices renumbered by descending
  // out-degree (ties by id); either the built CSC's rows with every source
  // and destination replaced by its new id (row bits shared with the
  // engine's), or the merged rows in new-id order with sources ascending
  bool pv_ready = false;
  uint32_t pv_R = 0;
  double pv_prepare_ms = 0;
  eg::DevBuf pv_perm;                // u32[V]  old id -> new id
  eg::DevBuf pv_inv;                 // u32[V]  new id -> old id
  eg::DevBuf pv_deg;                 // u32[V]  out-degree by new id
  eg::DevBuf pv_src;                 // u32[E + 4] new source id of every edge of the view's rows
  eg::DevBuf pv_row_dst;             // u32[R]  new destination id of each row
  eg::DevBuf pv_rowflag;             // u64[ceil(E/64)+1] row starts of the sorted-row form (the other form uses the engine's)
  eg::DevBuf pv_wordrank;            // u32[ceil(E/64)+2]
  int pv_form = -1;                  // the form pv_ready holds: 0 relabelled CSC rows, 1 rows sorted by new source id
  uint32_t pv_hot = 0;               // FP32C pull: values of new ids [0, pv_hot) staged in LDS per block
  uint32_t pv_hot_grid = 0;          // ... by a persistent grid of this many blocks
  // traversal view for the batched (64-root) programs: the CSC merged over
  // sub-partitions with each destination's in-edges sorted by source
  bool tv_ready = false;
  eg::DevBuf tv_ptr;                 // u32[V+1] row pointer (prefix of in-degree)
  eg::DevBuf tv_src;                 // u32[E + 4] (4 words of padding)
  eg::DevBuf tv_w;                   // f32[E] (12-byte records only)
  eg::DevBuf tv_items;               // u64 (v << 32 | chunk): in-edge chunks of at most 512 edges
  uint64_t tv_nitems = 0;
  eg::DevBuf tv_big;                 // first chunks of the vertices with more than kTvSmallRow in-edges
  uint64_t tv_nbig = 0;
  eg::DevBuf tv_big_rest;            // their other chunks
  uint64_t tv_nbig_rest = 0;
  double tv_prepare_ms = 0;
  // transposed view for the PPR adjoint: CSC slots grouped by source
  bool ad_ready = false;
  uint32_t ad_R = 0;
  eg::DevBuf ad_slot_dst;            // u32[E] destination of every CSC slot
  eg::DevBuf ad_slot;                // u32[E] slot of transposed position k
  eg::DevBuf ad_dst;                 // u32[E] destination of that slot (gather index)
  eg::DevBuf ad_rowflag, ad_wordrank;
  eg::DevBuf ad_row_src;             // u32[R] source of each transposed row
  eg::DevBuf e_ppr, e_ppr_adj;       // PPR work vectors
  eg::DevBuf e_sbfs;                 // single-root BFS state (bitmaps, push items, counters)
  eg::DevBuf e_multi;                // batched traversal masks, work items, counters
  MultiGraph multi_graph;            // captured steps of the device-driven programs
  MultiGraph sssp_graph;             // captured steps of the single-root SSSP
  eg::DevBuf e_dsssp;                // single-root SSSP state (bitmaps, counters)
  eg::DevBuf e_multi_roots;          // roots and per-root statistics
  double eng_prepare_ms = 0;

  // host copies of the small tables
  std::vector<uint32_t> page_offset, partition_offset;       // P+1
  std::vector<uint32_t> local_partition_offset;              // P*(S+1)
  std::vector<uint64_t> q_off_csc, q_off_csr;                // Q+1 (edges per sub-partition, prefix)
  std::vector<uint64_t> cidx_off, cidx_cnt;                  // Q (entries incl. sentinel)

  // Layout of `meta` (u64 words unless noted)
  struct Meta {
    uint64_t q_cnt_csc[256 + 1];
    uint64_t q_cnt_csr[256 + 1];
    uint64_t q_off_csc[256 + 1];
    uint64_t q_off_csr[256 + 1];
    uint64_t cidx_off[256];
    uint64_t cidx_cnt[256];
    uint32_t page_offset[256 + 1];
    uint32_t partition_offset[256 + 1];
    uint32_t local_page_offset[256 * 9];  // absolute page ids, P*(S+1)
    uint32_t local_partition_offset[256 * 9];
  };
  Meta* dmeta() const { return meta.as<Meta>(); }
};

namespace eg {
// Derives the engine's view after a build (first engine call; engine.hip).
int engine_ensure(eg_graph* g, hipStream_t s);
// The single-root push's hub chunks for the graph's current tune (engine.hip).
int ensure_hub_items(eg_graph* g, hipStream_t s);
// Rows of in-edges by destination with sources ascending and the pull work
// items over them (first traversal that needs it; multi.hip).
int ensure_tview(eg_graph* g, hipStream_t s);
// In-edges per pull work item, and the in-degree up to which a single-root
// BFS pull scans a row in one lane instead (one item's worth: most rows end
// at their first source; twitter-shaped BFS 3.76 ms at 32, 3.50 at 128,
// 3.43 at 512; RMAT-24 0.78 -> 0.75 ms).
constexpr uint32_t kTvPullEdges = 512;
// Out-degree up to which a single-root push relaxes a frontier vertex in one
// lane; the hub threshold of EngineTune may not go below it.
constexpr uint32_t kSbfsLaneEdges = 16;
constexpr uint32_t kTvSmallRow = 512;
}  // namespace eg
