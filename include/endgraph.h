/*
 * endgraph.h - C ABI of the MI355X-native EndGraph preprocessing and Gemini
 * edge-processing engine (libendgraph.so, gfx950).
 *
 * The reference boundary is the header-only C++ template Graph<EdgeData>
 * (core/graph.hpp:100-1819).  Each entry point below names the reference
 * member it replaces.  Conventions:
 *   - every function returns an eg_status (0 = EG_OK) and never aborts; the
 *     message of the last failure on the calling thread is eg_last_error();
 *   - pointers named d_* are device (HBM) pointers, h_* are host pointers;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     results are ready when the stream has drained, unless a function says
 *     it synchronises;
 *   - the graph handle owns every array it hands out; they stay valid until
 *     the next eg_graph_build or eg_graph_destroy on that handle.
 * Records are the reference's packed layouts (core/type.hpp:23-74):
 *   EdgeType<Empty> {u32 src; u32 dst}            8 bytes
 *   EdgeType<float> {u32 src; u32 dst; f32 w}     12 bytes
 *   CompressedAdjIndexType {u64 index; u32 vertex} 12 bytes, packed
 */
#ifndef ENDGRAPH_H
#define ENDGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eg_graph eg_graph;

enum eg_status {
  EG_OK = 0,
  EG_ERR_ARG = -1,    /* invalid argument or configuration */
  EG_ERR_RANGE = -2,  /* a vertex id >= vertices_num in the edge list */
  EG_ERR_NOMEM = -3,  /* device allocation failed */
  EG_ERR_HIP = -4,    /* HIP runtime error (message names the call) */
  EG_ERR_STATE = -5   /* handle not built / wrong record type for the call */
};

/* Message of the last failure on this thread ("" if none). */
const char* eg_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int eg_version(void);

/* ---- device memory helpers (for hosts without their own HIP runtime) -- */
int eg_device_alloc(void** d_ptr, uint64_t bytes);
int eg_device_free(void* d_ptr);
int eg_memcpy_h2d(void* d_dst, const void* h_src, uint64_t bytes, void* stream);
int eg_memcpy_d2h(void* h_dst, const void* d_src, uint64_t bytes, void* stream); /* synchronises */
int eg_memcpy_d2d(void* d_dst, const void* d_src, uint64_t bytes, void* stream);
int eg_stream_synchronize(void* stream);
/* fill_vertex_array (core/graph.hpp:289-294) for any 1/2/4/8-byte element:
 * n copies of *h_value.  eg_device_iota: d[i] = i (cc.cpp:35-41's label). */
int eg_device_fill(void* d_ptr, uint64_t n, const void* h_value, int value_bytes, void* stream);
int eg_device_iota(uint32_t* d_ptr, uint32_t n, void* stream);

/* ---- kernel timing (no reference counterpart; the reference prints
 *      wall-clock timers only, core/graph.hpp:1367-1371) -------------------
 * While enabled, selected launches are bracketed by HIP events on their own
 * stream.  eg_profile_read returns, for a kernel class ("radix_upsweep",
 * "radix_downsweep", "scan", "degree", "cindex", "csr_index", "pr_pull",
 * "bfs"), the summed event time, the launch count and the summed algorithmic
 * bytes of those launches.  Not thread-safe. */
int eg_profile_enable(int on);
int eg_profile_reset(void);
int eg_profile_read(const char* kernel, double* total_ms, uint64_t* launches, uint64_t* bytes);

/* ---- radix-sort variant (no reference counterpart) ------------------------
 * Every stable sort of the build is a segmented LSD radix sort with three
 * pass schemes (0: upsweep/scan/downsweep with tiles in block order, 1: one
 * pass per digit with decoupled look-back, 2: as 0 with each XCD taking one
 * contiguous range of tiles), a tile of threads x 16 records (threads 256, 512 or 1024;
 * 1024 only for 8-byte records) and digits of up to 8 or 9 bits (9 only with
 * 512 threads, and taken by a sort only where it saves a pass, as for 25-27
 * bit keys).  Process-wide per record size; chosen by eg_autotune_sort or
 * set here.  Results never depend on the variant. */
int eg_set_sort_variant(int rec_bytes, int scheme, int threads, int digit_bits);
int eg_get_sort_variant(int rec_bytes, int* scheme, int* threads, int* digit_bits);
/* Builds g from the given list once per candidate variant (twice each, the
 * second timed), keeps the fastest for g's record size and leaves g built
 * with it.  Synchronises. */
int eg_autotune_sort(eg_graph* g, const void* d_edges, uint64_t E0, int symmetrize, int* h_scheme, int* h_threads,
                     int* h_digit_bits, double* h_ms, void* stream);

/* ---- synthetic inputs (no reference counterpart: the reference reads
 *      binary edge files, README.md:40-41; these write the same records) -- */
/* R-MAT: 2^scale vertices, E records, quadrant probabilities a,b,c
 * (d = 1-a-b-c), counter-based RNG keyed by seed.  rec_bytes 12 adds a
 * float weight uniform in [w_lo, w_hi). */
int eg_gen_rmat(void* d_edges, uint64_t E, uint32_t scale, double a, double b, double c, uint64_t seed,
                int rec_bytes, float w_lo, float w_hi, void* stream);
/* Power-law ("twitter-shaped"): src and dst ranks drawn from (rank+c)^-beta
 * weights with beta = 1/(alpha-1), alpha_out for src and alpha_in for dst,
 * offsets c chosen so that the expected largest out/in degree is
 * max_out/max_in, ranks mapped to ids by seeded bijections of [0,V). */
int eg_gen_powerlaw(void* d_edges, uint64_t E, uint32_t V, double alpha_out, double alpha_in, double max_out,
                    double max_in, uint64_t seed, int rec_bytes, void* stream);

/* Erdos-Renyi batch: n_graphs graphs back to back, graph g with
 * h_vertices[g] vertices and h_edges[g] records of uniform local ids
 * (rec_bytes 12 adds a weight uniform in [0,1)).  Synchronises. */
int eg_gen_er_batch(void* d_edges, int n_graphs, const uint32_t* h_vertices, const uint64_t* h_edges, uint64_t seed,
                    int rec_bytes, void* stream);

/* ---- graph construction ------------------------------------------------ */
/* Graph<EdgeData>(eta) (core/graph.hpp:202) with the partition/socket
 * counts of the MPI world / NUMA topology made explicit: P partitions (ranks)
 * and S sockets per partition, all on this GPU.  m multiplies the degree in
 * the CPBT page weight eta*(P-1 + m*deg) (core/graph.hpp:733-738: m =
 * iterations for pull workloads, 1 otherwise).  rec_bytes: 8 or 12.
 * Limits: 1 <= P, 1 <= S <= 8, P*S <= 256. */
int eg_graph_create(uint32_t vertices_num, int P, int S, uint64_t eta, uint32_t m, int rec_bytes,
                    eg_graph** out);
int eg_graph_destroy(eg_graph* g);

/* load_directed / load_undirected_from_directed (core/graph.hpp:1358, 629)
 * from an in-memory device edge list of E0 records in file order (the
 * reference's timed prep region, split_graph, :981-1008).  symmetrize=1
 * appends every rank slice's reversed copy (:1057-1078).  Builds both
 * directions (CSC: prepare_pull_mode :1098; CSR: prepare_push_mode :1248),
 * the degree arrays and the partition offsets.  Synchronises `stream`.
 * Limit: E (2*E0 when symmetrised) < 2^32 records; the engine's edge
 * positions are u32 (EG_ERR_ARG beyond). */
int eg_graph_build(eg_graph* g, const void* d_edges, uint64_t E0, int symmetrize, void* stream);
/* load_directed / load_undirected_from_directed from a binary edge file
 * (preload_graph_file, core/graph.hpp:1011-1084, then the build): the file's
 * records (eg_edge_file_records of them, at most capacity_records) are read
 * into d_edges rank slice by rank slice (:1024-1030), and each slice's
 * reversed copy and first sorts run on the GPU while the next slice is read;
 * the result equals eg_graph_build on the same records.  With P = 1 the one
 * slice is read whole first.  Synchronises `stream`. */
int eg_graph_build_file(eg_graph* g, const char* path, void* d_edges, uint64_t capacity_records, int symmetrize,
                        uint64_t* h_records, void* stream);

typedef struct eg_graph_info {
  uint32_t vertices_num;
  int partitions, sockets, record_bytes;
  uint64_t edges_num;  /* after symmetrisation */
  uint32_t pages_num;
  int symmetrized;
} eg_graph_info;
int eg_graph_get_info(const eg_graph* g, eg_graph_info* out);

/* partition_offset[P+1] (core/graph.hpp:744-749, boundaries clamped to V). */
int eg_graph_partition_offset(const eg_graph* g, uint32_t* h_out);
/* local_partition_offset[S+1] of partition p (core/graph.hpp:774-780). */
int eg_graph_local_partition_offset(const eg_graph* g, int p, uint32_t* h_out);
/* degree (out-degree, core/graph.hpp:697-706) and in-degree, u32[V]. */
int eg_graph_degree(const eg_graph* g, const uint32_t** d_out_degree, const uint32_t** d_in_degree);
/* Per-page CPBT weights as an exclusive prefix, u64[pages+1]. */
int eg_graph_page_prefix(const eg_graph* g, const uint64_t** d_prefix);
/* incoming_adj_list[s] and compressed_incoming_adj_index[s] of partition p
 * (core/graph.hpp:1126-1188): n_edges records, n_cindex index entries
 * including the sentinel {n_edges, V}. */
int eg_graph_csc(const eg_graph* g, int p, int s, const void** d_list, uint64_t* n_edges, const void** d_cindex,
                 uint64_t* n_cindex);
/* outgoing_adj_list[s], outgoing_adj_index[s] (u64[V+1]) and
 * outgoing_adj_bitmap[s] (u64[(V>>6)+1]) of partition p
 * (core/graph.hpp:1282-1353).  Unlike the reference, every index entry is
 * defined: index[u] == index[u+1] for u without edges. */
int eg_graph_csr(const eg_graph* g, int p, int s, const void** d_list, uint64_t* n_edges, const uint64_t** d_index,
                 const uint64_t** d_bitmap);

/* ---- files ------------------------------------------------------------
 * Binary edge lists (README.md:40-41: packed EdgeType records in file order)
 * and vertex arrays (V * elem_bytes raw bytes), as the reference reads and
 * writes them.  eg_edge_file_records: records in the file (its size must be
 * a whole number of records).  eg_load_edge_file replaces
 * preload_graph_file's read (core/graph.hpp:1011-1055): 64 MiB chunks are
 * read into pinned host buffers and copied to d_edges (capacity_records
 * records) while the next chunk is read.  eg_dump_vertex_array /
 * eg_restore_vertex_array replace dump_vertex_array / restore_vertex_array
 * (core/graph.hpp:324-375) for the whole array.  All synchronise. */
int eg_edge_file_records(const char* path, int rec_bytes, uint64_t* h_records);
int eg_load_edge_file(const char* path, int rec_bytes, void* d_edges, uint64_t capacity_records, uint64_t* h_records,
                      void* stream);
int eg_dump_vertex_array(const void* d_array, uint64_t elem_bytes, uint32_t V, const char* path, void* stream);
int eg_restore_vertex_array(void* d_array, uint64_t elem_bytes, uint32_t V, const char* path, void* stream);

/* ---- batched CSR construction (no reference counterpart: one process of
 *      the reference builds one graph) -------------------------------------
 * n_graphs independent graphs stored back to back in d_edges (graph g: local
 * ids in [0, h_vertices[g]), h_edges[g] records) are built together; graph
 * g's result equals prepare_push_mode's CSR at P = S = 1
 * (core/graph.hpp:1248-1355): records ordered by (src, input position) and a
 * dense u64 row pointer of h_vertices[g]+1 entries.  Synchronises. */
typedef struct eg_batch eg_batch;
int eg_batch_csr_build(const void* d_edges, int rec_bytes, int n_graphs, const uint32_t* h_vertices,
                       const uint64_t* h_edges, eg_batch** out, void* stream);
/* Builds again from new records of the same sizes, reusing the batch's
 * device memory (no allocation after the first build).  Synchronises. */
int eg_batch_csr_rebuild(eg_batch* b, const void* d_edges, void* stream);
int eg_batch_csr_get(const eg_batch* b, int g, const void** d_list, uint64_t* n_edges, const uint64_t** d_rowptr,
                     uint32_t* n_vertices);
int eg_batch_destroy(eg_batch* b);
/* Allocation-free batched CSR build.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The result
 * equals eg_batch_csr_build's: d_out_list (E records, at the input offsets)
 * gets every graph's CSR list and d_out_rowptr (sum of h_vertices[g] + 1
 * u64) the row pointers, graph g's h_vertices[g] + 1 of them starting at
 * entry (vertices of graphs < g) + g.  Every temporary lives in one 32 MiB
 * device array of the library, one per GPU, held until the process ends; a
 * batch whose temporaries do not fit (about E * rec_bytes + 2 KiB per 4096
 * records and per graph) is rejected with EG_ERR_ARG and the size it needs.
 * Calls on one GPU share that array: calls that overlap on one GPU (other
 * streams or threads) race.  Synchronises `stream`. */
int eg_batch_csr_build_noalloc(const void* d_edges, int rec_bytes, int n_graphs, const uint32_t* h_vertices,
                               const uint64_t* h_edges, void* d_out_list, uint64_t* d_out_rowptr, void* stream);

/* ---- engine (Gemini process_vertices / process_edges) --------------------
 * The reference's process_edges takes four std::function lambdas
 * (core/graph.hpp:1473); a C ABI cannot carry them, so every signal/slot pair
 * the toolkits use is one entry point.  Mode policy for process_edges
 * (core/graph.hpp:1483-1498, repaired names):
 *   EG_POLICY_REFERENCE  iterative programs dense, rooted programs sparse,
 *                        undirected (CC) sparse iff active out-edges < E/20
 *   EG_POLICY_AUTO       sparse iff active out-edges < E/20 for every program
 *   EG_POLICY_PUSH / EG_POLICY_PULL force one mode.
 * Bitmaps are core/bitmap.hpp's layout: u64 words, (V>>6)+1 of them. */
enum eg_policy { EG_POLICY_REFERENCE = 0, EG_POLICY_AUTO = 1, EG_POLICY_PUSH = 2, EG_POLICY_PULL = 3 };
enum eg_precision { EG_PR_FP64 = 0, EG_PR_FP32C = 1 };
/* Layout flags for eg_pagerank, OR-ed into the precision (no reference
 * counterpart).  EG_PR_ORDERED runs the pull on a degree-ordered view:
 * vertices renumbered by descending out-degree, so the most-gathered values
 * share cache lines; one row per destination.  With P = S = 1 its rows are
 * the built CSC's, relabelled (the same terms in the same order; one gather
 * per edge to derive); otherwise the renumbered graph's rows with sources
 * ascending by new id.  EG_PR_SORTED asks for that sorted-row form on any
 * graph: costlier to derive (a row copy and a sort by destination) and
 * faster per pull, since a gather's lanes share cache lines.  EG_PR_PLAIN
 * runs the pull on the built CSC in original ids.  Defaults: FP64 plain,
 * FP32C ordered.  The view is derived on first use after a build, and again
 * when another form is asked for. */
enum eg_pr_layout { EG_PR_ORDERED = 0x10, EG_PR_PLAIN = 0x20, EG_PR_SORTED = 0x40 };
/* The four toolkit programs, for the mode rule and the message counts. */
enum eg_program { EG_PROGRAM_PAGERANK = 0, EG_PROGRAM_BFS = 1, EG_PROGRAM_CC = 2, EG_PROGRAM_SSSP = 3 };

/* The reference's byte counters (Graph::prep_data, Graph::exe_data,
 * core/graph.hpp:128-129), printed by its toolkits as "algo data / pre data".
 * The P partitions share one GPU here, so nothing is sent: these count what
 * the reference's ranks would exchange.
 * eg_prep_data: every rank's allreduce of each degree array (V x 4 B; the
 * build keeps out- and in-degrees, :716 counts the one it computes) plus,
 * for each shuffle (by source for the CSC, by destination for the CSR), the
 * records of each rank slice owned by another partition (:906-908).
 * eg_mode_sparse: whether a step of `program` with active_edges active
 * out-edges runs sparse under `policy` (the rule every step entry point
 * applies; PageRank is always dense).
 * eg_dense_messages: the messages a dense step of `program` sends between
 * partitions (:1736, MsgUnit of 4 + sizeof(message) bytes each): pull rows
 * (one per sub-partition and destination) whose destination another
 * partition owns and whose dense signal emits, given the state at the step's
 * start: PageRank always (pagerank.cpp:64-70; counted once per build), BFS
 * when the destination is not in d_visited and a source is in d_active
 * (bfs.cpp:59-67), CC when a source's label (d_values, u32) is below the
 * destination id (cc.cpp:62-72), SSSP when a source's distance (d_values,
 * f32) plus the weight is below 1e9 (sssp.cpp:62-74).  Synchronises. */
int eg_prep_data(const eg_graph* g, uint64_t* h_bytes);
int eg_mode_sparse(const eg_graph* g, int policy, int program, uint64_t active_edges, int* h_sparse);
int eg_dense_messages(eg_graph* g, int program, const uint64_t* d_active, const uint64_t* d_visited,
                      const void* d_values, uint64_t* h_count, void* stream);

/* Derives the engine's view of the built arrays (SoA source/target ids,
 * row-start bits, a CSR merged over sub-partitions).  Runs on the first
 * engine call after a build; calling it directly separates its cost
 * (the reference does the same work inside prep: tuned_chunks_dense,
 * core/graph.hpp:1191-1228).  h_ms (optional) gets its wall time. */
int eg_engine_prepare(eg_graph* g, double* h_ms, void* stream);

/* BitmapType::clear / fill / set_bit (core/bitmap.hpp:36-65). */
int eg_bitmap_clear(uint64_t* d_bm, uint32_t V, void* stream);
int eg_bitmap_fill(uint64_t* d_bm, uint32_t V, void* stream);
int eg_bitmap_set_bit(uint64_t* d_bm, uint32_t v, void* stream); /* synchronises */
/* d_dst |= d_src (bfs.cpp:80-86's visited update); number of set bits
 * (process_vertices returning 1 per active vertex); sum of out-degrees over
 * the set bits (the mode switch input, core/graph.hpp:1484-1490). */
int eg_bitmap_or(uint64_t* d_dst, const uint64_t* d_src, uint32_t V, void* stream);
int eg_bitmap_count(const uint64_t* d_bm, uint32_t V, uint64_t* h_count, void* stream); /* synchronises */
int eg_bitmap_degree_sum(eg_graph* g, const uint64_t* d_bm, uint64_t* h_sum, void* stream); /* synchronises */

/* PageRank, toolkits/pagerank.cpp:26-103.  eg_pr_init: process_vertices
 * x = 1/deg (1 if deg == 0), also written as float to d_x32 if non-NULL.
 * eg_pr_pull: process_edges dense step, d_next = sum over in-edges of x[src]
 * (EG_PR_FP64: d_x is double*; EG_PR_FP32C: d_x is float* and the per-vertex
 * sums are compensated FP32, handed over in FP64).  eg_pr_apply:
 * process_vertices next = 0.15 + 0.85*next, divided by deg unless last;
 * adds sum |next - x| * deg to *d_delta (device double, not divided by V);
 * writes a float copy to d_next32 if non-NULL.
 * eg_pagerank runs the whole loop (precision may carry an eg_pr_layout
 * flag; d_out is in original ids either way); h_deltas (optional) gets
 * iterations+1 values, delta(i) of the reference (already divided by V).
 * eg_pagerank_view_prepare derives the degree-ordered view now (h_ms gets
 * its wall time) so that its one-time cost is kept out of a timed run.
 * Reproducibility: a pull row cut by two wave ranges (or, with several
 * sub-partitions, every row of the plain layout) is completed with FP64
 * atomic adds, whose order varies between runs, so ranks may differ in the
 * last bits from run to run (about 1e-19 relative); the integer programs
 * are bit-reproducible. */
int eg_pr_init(eg_graph* g, double* d_x, float* d_x32, void* stream);
int eg_pr_pull(eg_graph* g, int precision, const void* d_x, double* d_next, void* stream);
int eg_pr_apply(eg_graph* g, const double* d_x, double* d_next, float* d_next32, int last, double* d_delta,
                void* stream);
int eg_pagerank(eg_graph* g, int iterations, int precision, double* d_out, double* h_deltas, void* stream);
int eg_pagerank_view_prepare(eg_graph* g, double* h_ms, void* stream);
/* The same for layout EG_PR_ORDERED (the graph's default form) or
 * EG_PR_SORTED (the sorted-row form). */
int eg_pagerank_view_prepare_layout(eg_graph* g, int layout, double* h_ms, void* stream);
/* Frees the degree-ordered view (no reference counterpart); the next ordered
 * run derives it again.  Synchronises the device. */
int eg_pagerank_view_release(eg_graph* g);

/* BFS, toolkits/bfs.cpp:22-112.  Parent rule (no reference counterpart: the
 * reference's CAS race leaves it open): the smallest frontier id among a
 * vertex's in-neighbours.  eg_bfs_level runs one process_edges +
 * process_vertices step: d_next = newly reached vertices, parents set,
 * d_visited |= d_next, d_depth (optional) = level+1 for them;
 * frontier_edges = sum of out-degrees of the frontier (mode switch input);
 * returns the new vertex count and their out-edge sum.  eg_bfs runs from
 * root; h_edges gets the number of out-edges of reached vertices.  Its
 * levels run on the device as a captured step graph (direction chosen per
 * level on the device, the host checking a pinned mailbox once per 4
 * levels); it returns when the run is done.  Dense
 * (pull) steps and the batched programs read the traversal view: in-edges
 * by destination with sources ascending, derived on first use after a
 * build; eg_traversal_view_prepare derives it now (h_ms gets its wall time)
 * so that its one-time cost is kept out of a timed run. */
int eg_bfs_level(eg_graph* g, int policy, uint32_t level, const uint64_t* d_frontier, uint64_t* d_next,
                 uint64_t* d_visited, uint32_t* d_parent, uint32_t* d_depth, uint64_t frontier_edges,
                 uint64_t* h_new_vertices, uint64_t* h_new_edges, int* h_was_sparse, void* stream);
int eg_bfs(eg_graph* g, uint32_t root, int policy, uint32_t* d_parent, uint32_t* d_depth, uint32_t* h_levels,
           uint64_t* h_edges, void* stream);
int eg_traversal_view_prepare(eg_graph* g, double* h_ms, void* stream);

/* CC, toolkits/cc.cpp:22-112 (graph built with symmetrize=1): label =
 * smallest id of the component.  One iteration or the whole loop.
 * eg_cc_iteration / eg_sssp_iteration / eg_bfs_level take the active set's
 * out-edge sum (the mode switch input, core/graph.hpp:1484-1490) and report
 * in *h_was_sparse (optional) the mode the step ran in. */
int eg_cc_iteration(eg_graph* g, int policy, const uint64_t* d_active, uint64_t* d_next, uint32_t* d_label,
                    uint64_t active_edges, uint64_t* h_new_active, uint64_t* h_new_edges, int* h_was_sparse,
                    void* stream);
int eg_cc(eg_graph* g, int policy, uint32_t* d_label, uint32_t* h_iterations, void* stream);

/* SSSP, toolkits/sssp.cpp:24-108 (12-byte records): float distances, 1e9
 * for unreached, relax = dist[src] + w in one round-to-nearest fp32 add.
 * eg_sssp_iteration runs one process_edges step from the host; eg_sssp runs
 * the whole program on the device as a captured step graph (mode chosen
 * per step on the device; the auto policy goes dense at E/80 active
 * out-edges) and returns when it is done.  Distances do not depend on the
 * policy or the step order (the least fixpoint of the relaxation). */
int eg_sssp_iteration(eg_graph* g, int policy, const uint64_t* d_active, uint64_t* d_next, float* d_dist,
                      uint64_t active_edges, uint64_t* h_new_active, uint64_t* h_new_edges, int* h_was_sparse,
                      void* stream);
int eg_sssp(eg_graph* g, uint32_t root, int policy, float* d_dist, uint32_t* h_iterations, void* stream);

/* Personalized PageRank with per-edge weights and its adjoint (no reference
 * counterpart).  Edge slots are the engine's CSC order (the sub-partitions'
 * incoming lists back to back); eg_ppr_slots writes the source, destination
 * and record weight (1 for 8-byte records) of every slot (each pointer may be
 * NULL).  d_w holds one FP64 weight per slot.
 * eg_ppr_forward: x_{k+1} = d * A x_k + (1-d) * p from x_0 = p, `iterations`
 *   times, A[v][u] = sum of the weights of the slots u -> v; FP64.
 * eg_ppr_adjoint: for a loss L with dL/dx = d_c at the forward result d_x:
 *   lambda = c + d * A^T lambda (same iteration count, from lambda = c),
 *   d_grad_p = (1-d) * lambda, d_grad_w[e] = d * lambda[dst_e] * x[src_e]
 *   (either output may be NULL).  The first call after a build derives a
 *   source-grouped copy of the slots.  All three synchronise. */
int eg_ppr_slots(eg_graph* g, uint32_t* d_src, uint32_t* d_dst, float* d_w, void* stream);
int eg_ppr_forward(eg_graph* g, const double* d_w, const double* d_p, double damping, int iterations, double* d_x,
                   void* stream);
int eg_ppr_adjoint(eg_graph* g, const double* d_w, const double* d_x, const double* d_c, double damping, int iterations,
                   double* d_grad_p, double* d_grad_w, void* stream);

/* Batched BFS and SSSP from n_roots (1..64) roots at once (no reference
 * counterpart: bfs.cpp / sssp.cpp run one root per process); each level or
 * relaxation round serves all roots, directions are decided on the device.
 * Root r is lane r of a wavefront; results are vertex-major, 64 per vertex:
 * d_parent[v*64 + r], d_depth[v*64 + r] (u32, V*64 each), d_dist[v*64 + r]
 * (f32, V*64).  Each root's column equals eg_bfs / eg_sssp from that root bit
 * for bit (same min-id parent rule, same policy meanings); lanes >= n_roots
 * hold the unreached values (parent V, depth 0xffffffff, dist 1e9).
 * h_levels / h_edges (optional, n_roots each) get what eg_bfs reports per
 * root; h_steps (optional) the number of levels / rounds run.
 * The first call after a build derives a source-sorted copy of the CSC
 * (the pull side of the batched programs). */
/* Sizes and auto-policy thresholds of the device-driven traversals, per
 * graph (no reference counterpart: the reference's schedule is its fixed
 * thread chunks, and its mode switch a fixed edges/20, core/graph.hpp:1491).
 * Under EG_POLICY_AUTO eg_bfs goes dense once the frontier's out-edges reach
 * E / bfs_dense_div (eg_sssp: sssp_dense_div; the batched programs:
 * bfs_multi_dense_div, sssp_multi_dense_div); the single-root push cuts
 * vertices of more than hub_degree out-edges into chunks of hub_chunk; the
 * batched push's work items hold multi_item_edges out-edges.  No value
 * changes a result.  Defaults: 20, 80, 20, 2, 4096, 1024, 4096.
 * eg_autotune_engine times the programs from n_roots sample roots (the
 * max-out-degree vertex and seeded vertices with out-edges) over candidate
 * values, keeps the fastest for g and returns them (h_ms, optional: the best
 * times of single BFS, single BFS+SSSP hub split, single SSSP, batched BFS,
 * batched SSSP, in ms).  Synchronises. */
typedef struct eg_engine_tune {
  uint32_t bfs_dense_div, sssp_dense_div, bfs_multi_dense_div, sssp_multi_dense_div;
  uint32_t hub_degree, hub_chunk, multi_item_edges;
} eg_engine_tune;
int eg_get_engine_tune(const eg_graph* g, eg_engine_tune* out);
int eg_set_engine_tune(eg_graph* g, const eg_engine_tune* tune);
int eg_autotune_engine(eg_graph* g, int n_roots, uint64_t seed, eg_engine_tune* h_out, double* h_ms, void* stream);

int eg_bfs_multi(eg_graph* g, const uint32_t* h_roots, int n_roots, int policy, uint32_t* d_parent, uint32_t* d_depth,
                 uint32_t* h_levels, uint64_t* h_edges, uint32_t* h_steps, void* stream);
int eg_sssp_multi(eg_graph* g, const uint32_t* h_roots, int n_roots, int policy, float* d_dist, uint32_t* h_steps,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ENDGRAPH_H */
