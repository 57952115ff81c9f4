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
 * the degree arrays and the partition offsets.  Synchronises `stream`. */
int eg_graph_build(eg_graph* g, const void* d_edges, uint64_t E0, int symmetrize, void* stream);

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

#ifdef __cplusplus
}
#endif

#endif /* ENDGRAPH_H */
