// Synthetic edge lists written in the reference's binary record layout
// (core/type.hpp:26-40; README.md:40-41 "<src, dst, edge data> tuples").
// Counter-based RNG: every record is a pure function of (seed, index), so a
// list is identical whatever the launch shape, and any slice can be
// regenerated on the host for checking.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace eg {
namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {  // SplitMix64 output function
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// 64 random bits number k of record e.
__device__ inline uint64_t draw(uint64_t key, uint64_t e, uint32_t k) { return mix64(key + e * 64 + k); }

__device__ inline float weight_of(uint64_t key, uint64_t e, float lo, float hi) {
  const float u = float(draw(key, e, 63) >> 40) * (1.0f / 16777216.0f);  // exact, in [0,1)
  return __fadd_rn(lo, __fmul_rn(hi - lo, u));
}

template <int RB>
__device__ inline void store(void* out, uint64_t e, uint32_t src, uint32_t dst, float w) {
  uint32_t* p = static_cast<uint32_t*>(out) + e * (RB / 4);
  p[0] = src;
  p[1] = dst;
  if (RB == 12) p[2] = __float_as_uint(w);
}

// R-MAT: level l (0 = most significant bit) takes 32 random bits r and picks
// quadrant a (r < ta), b (< tab: dst bit set), c (< tabc: src bit set) or d.
template <int RB>
__global__ void k_rmat(void* out, uint64_t E, uint32_t scale, uint64_t ta, uint64_t tab, uint64_t tabc,
                       uint64_t key, float wlo, float whi) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t src = 0, dst = 0;
    uint64_t bits = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) bits = draw(key, e, l >> 1);
      const uint64_t r = (l & 1) ? (bits >> 32) : (bits & 0xffffffffull);
      const uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        dst |= bit;
      } else if (r < tabc) {
        src |= bit;
      } else {
        src |= bit;
        dst |= bit;
      }
    }
    store<RB>(out, e, src, dst, RB == 12 ? weight_of(key, e, wlo, whi) : 0.f);
  }
}

// Seeded bijection of [0, n): 4-round Feistel network on 2*half bits with
// cycle walking (terminates: the walk follows x's own cycle, which contains x).
__device__ inline uint32_t permute(uint32_t x, uint32_t n, int half, uint64_t key) {
  const uint32_t mask = (1u << half) - 1;
  do {
    uint32_t l = x >> half, r = x & mask;
    for (int round = 0; round < 4; ++round) {
      const uint32_t f = uint32_t(mix64(key + uint64_t(round) * 0x632BE59BD9B4E019ull + r)) & mask;
      const uint32_t t = l ^ f;
      l = r;
      r = t;
    }
    x = (l << half) | r;
  } while (x >= n);
  return x;
}

// Inverse CDF of the continuous law with density ~ (x + c)^-beta on [0, n).
__device__ inline uint32_t powerlaw_rank(double u, double n, double beta, double c) {
  double x;
  if (fabs(1.0 - beta) < 1e-9) {
    x = c * pow((n + c) / c, u) - c;
  } else {
    const double g = 1.0 - beta;
    const double lo = pow(c, g), hi = pow(n + c, g);
    x = pow(lo + u * (hi - lo), 1.0 / g) - c;
  }
  if (!(x >= 0.0)) x = 0.0;
  if (x >= n) x = n - 1;
  return uint32_t(x);
}

template <int RB>
__global__ void k_powerlaw(void* out, uint64_t E, uint32_t V, int half, double bo, double co, double bi, double ci,
                           uint64_t key) {
  const uint64_t kp_out = mix64(key ^ 0x0A0A0A0A0A0A0A0Aull), kp_in = mix64(key ^ 0x5050505050505050ull);
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    const double u1 = double(draw(key, e, 0) >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = double(draw(key, e, 1) >> 11) * (1.0 / 9007199254740992.0);
    const uint32_t rs = powerlaw_rank(u1, double(V), bo, co);
    const uint32_t rd = powerlaw_rank(u2, double(V), bi, ci);
    store<RB>(out, e, permute(rs, V, half, kp_out), permute(rd, V, half, kp_in),
              RB == 12 ? weight_of(key, e, 0.f, 1.f) : 0.f);
  }
}

// Erdos-Renyi batch: record e of graph g (ebase[g] <= e < ebase[g+1]) joins
// two ids drawn uniformly from [0, V_g) (32-bit draw scaled by V_g).
template <int RB>
__global__ void k_er_batch(void* out, uint64_t E, const uint64_t* __restrict__ ebase,
                           const uint32_t* __restrict__ nv, int n, uint64_t key) {
  for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += uint64_t(gridDim.x) * blockDim.x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (ebase[mid] <= e)
        lo = mid;
      else
        hi = mid;
    }
    const uint64_t V = nv[lo];
    const uint64_t r = draw(key, e, 0);
    const uint32_t src = uint32_t(((r & 0xffffffffull) * V) >> 32);
    const uint32_t dst = uint32_t(((r >> 32) * V) >> 32);
    store<RB>(out, e, src, dst, RB == 12 ? weight_of(key, e, 0.f, 1.f) : 0.f);
  }
}

uint64_t threshold(double p) {
  double t = std::floor(p * 4294967296.0);
  if (t < 0) t = 0;
  if (t > 4294967296.0) t = 4294967296.0;
  return uint64_t(t);
}

// Probability of rank 0 under (x+c)^-beta on [0,n).
double top_mass(double n, double beta, double c) {
  if (std::fabs(1.0 - beta) < 1e-9) return std::log((1 + c) / c) / std::log((n + c) / c);
  const double g = 1.0 - beta;
  return (std::pow(1 + c, g) - std::pow(c, g)) / (std::pow(n + c, g) - std::pow(c, g));
}

// Offset c giving an expected top degree of max_deg (top_mass decreases in c).
double solve_offset(double n, double beta, double E, double max_deg) {
  double lo = std::log(1e-9), hi = std::log(1e15);
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    if (E * top_mass(n, beta, std::exp(mid)) > max_deg)
      lo = mid;
    else
      hi = mid;
  }
  return std::exp(0.5 * (lo + hi));
}

unsigned gen_grid(uint64_t E) {
  uint64_t b = ceil_div(E, 256);
  return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

}  // namespace
}  // namespace eg

using namespace eg;

extern "C" {

int eg_gen_rmat(void* d_edges, uint64_t E, uint32_t scale, double a, double b, double c, uint64_t seed, int rec_bytes,
                float w_lo, float w_hi, void* stream) {
  if (!d_edges && E) return fail(EG_ERR_ARG, "eg_gen_rmat: d_edges is NULL");
  if (scale < 1 || scale > 32) return fail(EG_ERR_ARG, "eg_gen_rmat: scale must be in [1, 32]");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return fail(EG_ERR_ARG, "eg_gen_rmat: bad quadrant probabilities");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  if (!E) return EG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t key = mix64(seed);
  const uint64_t ta = threshold(a), tab = threshold(a + b), tabc = threshold(a + b + c);
  if (rec_bytes == 8)
    hipLaunchKernelGGL(k_rmat<8>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, scale, ta, tab, tabc, key, w_lo, w_hi);
  else
    hipLaunchKernelGGL(k_rmat<12>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, scale, ta, tab, tabc, key, w_lo,
                       w_hi);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

int eg_gen_powerlaw(void* d_edges, uint64_t E, uint32_t V, double alpha_out, double alpha_in, double max_out,
                    double max_in, uint64_t seed, int rec_bytes, void* stream) {
  if (!d_edges && E) return fail(EG_ERR_ARG, "eg_gen_powerlaw: d_edges is NULL");
  if (V < 2 || alpha_out <= 1.0 || alpha_in <= 1.0 || max_out < 1 || max_in < 1)
    return fail(EG_ERR_ARG, "eg_gen_powerlaw: need V >= 2, alpha > 1, max degrees >= 1");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  if (!E) return EG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double bo = 1.0 / (alpha_out - 1.0), bi = 1.0 / (alpha_in - 1.0);
  const double co = solve_offset(V, bo, double(E), max_out), ci = solve_offset(V, bi, double(E), max_in);
  int bits = counting_bits(V - 1);
  if (bits < 2) bits = 2;
  const int half = (bits + 1) / 2;
  const uint64_t key = mix64(seed);
  if (rec_bytes == 8)
    hipLaunchKernelGGL(k_powerlaw<8>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, V, half, bo, co, bi, ci, key);
  else
    hipLaunchKernelGGL(k_powerlaw<12>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, V, half, bo, co, bi, ci, key);
  EG_HIP(hipGetLastError());
  return EG_OK;
}

int eg_gen_er_batch(void* d_edges, int n_graphs, const uint32_t* h_vertices, const uint64_t* h_edges, uint64_t seed,
                    int rec_bytes, void* stream) {
  if (n_graphs < 1 || !h_vertices || !h_edges) return fail(EG_ERR_ARG, "eg_gen_er_batch: bad graph list");
  if (rec_bytes != 8 && rec_bytes != 12) return fail(EG_ERR_ARG, "rec_bytes must be 8 or 12");
  std::vector<uint64_t> eb(n_graphs + 1, 0);
  for (int g = 0; g < n_graphs; ++g) {
    if (h_vertices[g] == 0 && h_edges[g] > 0) return fail(EG_ERR_ARG, "graph %d has edges but no vertices", g);
    eb[g + 1] = eb[g] + h_edges[g];
  }
  const uint64_t E = eb[n_graphs];
  if (!E) return EG_OK;
  if (!d_edges) return fail(EG_ERR_ARG, "eg_gen_er_batch: d_edges is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevBuf d_eb, d_nv;
  EG_TRY(d_eb.resize((n_graphs + 1) * 8));
  EG_TRY(d_nv.resize(n_graphs * 4));
  EG_HIP(hipMemcpyAsync(d_eb.ptr, eb.data(), (n_graphs + 1) * 8, hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(d_nv.ptr, h_vertices, n_graphs * 4, hipMemcpyHostToDevice, s));
  const uint64_t key = mix64(seed);
  if (rec_bytes == 8)
    hipLaunchKernelGGL(k_er_batch<8>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, d_eb.as<uint64_t>(),
                       d_nv.as<uint32_t>(), n_graphs, key);
  else
    hipLaunchKernelGGL(k_er_batch<12>, dim3(gen_grid(E)), dim3(256), 0, s, d_edges, E, d_eb.as<uint64_t>(),
                       d_nv.as<uint32_t>(), n_graphs, key);
  EG_HIP(hipGetLastError());
  EG_HIP(hipStreamSynchronize(s));  // the size tables are freed on return
  return EG_OK;
}

}  // extern "C"
