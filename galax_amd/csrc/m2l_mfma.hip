// m2l_mfma.hip — M2L of the reference variant as batched dense contractions
// on the matrix cores.
//
// galax-core's M2L (galax-core/src/lib.rs:1199-1259) uses, for every pair of a
// target node t and source node s, the derivative table of the first pair with
// the same key (dx as i64, dy as i64). The operator of a pair therefore
// depends on its key K only:
//     L_t[I,J] += sum_{i,j} (-1)^{i+j} / (I! J!) * G_K[i+I, j+J] * M_s[i,j]
// i.e. L_t += T_K M_s with T_K an S x S matrix. Two consequences are used:
//  * the sources of one target that share a key are summed first (a group):
//    L_t += T_K (sum_s M_s);
//  * 16 consecutive targets form a tile whose columns are those targets; the
//    tile steps through the distinct keys of its groups, and at each step one
//    16-column product T_K B runs on the matrix cores, B's column c holding
//    the group sum of target c for key K (zero when c has no such group).
// Results accumulate in registers: no per-pair result buffer and no gather
// pass. A work item is at most MX_ITEM_STEPS steps of one tile, run by one
// block whose four waves take the steps round-robin and are summed in wave
// order through LDS; a tile with several items writes partials that a combine
// pass sums in item order. Results therefore do not depend on scheduling.
// Blocks are mapped so that each XCD (and its L2) takes runs of XCD_RUN
// consecutive items, i.e. of neighbouring tiles that share sources and keys.
//
// The adjoint (VJP) uses the same machinery with rows = sources and the
// transposed FP64 operator: mbar_s += T_K^T (sum of lbar_t over its group).
//
// Forms:
//  * FP64: T_K in FP64 on v_mfma_f64_16x16x4_f64, accumulated across steps in
//    the matrix accumulators.
//  * MIXED_FP32: a scaled FP32 operator per key
//        T^_K[IJ][ij] = T_K[IJ][ij] * sig^(i+j) * sig^(I+J+1),
//    sig the representative's softened distance; columns M^_ij = M_ij *
//    sig^-(i+j); each step's product (v_mfma_f32_16x16x4_f32) is unscaled by
//    sig^-(I+J+1) and added to an FP64 accumulator.
//  * MIXED_BF16X3: as MIXED_FP32 with both operands split into three bf16
//    terms and six v_mfma_f32_16x16x32_bf16 products per k-block.
#include <algorithm>
#include <cmath>

#include <type_traits>

#include "common.h"

namespace gx {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

enum { MX_F64 = 0, MX_F32 = 1, MX_BF16X3 = 2 };

static inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

static int bits_of(uint64_t n) {  // bits to hold values 0 .. n-1 (at least 1)
  int w = 1;
  while (w < 63 && (1ull << w) < n) ++w;
  return w;
}

__host__ __device__ constexpr int row_blocks(int p) { return (num_moments(p) + 15) / 16; }

template <int P>
__host__ __device__ __forceinline__ int order_of(int k) {
  int n = 0;
#pragma unroll
  for (int q = 0; q <= P; ++q)
    if (k >= (q + 1) * (q + 2) / 2) n = q + 1;
  return n < P ? n : P;
}

template <int P>
__device__ __forceinline__ void coef_ij2(int k, int& i, int& j) {
  int n = order_of<P>(k);
  j = k - n * (n + 1) / 2;
  i = n - j;
}

// Contraction index carried by slot e (0 <= e < KL) of lane quarter kq:
//     FP64 source rows (FP64 form):   k = 8 (e / 2) + 2 kq + (e % 2),
//     FP32 source rows (mixed forms): k = 16 (e / 4) + 4 kq + (e % 4).
// The 16x16x4 forms contract the KL = K/4 slots one per instruction; the
// 16x16x32 bf16 form takes slots 8c .. 8c + 7 of each lane in instruction c.
// A 16-byte load per lane of a source row's consecutive slots (2 doubles or 4
// floats) then reads 64 contiguous bytes of each row per instruction.
template <int P, int MODE>
struct Lanes {
  static constexpr int S = num_moments(P), RB = row_blocks(P), K = 16 * RB;
  static constexpr int KC = (K + 31) / 32;
  static constexpr int KL = MODE == MX_BF16X3 ? KC * 8 : K / 4;
  static __host__ __device__ __forceinline__ int k_of(int e, int kq) {
    return MODE == MX_F64 ? 8 * (e >> 1) + 2 * kq + (e & 1) : 16 * (e >> 2) + 4 * kq + (e & 3);
  }
  // operator values per key, in fragment order (below)
  static constexpr int PER_KEY = RB * KL * 64;
};

// ------------------------------------------------------------- setup ----

constexpr uint32_t MX_ITEM_STEPS = 64;  // steps per work item (32: +2% on the FP32 pass, 128: +5%)
constexpr uint32_t MX_WAVE_RUN = 256;   // longest run pre-summed by one wave; longer runs (hubs) go in chunks
constexpr uint32_t MX_HUB_CHUNK = 128;  // rows per chunk of a hub run

// row of each live pair: its target (CSR row) or its source
__global__ void k_mx_pair_rows(const uint32_t* __restrict__ off, uint32_t nt, uint32_t* __restrict__ tgt) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  for (uint32_t e = off[t]; e < off[t + 1]; ++e) tgt[e] = t;
}

__global__ void k_mx_pair_keys(const uint32_t* __restrict__ row_of_pair, const uint32_t* __restrict__ key_of_pair,
                               size_t ne, uint32_t n_rows, int kb, unsigned long long* __restrict__ ka,
                               uint32_t* __restrict__ va, uint32_t* __restrict__ err) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const uint32_t r = row_of_pair[e];
  if (r >= n_rows) atomicOr(err, 1u);
  ka[e] = ((unsigned long long)min(r, n_rows - 1) << kb) | key_of_pair[e];
  va[e] = (uint32_t)e;
}

__global__ void k_mx_run_flags(const unsigned long long* __restrict__ k, size_t n, int shift,
                               uint32_t* __restrict__ f) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || (k[i] >> shift) != (k[i - 1] >> shift)) ? 1u : 0u;
}

// run starts of equal (row, key) in the sorted pairs, as their own index
// (a max-scan then gives each pair the start of its run)
__global__ void k_mx_run_starts(const unsigned long long* __restrict__ k, size_t n, uint32_t* __restrict__ f) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || k[i] != k[i - 1]) ? (uint32_t)i : 0u;
}

// a group is one (row, key) run: flag its first pair
__global__ void k_mx_chunk_flags(const uint32_t* __restrict__ run_start, size_t n, uint32_t* __restrict__ f) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == run_start[i]) ? 1u : 0u;
}

// pairs sorted by (row, key, list position) -> the other end of each pair in
// that order, and groups (one per (row, key) run) with their tile composite
// (tile, key, rank 0, column) for the second sort.
// gid: exclusive scan of the group starts.
__global__ void k_mx_groups(const unsigned long long* __restrict__ ka, const uint32_t* __restrict__ va,
                            const uint32_t* __restrict__ run_start, const uint32_t* __restrict__ gid, size_t ne,
                            const uint32_t* __restrict__ other_of_pair, int kb, int rkb,
                            uint32_t* __restrict__ src_sorted, uint32_t* __restrict__ gfirst,
                            unsigned long long* __restrict__ g2k, uint32_t* __restrict__ g2v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  src_sorted[i] = other_of_pair[va[i]];
  if ((uint32_t)i == run_start[i]) {
    const uint32_t g = gid[i];
    const unsigned long long t = ka[i] >> kb, key = ka[i] & ((1ull << kb) - 1), rank = 0;
    gfirst[g] = (uint32_t)i;
    g2k[g] = ((t >> 4) << (kb + rkb + 4)) | (key << (rkb + 4)) | (rank << 4) | (t & 15);
    g2v[g] = g;
  }
}

// sid: inclusive scan of the step-start flags over the sorted groups
__global__ void k_mx_steps(const unsigned long long* __restrict__ g2k, const uint32_t* __restrict__ g2v,
                           const uint32_t* __restrict__ sid, size_t ng, int kb, int rkb,
                           uint32_t* __restrict__ step_key, uint32_t* __restrict__ step_grp,
                           uint32_t* __restrict__ tile_cnt) {
  size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ng) return;
  const unsigned long long c = g2k[j];
  const uint32_t st = sid[j] - 1, col = (uint32_t)(c & 15), g = g2v[j];
  step_grp[(size_t)st * 16 + col] = g;
  if (j == 0 || (c >> 4) != (g2k[j - 1] >> 4)) {
    step_key[st] = (uint32_t)((c >> (rkb + 4)) & ((1ull << kb) - 1));
    atomicAdd(&tile_cnt[c >> (kb + rkb + 4)], 1u);
  }
}

__global__ void k_mx_set_u32(uint32_t* __restrict__ p, uint32_t v) { *p = v; }

// groups of more than G pairs (long runs): index and source range
__global__ void k_mx_long_flags(const uint32_t* __restrict__ gfirst, size_t ng, uint32_t G, uint32_t* __restrict__ f) {
  size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ng) f[g] = gfirst[g + 1] - gfirst[g] > G ? 1u : 0u;
}

__global__ void k_mx_long_table(const uint32_t* __restrict__ gfirst, const uint32_t* __restrict__ lrid, size_t ng,
                                uint32_t G, uint32_t* __restrict__ lr_first, uint32_t* __restrict__ lr_len) {
  size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t c = gfirst[g + 1] - gfirst[g];
  if (c > G) {
    lr_first[lrid[g]] = gfirst[g];
    lr_len[lrid[g]] = c;
  }
}

// sort key of each long run: 32 - its length's bit length (longest class first), 6 bits
__global__ void k_mx_len_class(const uint32_t* __restrict__ lr_len, uint32_t n, unsigned long long* __restrict__ key,
                               uint32_t* __restrict__ val) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t len = lr_len[r];
  key[r] = (unsigned long long)(len ? __clz(len) : 32);
  val[r] = r;
}

__global__ void k_mx_hub_flags(const uint32_t* __restrict__ lr_len, size_t n, uint32_t* __restrict__ f) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) f[r] = lr_len[r] > MX_WAVE_RUN ? 1u : 0u;
}

__global__ void k_mx_hub_list(const uint32_t* __restrict__ lr_len, const uint32_t* __restrict__ pos, size_t n,
                              uint32_t* __restrict__ hubs) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && lr_len[r] > MX_WAVE_RUN) hubs[pos[r]] = (uint32_t)r;
}

// source ids of every step, [step][q][column] (q < G), NONE past a group's
// end; a long run's only entry is its pre-summed row, id n_in + its index
__global__ void k_mx_step_ids(const uint32_t* __restrict__ step_grp, const uint32_t* __restrict__ gfirst,
                              const uint32_t* __restrict__ lrid, const uint32_t* __restrict__ src, size_t ns,
                              uint32_t n_in, uint32_t G, const uint32_t* __restrict__ step_key,
                              uint32_t* __restrict__ ids, uint32_t* __restrict__ lr_key) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns * 16) return;
  const size_t st = i / 16, col = i % 16;
  const uint32_t g = step_grp[i];
  uint32_t f = 0, c = 0;
  if (g != NONE) {
    f = gfirst[g];
    c = gfirst[g + 1] - f;
  }
  if (c > G) lr_key[lrid[g]] = step_key[st];  // each long run is one (step, column)
  for (uint32_t q = 0; q < G; ++q) {
    uint32_t v = NONE;
    if (c > G)
      v = q == 0 ? n_in + lrid[g] : NONE;
    else if (q < c)
      v = src[f + q];
    ids[(st * G + q) * 16 + col] = v;
  }
}

// Mixed forms. Exponent of each node's FP32 multipole row: the binary
// exponent of its half width (rows hold M_ij 2^(-e (i+j)), i.e. the moments
// in units of the node's size, far from FP32's range limits at any depth).
__global__ void k_mx_node_exp(const double* __restrict__ hw, uint32_t nn, int8_t* __restrict__ nexp) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double w = hw[i];
  int e = (w > 0.0 && isfinite(w)) ? ilogb(w) : 0;
  nexp[i] = (int8_t)min(max(e, -100), 100);
}

// Per (step, column): the shift e_s - a of each of its G source rows from
// the row's own scale to the key's (a the key's exponent); 0 for a
// pre-summed row, which is written at the key's scale. Packed as int8; a
// shift outside +-127 (a node 2^127 times smaller or larger than its
// distance) sets *err.
__global__ void k_mx_step_shifts(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ step_key,
                                 const int* __restrict__ sig_exp, const int8_t* __restrict__ nexp, size_t ns,
                                 uint32_t n_in, uint32_t G, uint32_t* __restrict__ sh, uint32_t* __restrict__ err) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns * 16) return;
  const size_t st = i / 16, col = i % 16;
  const int a = sig_exp[step_key[st]];
  uint32_t packed = 0;
  for (uint32_t q = 0; q < G; ++q) {
    const uint32_t id = ids[(st * G + q) * 16 + col];
    int t = (id != NONE && id < n_in) ? (int)nexp[id] - a : 0;
    if (t < -127 || t > 127) {
      atomicOr(err, 1u);
      t = 0;
    }
    packed |= (uint32_t)(uint8_t)(int8_t)t << (8 * q);
  }
  sh[i] = packed;
}

// Lanes per run of the pre-sums: a lane per coefficient pair when S is even
// (each row then starts on a 16-byte boundary), else a lane per coefficient.
inline __host__ __device__ constexpr int sums_lanes(int s) {
  return s % 2 == 0 ? (s / 2 <= 16 ? 16 : s / 2 <= 32 ? 32 : 64) : (s <= 32 ? 32 : 64);
}

// Pre-sums of the long runs, every evaluation, in a fixed order. Runs of at
// most MX_WAVE_RUN pairs: sums_lanes(S) lanes per run, sources in list order.
// Longer runs (hubs: sources listed by very many rows, as in the adjoint) are
// summed in chunks by k_m2l_hub_chunks and k_m2l_hub_sum.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_run_sums_w(const uint32_t* __restrict__ lr_order,
                                                        const uint32_t* __restrict__ lr_first,
                                                        const uint32_t* __restrict__ lr_len, uint32_t nlr,
                                                        const uint32_t* __restrict__ src,
                                                        const double* __restrict__ in, double* __restrict__ virt) {
  // LW lanes per run (sums_lanes), RPW runs per wave
  constexpr int S = num_moments(P);
  constexpr int LW = sums_lanes(S), RPW = 64 / LW;
  const int lane = threadIdx.x & 63;
  const uint32_t q = (xcd_runs(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6)) * RPW + lane / LW;
  const int kl = lane % LW;
  if (q >= nlr) return;
  const uint32_t r = lr_order[q], f = lr_first[r], len = lr_len[r];
  if (len > MX_WAVE_RUN) return;
  if constexpr (S % 2 == 0) {
    // a lane per coefficient pair: each coefficient is still summed alone,
    // in list order
    for (int k2 = kl; k2 < S / 2; k2 += LW) {
      double ax = 0.0, ay = 0.0;
#pragma unroll 4
      for (uint32_t i = 0; i < len; ++i) {
        const double2 v = *reinterpret_cast<const double2*>(in + (size_t)src[f + i] * S + 2 * k2);
        ax += isfinite(v.x) ? v.x : 0.0;
        ay += isfinite(v.y) ? v.y : 0.0;
      }
      *reinterpret_cast<double2*>(virt + (size_t)r * S + 2 * k2) = make_double2(ax, ay);
    }
    return;
  }
  for (int k = kl; k < S; k += LW) {
    double acc = 0.0;
#pragma unroll 4
    for (uint32_t i = 0; i < len; ++i) {
      const double v = in[(size_t)src[f + i] * S + k];
      acc += isfinite(v) ? v : 0.0;
    }
    virt[(size_t)r * S + k] = acc;
  }
}

// Mixed forms: the same pre-sums from the FP32 node rows (K floats, zero past
// S, each at its node's scale 2^(-e ord)): K/4 lanes per run, a lane per four
// coefficients (one 16-byte load per row), each row rescaled to the key's
// scale 2^(-a ord) and summed in FP64 in list order, rounded to FP32 once.
// Runs of any length, in tab order: (first pair, length, pre-summed row);
// psh holds each pair's shift e - a, so a row's load waits only on its id.
template <int P>
__host__ __device__ constexpr int sums32_lanes() {
  return 4 * row_blocks(P);
}

template <int P>
__global__ void __launch_bounds__(256) k_m2l_run_sums32(const uint4* __restrict__ tab, uint32_t nlr,
                                                        const uint32_t* __restrict__ src,
                                                        const int8_t* __restrict__ psh,
                                                        const float* __restrict__ in32, float* __restrict__ virt32) {
  constexpr int K = 16 * row_blocks(P), LW = sums32_lanes<P>(), RPW = 64 / LW;
  static_assert(LW <= 64, "a run's row must fit one wave");
  const int lane = threadIdx.x & 63;
  const uint32_t q = (xcd_runs(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6)) * RPW + lane / LW;
  const int kl = lane % LW;
  if (lane >= RPW * LW || q >= nlr) return;
  const uint4 rt = tab[q];
  const uint32_t f = rt.x, len = rt.y, r = rt.z;
  int o[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) o[j] = order_of<P>(4 * kl + j);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (uint32_t i = 0; i < len; ++i) {
    const uint32_t sn = src[f + i];
    const int t = psh[f + i];
    const float4 v = *reinterpret_cast<const float4*>(in32 + (size_t)sn * K + 4 * kl);
    acc[0] += ldexp((double)v.x, t * o[0]);
    acc[1] += ldexp((double)v.y, t * o[1]);
    acc[2] += ldexp((double)v.z, t * o[2]);
    acc[3] += ldexp((double)v.w, t * o[3]);
  }
  *reinterpret_cast<float4*>(virt32 + (size_t)r * K + 4 * kl) =
      make_float4((float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]);
}

// The pre-sum table of the mixed forms: per run q of the pre-sum order its
// (first pair, length, row), and per pair of the run its shift e_s - a (the
// node's row exponent less the key's); a shift outside +-127 sets *err.
__global__ void k_mx_run_tab32(const uint32_t* __restrict__ lr_order, const uint32_t* __restrict__ lr_first,
                               const uint32_t* __restrict__ lr_len, const uint32_t* __restrict__ lr_key, uint32_t nlr,
                               const uint32_t* __restrict__ src, const int* __restrict__ sig_exp,
                               const int8_t* __restrict__ nexp, uint4* __restrict__ tab, int8_t* __restrict__ psh,
                               uint32_t* __restrict__ err) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nlr) return;
  const uint32_t r = lr_order[q], f = lr_first[r], len = lr_len[r];
  tab[q] = make_uint4(f, len, r, 0u);
  const int a = sig_exp[lr_key[r]];
  for (uint32_t i = 0; i < len; ++i) {
    int t = (int)nexp[src[f + i]] - a;
    if (t < -127 || t > 127) {
      atomicOr(err, 1u);
      t = 0;
    }
    psh[f + i] = (int8_t)t;
  }
}

// Hub runs in chunks of MX_HUB_CHUNK rows: sums_lanes(S) lanes per chunk sum
// its rows in list order into a partial row; then one thread per (hub,
// coefficient) adds the hub's partials in chunk order (a fixed order).
template <int P>
__global__ void __launch_bounds__(256) k_m2l_hub_chunks(const uint32_t* __restrict__ ch_first,
                                                        const uint32_t* __restrict__ ch_len, uint32_t nch,
                                                        const uint32_t* __restrict__ src,
                                                        const double* __restrict__ in, double* __restrict__ part) {
  constexpr int S = num_moments(P);
  constexpr int LW = sums_lanes(S), RPW = 64 / LW;
  const int lane = threadIdx.x & 63;
  const uint32_t q = (blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW + lane / LW;
  const int kl = lane % LW;
  if (q >= nch) return;
  const uint32_t f = ch_first[q], len = ch_len[q];
  if constexpr (S % 2 == 0) {
    for (int k2 = kl; k2 < S / 2; k2 += LW) {
      double ax = 0.0, ay = 0.0;
#pragma unroll 4
      for (uint32_t i = 0; i < len; ++i) {
        const double2 v = *reinterpret_cast<const double2*>(in + (size_t)src[f + i] * S + 2 * k2);
        ax += isfinite(v.x) ? v.x : 0.0;
        ay += isfinite(v.y) ? v.y : 0.0;
      }
      *reinterpret_cast<double2*>(part + (size_t)q * S + 2 * k2) = make_double2(ax, ay);
    }
    return;
  }
  for (int k = kl; k < S; k += LW) {
    double acc = 0.0;
#pragma unroll 4
    for (uint32_t i = 0; i < len; ++i) {
      const double v = in[(size_t)src[f + i] * S + k];
      acc += isfinite(v) ? v : 0.0;
    }
    part[(size_t)q * S + k] = acc;
  }
}

template <int P>
__global__ void __launch_bounds__(256) k_m2l_hub_sum(const uint32_t* __restrict__ hubs,
                                                     const uint32_t* __restrict__ coff, uint32_t nhub,
                                                     const double* __restrict__ part, double* __restrict__ virt) {
  constexpr int S = num_moments(P);
  const size_t tt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= (size_t)nhub * S) return;
  const uint32_t hb = (uint32_t)(tt / S);
  const int k = (int)(tt % S);
  double acc = 0.0;
  for (uint32_t c = coff[hb]; c < coff[hb + 1]; ++c) acc += part[(size_t)c * S + k];
  virt[(size_t)hubs[hb] * S + k] = acc;
}

// Operators in the product kernel's fragment order. Lane l = 16 kq + c
// feeds A[16 rb + c][k_of(e, kq)] as slot e of row block rb; a key's values
// are stored as [rb][e / V][lane][e % V] with V = 16 / sizeof(value), so each
// 16-byte load of a wave reads 1 KB of contiguous memory. Zero padded.
template <int P, int MODE>
__host__ __device__ __forceinline__ void frag_pos(int t, int& r, int& k) {
  typedef Lanes<P, MODE> LN;
  constexpr int V = MODE == MX_F64 ? 2 : 4;
  const int v = t % V, lane = (t / V) % 64, eq = (t / (V * 64)) % (LN::KL / V), rb = t / (V * 64 * (LN::KL / V));
  r = 16 * rb + (lane & 15);
  k = LN::k_of(eq * V + v, lane >> 4);
}

// FP64 operator of each key in fragment order; TR: the transposed operator
// (A = T_K^T, the adjoint)
template <int P, bool TR>
__global__ void k_mx_operators64(const double* __restrict__ gtab, uint32_t nk, double* __restrict__ t64) {
  typedef Lanes<P, MX_F64> LN;
  constexpr int S = num_moments(P), ST = 2 * P + 1, GS = ST * ST;
  uint32_t key = blockIdx.x;
  if (key >= nk) return;
  const double* G = gtab + (size_t)key * GS;
  double fact[P + 1];
  fact[0] = 1.0;
  for (int q = 1; q <= P; ++q) fact[q] = fact[q - 1] * q;
  for (int t = threadIdx.x; t < LN::PER_KEY; t += blockDim.x) {
    int r, c;
    frag_pos<P, MX_F64>(t, r, c);
    if (TR) {
      const int q = r;
      r = c;
      c = q;
    }
    double v = 0.0;
    if (r < S && c < S) {
      int I, J, i, j;
      coef_ij2<P>(r, I, J);
      coef_ij2<P>(c, i, j);
      double w = (((i + j) & 1) ? -1.0 : 1.0) / (fact[I] * fact[J]);
      v = w * G[(i + I + j + J) * ST + (j + J)];
      if (!isfinite(v)) v = 0.0;
    }
    t64[(size_t)key * LN::PER_KEY + t] = v;
  }
}

// scaled FP32 operator of each key in the fragment order of MODE, and the
// key's scale exponent
template <int P, int MODE>
__global__ void k_mx_operators32(const double* __restrict__ gtab, const double* __restrict__ dxy, uint32_t nk,
                                 double eps2, float* __restrict__ that, int* __restrict__ sig_exp) {
  typedef Lanes<P, MODE> LN;
  constexpr int S = num_moments(P), ST = 2 * P + 1, GS = ST * ST;
  uint32_t key = blockIdx.x;
  if (key >= nk) return;
  const double dx = dxy[2 * key], dy = dxy[2 * key + 1];
  double sig = sqrt(dx * dx + dy * dy + eps2);
  if (!(sig > 0.0) || !isfinite(sig)) sig = 1.0;
  // the scale is the power of two 2^a nearest sig, so scaling is exact
  constexpr int AMAX = 1000 / (2 * P + 2);
  const int ae = min(max((int)rint(log2(sig)), -AMAX), AMAX);
  double pw[2 * P + 2];
  for (int q = 0; q < 2 * P + 2; ++q) pw[q] = ldexp(1.0, ae * q);
  if (threadIdx.x == 0) sig_exp[key] = ae;
  const double* G = gtab + (size_t)key * GS;
  double fact[P + 1];
  fact[0] = 1.0;
  for (int q = 1; q <= P; ++q) fact[q] = fact[q - 1] * q;
  for (int t = threadIdx.x; t < LN::PER_KEY; t += blockDim.x) {
    int r, c;
    frag_pos<P, MODE>(t, r, c);
    float v = 0.f;
    if (r < S && c < S) {
      int I, J, i, j;
      coef_ij2<P>(r, I, J);
      coef_ij2<P>(c, i, j);
      double w = (((i + j) & 1) ? -1.0 : 1.0) / (fact[I] * fact[J]);
      double g = G[(i + I + j + J) * ST + (j + J)];
      v = (float)(w * g * pw[i + j] * pw[I + J + 1]);
      if (!isfinite(v)) v = 0.f;
    }
    that[(size_t)key * LN::PER_KEY + t] = v;
  }
}

// ----------------------------------------------------------- products ----

// 3-term bf16 split x = hi + mid + lo
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  float r = x - (float)h;
  m = (__bf16)r;
  r = r - (float)m;
  l = (__bf16)r;
}

struct TileArgs {
  const uint32_t *item_tile, *item_s0, *item_s1, *item_slot;
  uint32_t nitems, n_rows;
  const uint32_t *step_key, *step_ids;
  const double* t64;
  const float* that;
  const int* sig_exp;  // per key: operators scaled by 2^(a (n + m + 1)), n, m the orders
  const double* in;  // multipoles (forward) or lbar (adjoint), S per node
  const double* virt;  // pre-summed long runs, S per run; ids n_in + r
  const double* zero;  // a row of zeros, read for missing ids
  // mixed forms: FP32 source rows of K floats (node rows scaled by 2^(-e ord),
  // pre-summed rows by the key's 2^(-a ord)) and per (step, column) the four
  // sources' shifts e - a (int8, packed; 0 for a pre-summed row)
  const float* in32;  // node rows, then pre-summed rows (ids n_in + r)
  const uint32_t* step_sh;
  uint32_t n_in;
  double* out;       // locals (forward) or mbar (adjoint), S per row
  double* part;      // partials of tiles with several items: [slot][16][S]
};

// Longest (row, key) run a product kernel sums itself; longer runs are
// pre-summed into one row. At 2^20 Plummer bodies, p = 6, 82% of the runs
// have one pair: the FP32 kernel takes one source per column (G = 1: 0.112 ->
// 0.084 ms, the M2L pass 0.156 -> 0.138 ms against G = 4), the FP64 kernel,
// bound by its matrix cores, two.
constexpr uint32_t mx_group_of(int mode) { return mode == MX_F64 ? 2u : 1u; }

// One block (4 waves) per work item; wave w takes steps s0 + w, s0 + w + 4, ...
// Each column of a step sums up to G source rows (a (row, key) run of at most
// G pairs; longer runs arrive as one pre-summed row).
// Operand layouts (lane l, column/row c = l % 16, quarter kq = l / 16):
//  v_mfma_*_16x16x4: A lane l holds A[c][slot kq]; B lane l holds B[slot kq][c];
//  v_mfma_f32_16x16x32_bf16: 8 consecutive slots 8 kq + 0..7 per lane;
//  D (16x16): v_mfma_f32_*: lane l holds D[4 kq + j][c], j = 0..3;
//             v_mfma_f64_16x16x4_f64: lane l holds D[kq + 4 j][c] (rows interleaved).
// TR (adjoint, FP64 only): A = T_K^T.
// The mixed forms with two row blocks (p = 5, 6) are held to five waves per
// SIMD (96 VGPRs, 16 bytes of spills: 0.114 -> 0.110 ms at 2^20 Plummer
// bodies, p = 6); larger rows would spill heavily.
template <int P, int MODE, bool TR, uint32_t G>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(MODE != MX_F64 && row_blocks(P) == 2 ? 5 : 1))) k_m2l_tiles(TileArgs a) {
  typedef Lanes<P, MODE> LN;
  constexpr int S = LN::S, RB = LN::RB, K = LN::K, KL = LN::KL, KS = K / 4;
  static_assert(!TR || MODE == MX_F64, "the transposed (adjoint) product is FP64");
  __shared__ double red[4][16][17];
  // one block per item; runs of consecutive items share an XCD (xcd_runs)
  const uint32_t item = xcd_runs(blockIdx.x, gridDim.x);
  if (item >= a.nitems) return;  // uniform over the block
  const uint32_t tile = a.item_tile[item], s0 = a.item_s0[item], s1 = a.item_s1[item], slot = a.item_slot[item];
  // w is wave-uniform: readfirstlane lets the step records and operator addresses live in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
  d4 L[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) L[rb] = (d4){0.0, 0.0, 0.0, 0.0};
  // per-lane orders of the B coefficients and of the D rows, 5 bits each
  // (the power-of-two scales of the mixed forms)
  constexpr int NOUT = RB * 4;
  uint32_t ord_in[(KL + 5) / 6], ord_out[(NOUT + 5) / 6];
#pragma unroll
  for (int q = 0; q < (KL + 5) / 6; ++q) ord_in[q] = 0;
#pragma unroll
  for (int q = 0; q < (NOUT + 5) / 6; ++q) ord_out[q] = 0;
  if (MODE != MX_F64) {
#pragma unroll
    for (int e = 0; e < KL; ++e) ord_in[e / 6] |= (uint32_t)order_of<P>(LN::k_of(e, kq)) << (5 * (e % 6));
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        ord_out[(rb * 4 + j) / 6] |= (uint32_t)(order_of<P>(rb * 16 + 4 * kq + j) + 1) << (5 * ((rb * 4 + j) % 6));
  }
  uint32_t st = s0 + w;
  // the step record (key, and the source ids of this lane's column) is
  // loaded one step ahead, so each step waits only on its multipole rows
  uint32_t key_n = 0, id_n[G], sh_n = 0;
#pragma unroll
  for (uint32_t q = 0; q < G; ++q) id_n[q] = NONE;
  if (st < s1) {
    key_n = a.step_key[st];
#pragma unroll
    for (uint32_t q = 0; q < G; ++q) id_n[q] = a.step_ids[((size_t)st * G + q) * 16 + col];
    if (MODE != MX_F64) sh_n = a.step_sh[(size_t)st * 16 + col];
  }
  for (; st < s1; st += 4) {
    const uint32_t key = key_n, sh = sh_n;
    uint32_t id[G];
#pragma unroll
    for (uint32_t q = 0; q < G; ++q) id[q] = id_n[q];
    if (st + 4 < s1) {
      key_n = a.step_key[st + 4];
#pragma unroll
      for (uint32_t q = 0; q < G; ++q) id_n[q] = a.step_ids[((size_t)(st + 4) * G + q) * 16 + col];
      if (MODE != MX_F64) sh_n = a.step_sh[(size_t)(st + 4) * 16 + col];
    }
    // every load of the step is issued before any is used: the operator rows
    // (their address needs only the key), then the source rows (a missing id
    // reads row 0 and is masked; padding coefficients k >= S read k = S - 1)
    constexpr int KC = LN::KC;
    typedef typename std::conditional<MODE == MX_F64, double, float>::type TT;
    // the operator in fragment order: slot e of row block rb for this lane
    TT tv[RB][KL];
    if (MODE == MX_F64) {
      const double2* T = reinterpret_cast<const double2*>(a.t64 + (size_t)key * LN::PER_KEY);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int eh = 0; eh < KL / 2; ++eh) {
          const double2 w = T[(rb * (KL / 2) + eh) * 64 + lane];
          tv[rb][2 * eh] = (TT)w.x;
          tv[rb][2 * eh + 1] = (TT)w.y;
        }
    } else {
      const float4* T = reinterpret_cast<const float4*>(a.that + (size_t)key * LN::PER_KEY);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int eq = 0; eq < KL / 4; ++eq) {
          const float4 w = T[(rb * (KL / 4) + eq) * 64 + lane];
          tv[rb][4 * eq] = (TT)w.x;
          tv[rb][4 * eq + 1] = (TT)w.y;
          tv[rb][4 * eq + 2] = (TT)w.z;
          tv[rb][4 * eq + 3] = (TT)w.w;
        }
    }
    const int nsx = MODE == MX_F64 ? 0 : -a.sig_exp[key];  // output scale 2^(-a (ord + 1)) per row
    // missing ids and padding coefficients k >= S load nothing (inactive
    // lanes make no cache accesses); the rows are finite (P2M and M2M map
    // non-finite sums to 0)
    // FP64 form: the group's sources summed in FP64 (list order)
    double bs[MODE == MX_F64 ? KL : 1];
    // mixed forms: B's column in FP32, each source row rescaled from its own
    // scale to the key's (an exact power of two), summed in list order
    float b32[MODE == MX_F64 ? 1 : KL];
    if constexpr (MODE == MX_F64) {
      double v[G][KL];
#pragma unroll
      for (uint32_t q = 0; q < G; ++q) {
        const uint32_t iq = id[q];
#pragma unroll
        for (int e = 0; e < KL; ++e) v[q][e] = 0.0;
        if (iq != NONE) {
          const double* row = iq < a.n_in ? a.in + (size_t)iq * S : a.virt + (size_t)(iq - a.n_in) * S;
          if constexpr (S % 2 == 0) {
            // even rows on 16-byte aligned bases (checked by the host): each
            // even k starts an aligned pair, and k < S holds for both or neither
#pragma unroll
            for (int e = 0; e < KL; e += 2) {
              const int k = LN::k_of(e, kq);
              if (k < S) {
                const double2 w = *reinterpret_cast<const double2*>(row + k);
                v[q][e] = w.x;
                v[q][e + 1] = w.y;
              }
            }
          } else {
#pragma unroll
            for (int e = 0; e < KL; ++e) {
              const int k = LN::k_of(e, kq);
              if (k < S) v[q][e] = row[k];
            }
          }
        }
      }
#pragma unroll
      for (int e = 0; e < KL; ++e) {  // a group's sources in list order
        bs[e] = v[0][e];
#pragma unroll
        for (uint32_t q = 1; q < G; ++q) bs[e] += v[q][e];
      }
    } else {
      // rows of K floats, zero past S: slots 4m .. 4m + 3 are one aligned
      // 16-byte load (k < K holds for all four or none)
      // a missing source loads nothing (inactive lanes make no cache
      // accesses, and the kernel is bound by the L1's accesses as much as by
      // its VALU work: loading a zero row instead was slower)
      float v[G][KL];
#pragma unroll
      for (uint32_t q = 0; q < G; ++q) {
        const uint32_t iq = id[q];
#pragma unroll
        for (int e = 0; e < KL; ++e) v[q][e] = 0.f;
        if (iq != NONE) {
          const float* row = a.in32 + (size_t)iq * K;
#pragma unroll
          for (int e = 0; e < KL; e += 4) {
            const int k = LN::k_of(e, kq);
            if (MODE == MX_F32 || k < K) {  // F32: k < K always (KL = K / 4)
              const float4 w = *reinterpret_cast<const float4*>(row + k);
              v[q][e] = w.x;
              v[q][e + 1] = w.y;
              v[q][e + 2] = w.z;
              v[q][e + 3] = w.w;
            }
          }
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < G; ++q) {
        const int t = (int)(int8_t)(uint8_t)(sh >> (8 * q));
#pragma unroll
        for (int e = 0; e < KL; ++e) {
          const float w = ldexpf(v[q][e], __mul24(t, (int)((ord_in[e / 6] >> (5 * (e % 6))) & 31u)));
          b32[e] = q == 0 ? w : b32[e] + w;
        }
      }
    }
    if constexpr (MODE == MX_F64) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int e = 0; e < KS; ++e)
          L[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)tv[rb][e], bs[e], L[rb], 0, 0, 0);
      }
    } else if constexpr (MODE == MX_F32) {
      // the row blocks' product chains interleaved: each MFMA's successor in
      // issue order is independent of it
      f4 acc[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) acc[rb] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < KS; ++e)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)tv[rb][e], b32[e], acc[rb], 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          L[rb][j] += ldexp((double)acc[rb][j], __mul24(nsx, (int)((ord_out[(rb * 4 + j) / 6] >> (5 * ((rb * 4 + j) % 6))) & 31u)));
    } else {
      bf8 bh[KC], bm[KC], bl[KC];
#pragma unroll
      for (int e = 0; e < KL; ++e) {
        __bf16 hh, mm, ll;
        split3(b32[e], hh, mm, ll);
        bh[e >> 3][e & 7] = hh;
        bm[e >> 3][e & 7] = mm;
        bl[e >> 3][e & 7] = ll;
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
          bf8 ah, am, al;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            __bf16 hh, mm, ll;
            split3((float)tv[rb][kc * 8 + e], hh, mm, ll);
            ah[e] = hh;
            am[e] = mm;
            al[e] = ll;
          }
          // small terms first
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[kc], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[kc], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[kc], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[kc], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[kc], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[kc], acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
          L[rb][j] += ldexp((double)acc[j], __mul24(nsx, (int)((ord_out[(rb * 4 + j) / 6] >> (5 * ((rb * 4 + j) % 6))) & 31u)));
      }
    }
  }
  // sum the four waves' partial results in wave order
  const uint32_t t = tile * 16 + col;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
    for (int j = 0; j < 4; ++j) red[w][MODE == MX_F64 ? kq + 4 * j : 4 * kq + j][col] = L[rb][j];
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = 4 * kq + j, row = rb * 16 + r;
        const double v = ((red[0][r][col] + red[1][r][col]) + red[2][r][col]) + red[3][r][col];
        if (row < S) {
          if (slot == NONE) {
            if (t < a.n_rows) a.out[(size_t)t * S + row] = v;
          } else {
            a.part[((size_t)slot * 16 + col) * S + row] = v;
          }
        }
      }
    }
    __syncthreads();
  }
}

// tiles with several items: sum their partials in item order
__global__ void __launch_bounds__(256) k_m2l_tiles_combine(const uint32_t* __restrict__ multi_tile,
                                                           const uint32_t* __restrict__ multi_slot,
                                                           const uint32_t* __restrict__ multi_cnt, int S,
                                                           uint32_t n_rows, const double* __restrict__ part,
                                                           double* __restrict__ out) {
  const uint32_t q = blockIdx.x;
  const uint32_t tile = multi_tile[q], slot0 = multi_slot[q], nit = multi_cnt[q];
  for (int idx = threadIdx.x; idx < 16 * S; idx += blockDim.x) {
    const int col = idx / S, row = idx % S;
    const uint32_t t = tile * 16 + col;
    if (t >= n_rows) continue;
    double v = 0.0;
#pragma unroll 8
    for (uint32_t k = 0; k < nit; ++k) v += part[((size_t)(slot0 + k) * 16 + col) * S + row];
    out[(size_t)t * S + row] = v;
  }
}

// ------------------------------------------------------- host drivers ----

static int upload_u32(DBuf& b, const std::vector<uint32_t>& v, hipStream_t s) {
  GX_CHECK(dbuf_reserve(b, std::max<size_t>(v.size(), 1) * 4, s));
  if (!v.empty()) GX_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * 4, hipMemcpyHostToDevice, s));
  return 0;
}

// Builds tiles over n_rows rows from the live pairs: row_of_pair[e] names the
// row of pair e, other_of_pair[e] the node whose coefficients it reads.
static int build_tiles(galax_fmm* h, const uint32_t* row_of_pair, const uint32_t* other_of_pair, uint32_t n_rows,
                       M2LTiles& T) {
  hipStream_t s = h->stream;
  const uint32_t nk = (uint32_t)h->n_keys, ne = (uint32_t)h->n_live_pairs;
  const uint32_t ntiles = (n_rows + 15) / 16;
  T.n_rows = n_rows;
  T.ntiles = ntiles;
  T.nlong = 0;
  T.nhub = 0;
  T.ngroups = 0;
  T.nsteps = 0;
  GX_CHECK(dbuf_reserve(T.tile_off, (size_t)(ntiles + 1) * 4, s));
  uint32_t* tile_off = T.tile_off.as<uint32_t>();
  GX_HIP(hipMemsetAsync(tile_off, 0, (size_t)(ntiles + 1) * 4, s));
  if (ne > 0 && n_rows > 0) {
    const int kb = bits_of(nk);
    const int tb = bits_of(n_rows);
    if (kb + tb > 59) return set_error(GALAX_ERR_UNSUPPORTED, "M2L tiles: too many keys and rows");
    Workspace& ws = h->ws;
    // 1. pairs sorted by (row, key), list order kept inside a group
    GX_CHECK(dbuf_reserve(ws.sort_ka, (size_t)ne * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_kb, (size_t)ne * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_va, (size_t)ne * 4, s));
    GX_CHECK(dbuf_reserve(ws.sort_vb, (size_t)ne * 4, s));
    GX_CHECK(dbuf_reserve(h->flag, 64, s));
    auto* ka = ws.sort_ka.as<unsigned long long>();
    auto* kb2 = ws.sort_kb.as<unsigned long long>();
    uint32_t* va = ws.sort_va.as<uint32_t>();
    uint32_t* vb = ws.sort_vb.as<uint32_t>();
    uint32_t* err = h->flag.as<uint32_t>() + 12;
    GX_HIP(hipMemsetAsync(err, 0, 4, s));
    k_mx_pair_keys<<<nblk(ne), 256, 0, s>>>(row_of_pair, h->key_of_pair.as<uint32_t>(), ne, n_rows, kb, ka, va, err);
    GX_HIP(hipGetLastError());
    uint32_t herr = 0;
    GX_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (herr) return set_error(GALAX_ERR_INVALID, "M2L tiles: a pair's row lies outside the live nodes");
    GX_CHECK(radix_sort_bits(h, reinterpret_cast<uint64_t*>(ka), va, reinterpret_cast<uint64_t*>(kb2), vb, ne,
                             kb + tb));
    // 2. groups: one per (row, key) run
    GX_CHECK(dbuf_reserve(ws.a, (size_t)ne * 4, s));
    GX_CHECK(dbuf_reserve(ws.h, (size_t)ne * 4, s));
    uint32_t* gid = ws.a.as<uint32_t>();
    uint32_t* run_start = ws.h.as<uint32_t>();
    k_mx_run_starts<<<nblk(ne), 256, 0, s>>>(ka, ne, run_start);
    GX_HIP(hipGetLastError());
    GX_CHECK(scan_u32(h, run_start, run_start, ne, true, true, nullptr));
    k_mx_chunk_flags<<<nblk(ne), 256, 0, s>>>(run_start, ne, gid);
    GX_HIP(hipGetLastError());
    uint32_t ng = 0;  // exclusive: a group start i is group gid[i]
    GX_CHECK(scan_u32(h, gid, gid, ne, false, false, &ng));
    T.ngroups = ng;
    const int rkb = 1;  // every (row, key) run is one group: its rank field is always 0
    if (bits_of(ntiles) + kb + rkb + 4 > 64) return set_error(GALAX_ERR_UNSUPPORTED, "M2L tiles: key space too wide");
    GX_CHECK(dbuf_reserve(T.src, (size_t)ne * 4, s));
    GX_CHECK(dbuf_reserve(ws.b, (size_t)(ng + 1) * 4, s));
    GX_CHECK(dbuf_reserve(ws.c, (size_t)ng * 8, s));
    GX_CHECK(dbuf_reserve(ws.d, (size_t)ng * 8, s));
    GX_CHECK(dbuf_reserve(ws.e, (size_t)ng * 4, s));
    GX_CHECK(dbuf_reserve(ws.g, (size_t)ng * 4, s));
    uint32_t* gfirst = ws.b.as<uint32_t>();
    auto* g2k = ws.c.as<unsigned long long>();
    auto* g2kb = ws.d.as<unsigned long long>();
    uint32_t* g2v = ws.e.as<uint32_t>();
    uint32_t* g2vb = ws.g.as<uint32_t>();
    k_mx_groups<<<nblk(ne), 256, 0, s>>>(ka, va, run_start, gid, ne, other_of_pair, kb, rkb, T.src.as<uint32_t>(),
                                         gfirst, g2k, g2v);
    k_mx_set_u32<<<1, 1, 0, s>>>(gfirst + ng, ne);
    GX_HIP(hipGetLastError());
    // 3. groups sorted by (tile, key, rank, column): runs of equal
    // (tile, key, rank) are steps
    GX_CHECK(radix_sort_bits(h, reinterpret_cast<uint64_t*>(g2k), g2v, reinterpret_cast<uint64_t*>(g2kb), g2vb, ng,
                             bits_of(ntiles) + kb + rkb + 4));
    uint32_t* sid = ws.h.as<uint32_t>();  // run_start is no longer needed
    k_mx_run_flags<<<nblk(ng), 256, 0, s>>>(g2k, ng, 4, sid);
    GX_HIP(hipGetLastError());
    uint32_t ns = 0;  // inclusive: entry j belongs to step sid[j] - 1
    GX_CHECK(scan_u32(h, sid, sid, ng, true, false, &ns));
    T.nsteps = ns;
    GX_CHECK(dbuf_reserve(T.step_key, (size_t)ns * 4, s));
    GX_CHECK(dbuf_reserve(T.step_first, (size_t)ns * 64, s));  // group id per (step, column)
    GX_HIP(hipMemsetAsync(T.step_first.p, 0xFF, (size_t)ns * 64, s));
    k_mx_steps<<<nblk(ng), 256, 0, s>>>(g2k, g2v, sid, ng, kb, rkb, T.step_key.as<uint32_t>(),
                                        T.step_first.as<uint32_t>(), tile_off);
    GX_HIP(hipGetLastError());
    // long runs (more than T.group pairs) become one pre-summed row each
    uint32_t* lrid = sid;  // the step ids are no longer needed
    k_mx_long_flags<<<nblk(ng), 256, 0, s>>>(gfirst, ng, T.group, lrid);
    GX_HIP(hipGetLastError());
    uint32_t nlr = 0;
    GX_CHECK(scan_u32(h, lrid, lrid, ng, false, false, &nlr));
    T.nlong = nlr;
    GX_CHECK(dbuf_reserve(T.lr_first, (size_t)std::max<uint32_t>(nlr, 1) * 4, s));
    GX_CHECK(dbuf_reserve(T.lr_len, (size_t)std::max<uint32_t>(nlr, 1) * 4, s));
    k_mx_long_table<<<nblk(ng), 256, 0, s>>>(gfirst, lrid, ng, T.group, T.lr_first.as<uint32_t>(),
                                             T.lr_len.as<uint32_t>());
    GX_HIP(hipGetLastError());
    // runs too long for one wave (hubs) get a block each
    T.nhub = 0;
    if (nlr) {
      GX_CHECK(dbuf_reserve(ws.g, (size_t)nlr * 4, s));
      uint32_t* hf = ws.g.as<uint32_t>();
      k_mx_hub_flags<<<nblk(nlr), 256, 0, s>>>(T.lr_len.as<uint32_t>(), nlr, hf);
      GX_HIP(hipGetLastError());
      uint32_t nh = 0;
      GX_CHECK(scan_u32(h, hf, hf, nlr, false, false, &nh));
      T.nhub = nh;
      GX_CHECK(dbuf_reserve(T.hubs, (size_t)std::max<uint32_t>(nh, 1) * 4, s));
      k_mx_hub_list<<<nblk(nlr), 256, 0, s>>>(T.lr_len.as<uint32_t>(), hf, nlr, T.hubs.as<uint32_t>());
      GX_HIP(hipGetLastError());
      // the pre-sum kernel's order: decreasing power-of-two length class
      // (a stable sort on the device), so that the runs of a wave are alike,
      // the longest start first and neighbouring runs of one class stay together
      k_mx_len_class<<<nblk(nlr), 256, 0, s>>>(T.lr_len.as<uint32_t>(), nlr, ka, va);
      GX_HIP(hipGetLastError());
      GX_CHECK(radix_sort_bits(h, reinterpret_cast<uint64_t*>(ka), va, reinterpret_cast<uint64_t*>(kb2), vb, nlr, 6));
      GX_CHECK(dbuf_reserve(T.lr_order, (size_t)nlr * 4, s));
      GX_HIP(hipMemcpyAsync(T.lr_order.p, va, (size_t)nlr * 4, hipMemcpyDeviceToDevice, s));
      // hub runs in chunks of MX_HUB_CHUNK rows
      T.nhub_chunks = 0;
      if (nh) {
        std::vector<uint32_t> len(nlr), hubs(nh), first(nlr), cf, cl, coff(nh + 1, 0);
        GX_HIP(hipMemcpyAsync(len.data(), T.lr_len.p, (size_t)nlr * 4, hipMemcpyDeviceToHost, s));
        GX_HIP(hipMemcpyAsync(hubs.data(), T.hubs.p, (size_t)nh * 4, hipMemcpyDeviceToHost, s));
        GX_HIP(hipMemcpyAsync(first.data(), T.lr_first.p, (size_t)nlr * 4, hipMemcpyDeviceToHost, s));
        GX_HIP(hipStreamSynchronize(s));
        for (uint32_t hb = 0; hb < nh; ++hb) {
          const uint32_t r = hubs[hb];
          coff[hb] = (uint32_t)cf.size();
          for (uint32_t i = 0; i < len[r]; i += MX_HUB_CHUNK) {
            cf.push_back(first[r] + i);
            cl.push_back(std::min(MX_HUB_CHUNK, len[r] - i));
          }
        }
        coff[nh] = (uint32_t)cf.size();
        T.nhub_chunks = cf.size();
        GX_CHECK(upload_u32(T.hub_first, cf, s));
        GX_CHECK(upload_u32(T.hub_len, cl, s));
        GX_CHECK(upload_u32(T.hub_coff, coff, s));
        GX_CHECK(dbuf_reserve(T.hub_part, T.nhub_chunks * num_moments(MAX_P) * 8, s));
      }
      GX_HIP(hipStreamSynchronize(s));
    }
    GX_CHECK(dbuf_reserve(T.step_ids, (size_t)ns * T.group * 16 * 4, s));
    GX_CHECK(dbuf_reserve(T.lr_key, (size_t)std::max<uint32_t>(nlr, 1) * 4, s));
    k_mx_step_ids<<<nblk((size_t)ns * 16), 256, 0, s>>>(T.step_first.as<uint32_t>(), gfirst, lrid,
                                                        T.src.as<uint32_t>(), ns, (uint32_t)h->tree.n_nodes,
                                                        T.group, T.step_key.as<uint32_t>(), T.step_ids.as<uint32_t>(),
                                                        T.lr_key.as<uint32_t>());
    GX_HIP(hipGetLastError());
  }
  uint32_t tot = 0;
  GX_CHECK(scan_u32(h, tile_off, tile_off, ntiles, false, false, &tot));
  k_mx_set_u32<<<1, 1, 0, s>>>(tile_off + ntiles, tot);
  GX_HIP(hipGetLastError());
  if (tot != T.nsteps) return set_error(GALAX_ERR_DEVICE, "M2L tiles: %u steps counted, %zu emitted", tot, T.nsteps);
  // 4. work items (host): at most MX_ITEM_STEPS steps each, one per empty tile
  std::vector<uint32_t> off(ntiles + 1);
  GX_HIP(hipMemcpyAsync(off.data(), tile_off, (ntiles + 1) * 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> it_tile, it_s0, it_s1, it_slot, m_tile, m_slot, m_cnt;
  uint32_t nslots = 0;
  for (uint32_t tt = 0; tt < ntiles; ++tt) {
    const uint32_t a0 = off[tt], a1 = off[tt + 1];
    const uint32_t nit = std::max<uint32_t>(1, (a1 - a0 + MX_ITEM_STEPS - 1) / MX_ITEM_STEPS);
    if (nit > 1) {
      m_tile.push_back(tt);
      m_slot.push_back(nslots);
      m_cnt.push_back(nit);
    }
    for (uint32_t k = 0; k < nit; ++k) {
      it_tile.push_back(tt);
      it_s0.push_back(std::min(a1, a0 + k * MX_ITEM_STEPS));
      it_s1.push_back(std::min(a1, a0 + (k + 1) * MX_ITEM_STEPS));
      it_slot.push_back(nit > 1 ? nslots + k : NONE);
    }
    if (nit > 1) nslots += nit;
  }
  // items by decreasing power-of-two step class (stable): the longest start
  // first instead of forming the last round of blocks, and neighbouring tiles
  // of one class stay together; each item writes its own slot, so the
  // combine's order does not depend on the order items run in
  {
    auto cls = [](uint32_t n) { return n ? 32 - __builtin_clz(n) : 0; };
    std::vector<uint32_t> ord(it_tile.size());
    for (uint32_t q = 0; q < ord.size(); ++q) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t u, uint32_t v) {
      return cls(it_s1[u] - it_s0[u]) > cls(it_s1[v] - it_s0[v]);
    });
    auto perm = [&](std::vector<uint32_t>& a) {
      std::vector<uint32_t> b(a.size());
      for (size_t q = 0; q < a.size(); ++q) b[q] = a[ord[q]];
      a.swap(b);
    };
    perm(it_tile);
    perm(it_s0);
    perm(it_s1);
    perm(it_slot);
  }
  T.nitems = it_tile.size();
  T.nmulti = m_tile.size();
  T.nslots = nslots;
  GX_CHECK(upload_u32(T.item_tile, it_tile, s));
  GX_CHECK(upload_u32(T.item_s0, it_s0, s));
  GX_CHECK(upload_u32(T.item_s1, it_s1, s));
  GX_CHECK(upload_u32(T.item_slot, it_slot, s));
  GX_CHECK(upload_u32(T.multi_tile, m_tile, s));
  GX_CHECK(upload_u32(T.multi_slot, m_slot, s));
  GX_CHECK(upload_u32(T.multi_cnt, m_cnt, s));
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

template <int P, bool TR>
static int operators64(galax_fmm* h) {
  constexpr int PK = Lanes<P, MX_F64>::PER_KEY;
  const uint32_t nk = (uint32_t)h->n_keys;
  DBuf& buf = TR ? h->mx_t64T : h->mx_t64;
  if (nk) {
    GX_CHECK(dbuf_reserve(buf, (size_t)nk * PK * 8, h->stream));
    k_mx_operators64<P, TR><<<nk, 256, 0, h->stream>>>(h->gtab.as<double>(), nk, buf.as<double>());
    GX_HIP(hipGetLastError());
  }
  return 0;
}

template <int P, int MODE>
static int operators32(galax_fmm* h) {
  constexpr int PK = Lanes<P, MODE>::PER_KEY;
  const uint32_t nk = (uint32_t)h->n_keys;
  hipStream_t s = h->stream;
  if (nk) {
    GX_CHECK(dbuf_reserve(h->mx_that, (size_t)nk * PK * 4, s));
    GX_CHECK(dbuf_reserve(h->mx_sig, (size_t)nk * 4, s));
    k_mx_operators32<P, MODE><<<nk, 256, 0, s>>>(h->gtab.as<double>(), h->key_dxy.as<double>(), nk,
                                                 h->cfg.softening * h->cfg.softening, h->mx_that.as<float>(),
                                                 h->mx_sig.as<int>());
    GX_HIP(hipGetLastError());
  }
  return 0;
}

// Mixed forms: per-node row exponents, per-step shifts, and the FP32 rows
// (multipoles of the live levels 0-1, which P2M and M2M write each
// evaluation, and the pre-summed long runs, whose padding stays zero).
template <int P>
static int rows32_setup(galax_fmm* h, M2LTiles& T) {
  constexpr int K = 16 * row_blocks(P);
  hipStream_t s = h->stream;
  const DevTree& t = h->tree;
  const size_t nn = t.n_nodes;
  GX_CHECK(dbuf_reserve(h->mx_nexp, std::max<size_t>(nn, 1), s));
  if (nn) {
    k_mx_node_exp<<<nblk(nn), 256, 0, s>>>(t.half_width.as<double>(), (uint32_t)nn, h->mx_nexp.as<int8_t>());
    GX_HIP(hipGetLastError());
  }
  // node rows (written by P2M and M2M for the live levels), then pre-summed
  // rows; all zero to begin with (the padding stays so)
  const size_t rows = nn + T.nlong;
  GX_CHECK(dbuf_reserve(h->mp32, rows * K * 4, s));
  GX_HIP(hipMemsetAsync(h->mp32.p, 0, rows * K * 4, s));
  GX_CHECK(dbuf_reserve(T.step_sh, std::max<size_t>(T.nsteps, 1) * 16 * 4, s));
  GX_CHECK(dbuf_reserve(T.lr_tab, std::max<size_t>(T.nlong, 1) * 16, s));
  GX_CHECK(dbuf_reserve(T.lr_psh, std::max<size_t>(h->n_live_pairs, 1), s));
  if (T.nsteps) {
    GX_CHECK(dbuf_reserve(h->flag, 64, s));
    uint32_t* err = h->flag.as<uint32_t>() + 13;
    GX_HIP(hipMemsetAsync(err, 0, 4, s));
    if (T.nlong) {
      k_mx_run_tab32<<<nblk(T.nlong), 256, 0, s>>>(T.lr_order.as<uint32_t>(), T.lr_first.as<uint32_t>(),
                                                   T.lr_len.as<uint32_t>(), T.lr_key.as<uint32_t>(),
                                                   (uint32_t)T.nlong, T.src.as<uint32_t>(), h->mx_sig.as<int>(),
                                                   h->mx_nexp.as<int8_t>(), T.lr_tab.as<uint4>(),
                                                   T.lr_psh.as<int8_t>(), err);
      GX_HIP(hipGetLastError());
    }
    k_mx_step_shifts<<<nblk(T.nsteps * 16), 256, 0, s>>>(T.step_ids.as<uint32_t>(), T.step_key.as<uint32_t>(),
                                                         h->mx_sig.as<int>(), h->mx_nexp.as<int8_t>(), T.nsteps,
                                                         (uint32_t)nn, T.group, T.step_sh.as<uint32_t>(), err);
    GX_HIP(hipGetLastError());
    uint32_t herr = 0;
    GX_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (herr) return set_error(GALAX_ERR_UNSUPPORTED, "M2L (mixed): a node's size and distance differ by over 2^127");
  }
  h->mx_rows32 = true;
  return 0;
}

template <int P>
static int setup_impl(galax_fmm* h) {
  hipStream_t s = h->stream;
  const uint32_t ne = (uint32_t)h->n_live_pairs, nt = (uint32_t)h->n_live_targets;
  GX_CHECK(dbuf_reserve(h->ws.f, (size_t)std::max<uint32_t>(ne, 1) * 4, s));
  uint32_t* tgt = h->ws.f.as<uint32_t>();
  if (nt) {
    k_mx_pair_rows<<<nblk(nt), 256, 0, s>>>(h->tree.m2l_off.as<uint32_t>(), nt, tgt);
    GX_HIP(hipGetLastError());
  }
  h->mx_fwd.group = h->cfg.precision == GALAX_PREC_FP64 ? mx_group_of(MX_F64) : mx_group_of(MX_F32);
  GX_CHECK(build_tiles(h, tgt, h->tree.m2l_dat.as<uint32_t>(), nt, h->mx_fwd));
  if (h->cfg.precision == GALAX_PREC_FP64) {
    GX_CHECK((operators64<P, false>(h)));
  } else {
    if (h->cfg.precision == GALAX_PREC_MIXED_BF16X3)
      GX_CHECK((operators32<P, MX_BF16X3>(h)));
    else
      GX_CHECK((operators32<P, MX_F32>(h)));
    GX_CHECK(rows32_setup<P>(h, h->mx_fwd));
  }
  GX_HIP(hipStreamSynchronize(s));
  h->mx_valid = true;
  return 0;
}

template <int P>
static int adjoint_setup_impl(galax_fmm* h) {
  hipStream_t s = h->stream;
  const uint32_t ne = (uint32_t)h->n_live_pairs, nt = (uint32_t)h->n_live_targets;
  GX_CHECK(dbuf_reserve(h->ws.f, (size_t)std::max<uint32_t>(ne, 1) * 4, s));
  uint32_t* tgt = h->ws.f.as<uint32_t>();
  if (nt) {
    k_mx_pair_rows<<<nblk(nt), 256, 0, s>>>(h->tree.m2l_off.as<uint32_t>(), nt, tgt);
    GX_HIP(hipGetLastError());
  }
  // rows are sources; every source of a live pair is a live node (same level)
  h->mx_adj.group = mx_group_of(MX_F64);
  GX_CHECK(build_tiles(h, h->tree.m2l_dat.as<uint32_t>(), tgt, nt, h->mx_adj));
  GX_CHECK((operators64<P, true>(h)));
  GX_HIP(hipStreamSynchronize(s));
  h->mx_adj_valid = true;
  return 0;
}

template <int P, int MODE, bool TR>
static int run_tiles(galax_fmm* h, M2LTiles& T, const double* in, double* out) {
  hipStream_t s = h->stream;
  if (!T.nitems) return 0;
  constexpr int S = num_moments(P);
  if (T.nslots) GX_CHECK(dbuf_reserve(T.part, T.nslots * 16 * S * 8, s));
  TileArgs a;
  a.item_tile = T.item_tile.as<uint32_t>();
  a.item_s0 = T.item_s0.as<uint32_t>();
  a.item_s1 = T.item_s1.as<uint32_t>();
  a.item_slot = T.item_slot.as<uint32_t>();
  a.nitems = (uint32_t)T.nitems;
  a.n_rows = (uint32_t)T.n_rows;
  a.step_key = T.step_key.as<uint32_t>();
  a.step_ids = T.step_ids.as<uint32_t>();
  a.n_in = (uint32_t)h->tree.n_nodes;
  constexpr bool rows32 = MODE != MX_F64;
  if (rows32 && !h->mx_rows32) return set_error(GALAX_ERR_INVALID, "M2L (mixed): FP32 rows not set up");
  if (rows32 && T.nlong) {
    k_m2l_run_sums32<P><<<nblk(T.nlong, 4 * (64 / sums32_lanes<P>())), 256, 0, s>>>(
        T.lr_tab.as<uint4>(), (uint32_t)T.nlong, T.src.as<uint32_t>(), T.lr_psh.as<int8_t>(), h->mp32.as<float>(),
        h->mp32.as<float>() + (size_t)h->tree.n_nodes * (16 * row_blocks(P)));
    GX_HIP(hipGetLastError());
  } else if (T.nlong) {
    GX_CHECK(dbuf_reserve(T.virt, T.nlong * S * 8, s));
    if ((uintptr_t)in & 15) return set_error(GALAX_ERR_INVALID, "M2L source rows need 16-byte aligned bases");
    k_m2l_run_sums_w<P><<<nblk(T.nlong, 4 * (64 / sums_lanes(S))), 256, 0, s>>>(T.lr_order.as<uint32_t>(),
                                                                       T.lr_first.as<uint32_t>(), T.lr_len.as<uint32_t>(),
                                                          (uint32_t)T.nlong, T.src.as<uint32_t>(), in,
                                                          T.virt.as<double>());
    if (T.nhub) {
      k_m2l_hub_chunks<P><<<nblk(T.nhub_chunks, 4 * (64 / sums_lanes(S))), 256, 0, s>>>(
          T.hub_first.as<uint32_t>(), T.hub_len.as<uint32_t>(), (uint32_t)T.nhub_chunks, T.src.as<uint32_t>(), in,
          T.hub_part.as<double>());
      GX_HIP(hipGetLastError());
      k_m2l_hub_sum<P><<<nblk(T.nhub * S), 256, 0, s>>>(T.hubs.as<uint32_t>(), T.hub_coff.as<uint32_t>(),
                                                         (uint32_t)T.nhub, T.hub_part.as<double>(),
                                                         T.virt.as<double>());
    }
    GX_HIP(hipGetLastError());
  }
  a.virt = T.virt.as<double>();
  if (!h->mx_zero.p) {
    GX_CHECK(dbuf_reserve(h->mx_zero, 16 * row_blocks(MAX_P) * 8, s));
    GX_HIP(hipMemsetAsync(h->mx_zero.p, 0, 16 * row_blocks(MAX_P) * 8, s));
  }
  a.zero = h->mx_zero.as<double>();
  a.t64 = (TR ? h->mx_t64T : h->mx_t64).as<double>();
  a.that = h->mx_that.as<float>();
  a.sig_exp = h->mx_sig.as<int>();
  a.in = in;
  a.in32 = h->mp32.as<float>();
  a.step_sh = T.step_sh.as<uint32_t>();
  if (((uintptr_t)a.in | (uintptr_t)a.virt) & 15)
    return set_error(GALAX_ERR_INVALID, "M2L source rows need 16-byte aligned bases");
  a.out = out;
  a.part = T.part.as<double>();
  if (!TR && h->timing) GX_HIP(hipEventRecord(h->ev[8], s));
  if (T.group != mx_group_of(MODE)) return set_error(GALAX_ERR_INVALID, "M2L tiles: built for another form");
  k_m2l_tiles<P, MODE, TR, mx_group_of(MODE)><<<a.nitems, 256, 0, s>>>(a);
  GX_HIP(hipGetLastError());
  if (!TR && h->timing) GX_HIP(hipEventRecord(h->ev[9], s));
  if (T.nmulti) {
    k_m2l_tiles_combine<<<(unsigned)T.nmulti, 256, 0, s>>>(T.multi_tile.as<uint32_t>(), T.multi_slot.as<uint32_t>(),
                                                           T.multi_cnt.as<uint32_t>(), S, (uint32_t)T.n_rows,
                                                           T.part.as<double>(), out);
    GX_HIP(hipGetLastError());
  }
  return 0;
}

template <int P>
static int run_impl(galax_fmm* h, const double* mp, double* loc) {
  switch (h->cfg.precision) {
    case GALAX_PREC_FP64:
      return run_tiles<P, MX_F64, false>(h, h->mx_fwd, mp, loc);
    case GALAX_PREC_MIXED_FP32:
      return run_tiles<P, MX_F32, false>(h, h->mx_fwd, mp, loc);
    case GALAX_PREC_MIXED_BF16X3:
      return run_tiles<P, MX_BF16X3, false>(h, h->mx_fwd, mp, loc);
  }
  return set_error(GALAX_ERR_INVALID, "M2L tiles: precision %u has no tiled form", h->cfg.precision);
}

#define GX_P_CASES(F) \
  F(0) F(1) F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15) F(16)

int m2l_mfma_setup(galax_fmm* h) {
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return setup_impl<PP>(h);
    GX_P_CASES(GX_CASE)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

int m2l_mfma_run(galax_fmm* h, const double* mp, double* loc) {
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return run_impl<PP>(h, mp, loc);
    GX_P_CASES(GX_CASE)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

int m2l_adjoint_setup(galax_fmm* h) {
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return adjoint_setup_impl<PP>(h);
    GX_P_CASES(GX_CASE)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

int m2l_adjoint_run(galax_fmm* h, const double* lbar, double* mbar) {
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return run_tiles<PP, MX_F64, true>(h, h->mx_adj, lbar, mbar);
    GX_P_CASES(GX_CASE)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

}  // namespace gx
