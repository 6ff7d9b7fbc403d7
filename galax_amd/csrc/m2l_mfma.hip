// m2l_mfma.hip — M2L of the reference variant as per-key batched dense
// contractions on the matrix cores (mixed-precision forms).
//
// galax-core's M2L (galax-core/src/lib.rs:1199-1259) uses, for every pair of a
// target node t and source node s, the derivative table of the first pair with
// the same key (dx as i64, dy as i64). The operator of a pair therefore
// depends on its key only:
//     L_t[I,J] += sum_{i,j} (-1)^{i+j} / (I! J!) * G_K[i+I, j+J] * M_s[i,j]
// so pairs sorted by key form one dense product per key: T_K (S x S) times a
// block of multipole columns. Here every key gets a scaled FP32 operator
//     T^_K[IJ][ij] = (-1)^{i+j}/(I!J!) G_K[i+I, j+J] * sig^(i+j) * sig^(I+J+1)
// with sig the representative's softened distance, columns are
// M^_ij = M_ij * sig^-(i+j), and results are unscaled by sig^-(I+J+1). Tiles of
// 16 pairs of one key run on v_mfma_f32_16x16x4_f32 (FP32 form) or on
// v_mfma_f32_16x16x32_bf16 with a 3-term bf16 split of both operands (BF16x3
// form). Each pair's result column is written to its own slot and a second
// kernel sums each target's pairs in list order (FP64), so results do not
// depend on scheduling.
#include <algorithm>
#include <cmath>

#include "common.h"

namespace gx {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

static inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

__host__ __device__ constexpr int row_blocks(int p) { return (num_moments(p) + 15) / 16; }

template <int P>
__device__ __forceinline__ int order_of(int k) {
  int n = 0;
#pragma unroll
  for (int q = 0; q <= P; ++q)
    if (k >= (q + 1) * (q + 2) / 2) n = q + 1;
  return n;
}

template <int P>
__device__ __forceinline__ void coef_ij2(int k, int& i, int& j) {
  int n = order_of<P>(k);
  j = k - n * (n + 1) / 2;
  i = n - j;
}

// ------------------------------------------------------------- setup ----

__global__ void k_mx_count(const uint32_t* __restrict__ key_of_pair, uint32_t n_pairs, uint32_t* __restrict__ cnt) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n_pairs) atomicAdd(&cnt[key_of_pair[e]], 1u);
}

__global__ void k_mx_pad(const uint32_t* __restrict__ cnt, uint32_t nk, uint32_t* __restrict__ padded,
                         uint32_t* __restrict__ ntile) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  uint32_t t = (cnt[k] + 15) / 16;
  ntile[k] = t;
  padded[k] = t * 16;
}

// positions: padded segment per key; any order inside a segment gives the same
// values (each pair's column is computed independently)
__global__ void k_mx_place(const uint32_t* __restrict__ key_of_pair, const uint32_t* __restrict__ dat,
                           uint32_t n_pairs, const uint32_t* __restrict__ koffp, uint32_t* __restrict__ cursor,
                           uint32_t* __restrict__ pos_of_pair, uint32_t* __restrict__ src_of_pos) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pairs) return;
  uint32_t k = key_of_pair[e];
  uint32_t pos = koffp[k] + atomicAdd(&cursor[k], 1u);
  pos_of_pair[e] = pos;
  src_of_pos[pos] = dat[e];
}

__global__ void k_mx_tiles(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ koffp,
                           const uint32_t* __restrict__ toff, uint32_t nk, uint32_t* __restrict__ tiles) {
  uint32_t k = blockIdx.x;
  if (k >= nk) return;
  uint32_t nt = (cnt[k] + 15) / 16;
  for (uint32_t j = threadIdx.x; j < nt; j += blockDim.x) {
    uint32_t tt = toff[k] + j;
    tiles[3 * tt] = k;
    tiles[3 * tt + 1] = koffp[k] + 16 * j;
    tiles[3 * tt + 2] = min(16u, cnt[k] - 16 * j);
  }
}

// scaled operator of each key (FP32, (16 RB)^2 row-major, zero padded) and
// the key's inverse scale powers sig^-n, n = 0..P+1 (FP64)
template <int P>
__global__ void k_mx_operators(const double* __restrict__ gtab, const double* __restrict__ dxy, uint32_t nk,
                               double eps2, float* __restrict__ that, double* __restrict__ sig_inv) {
  constexpr int S = num_moments(P), K = 16 * row_blocks(P), ST = 2 * P + 1, GS = ST * ST;
  uint32_t key = blockIdx.x;
  if (key >= nk) return;
  const double dx = dxy[2 * key], dy = dxy[2 * key + 1];
  double sig = sqrt(dx * dx + dy * dy + eps2);
  if (!(sig > 0.0) || !isfinite(sig)) sig = 1.0;
  double pw[2 * P + 2];
  pw[0] = 1.0;
  for (int q = 1; q < 2 * P + 2; ++q) pw[q] = pw[q - 1] * sig;
  if (threadIdx.x < P + 2) sig_inv[(size_t)key * (P + 2) + threadIdx.x] = 1.0 / pw[threadIdx.x];
  const double* G = gtab + (size_t)key * GS;
  double fact[P + 1];
  fact[0] = 1.0;
  for (int q = 1; q <= P; ++q) fact[q] = fact[q - 1] * q;
  for (int t = threadIdx.x; t < K * K; t += blockDim.x) {
    int r = t / K, c = t % K;
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
    that[(size_t)key * K * K + t] = v;
  }
}

// ----------------------------------------------------------- products ----

// One wave per tile of 16 pairs. Operand layouts of v_mfma_f32_16x16x4_f32:
// A (16x4): lane l holds A[l%16][l/16]; B (4x16): lane l holds B[l/16][l%16];
// D (16x16): lane l holds D[4*(l/16)+j][l%16], j = 0..3.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_mfma_f32(const uint32_t* __restrict__ tiles, uint32_t ntiles,
                                                      const uint32_t* __restrict__ src_of_pos,
                                                      const float* __restrict__ that,
                                                      const double* __restrict__ sig_inv,
                                                      const double* __restrict__ mp, float* __restrict__ R) {
  constexpr int S = num_moments(P), RB = row_blocks(P), K = 16 * RB, KS = K / 4;
  const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
  if (tile >= ntiles) return;
  const uint32_t key = tiles[3 * tile], base = tiles[3 * tile + 1], cnt = tiles[3 * tile + 2];
  const uint32_t pos = base + col;
  const bool live = (uint32_t)col < cnt;
  const uint32_t src = live ? src_of_pos[pos] : 0u;
  const double* sg = sig_inv + (size_t)key * (P + 2);
  float b[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = ks * 4 + kq;
    float v = 0.f;
    if (live && k < S) {
      double mv = mp[(size_t)src * S + k];
      if (!isfinite(mv)) mv = 0.0;
      v = (float)(mv * sg[order_of<P>(k)]);
    }
    b[ks] = v;
  }
  const float* T = that + (size_t)key * K * K;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const float* Trow = T + (size_t)(rb * 16 + col) * K;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(Trow[ks * 4 + kq], b[ks], acc, 0, 0, 0);
    if (live) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kk = rb * 16 + 4 * kq + j;
        if (kk < S) R[(size_t)pos * S + kk] = acc[j] * (float)sg[order_of<P>(kk) + 1];
      }
    }
  }
}

// 3-term bf16 split x = hi + mid + lo
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  float r = x - (float)h;
  m = (__bf16)r;
  r = r - (float)m;
  l = (__bf16)r;
}

// BF16x3 form. Operand layouts of v_mfma_f32_16x16x32_bf16: A (16x32): lane l
// holds A[l%16][8*(l/16) + 0..7]; B (32x16): lane l holds B[8*(l/16) + 0..7][l%16];
// D as above. Products kept: hh, hm, mh, hl, lh, mm (the terms of order <= 2
// in the split), accumulated in FP32.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_mfma_bf16x3(const uint32_t* __restrict__ tiles, uint32_t ntiles,
                                                         const uint32_t* __restrict__ src_of_pos,
                                                         const float* __restrict__ that,
                                                         const double* __restrict__ sig_inv,
                                                         const double* __restrict__ mp, float* __restrict__ R) {
  constexpr int S = num_moments(P), RB = row_blocks(P), K = 16 * RB, KC = (K + 31) / 32;
  const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
  if (tile >= ntiles) return;
  const uint32_t key = tiles[3 * tile], base = tiles[3 * tile + 1], cnt = tiles[3 * tile + 2];
  const uint32_t pos = base + col;
  const bool live = (uint32_t)col < cnt;
  const uint32_t src = live ? src_of_pos[pos] : 0u;
  const double* sg = sig_inv + (size_t)key * (P + 2);
  bf8 bh[KC], bm[KC], bl[KC];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kc * 32 + kq * 8 + e;
      float v = 0.f;
      if (live && k < S) {
        double mv = mp[(size_t)src * S + k];
        if (!isfinite(mv)) mv = 0.0;
        v = (float)(mv * sg[order_of<P>(k)]);
      }
      __bf16 h, m, l;
      split3(v, h, m, l);
      bh[kc][e] = h;
      bm[kc][e] = m;
      bl[kc][e] = l;
    }
  const float* T = that + (size_t)key * K * K;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    const float* Trow = T + (size_t)(rb * 16 + col) * K;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      bf8 ah, am, al;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = kc * 32 + kq * 8 + e;
        float v = k < K ? Trow[k] : 0.f;
        __bf16 h, m, l;
        split3(v, h, m, l);
        ah[e] = h;
        am[e] = m;
        al[e] = l;
      }
      // small terms first
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[kc], acc, 0, 0, 0);
    }
    if (live) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kk = rb * 16 + 4 * kq + j;
        if (kk < S) R[(size_t)pos * S + kk] = acc[j] * (float)sg[order_of<P>(kk) + 1];
      }
    }
  }
}

// Sum of each target's pair results in list order (FP64); one wave per target.
template <int P>
__global__ void __launch_bounds__(256) k_m2l_gather(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ pos_of_pair, uint32_t n_targets,
                                                    const float* __restrict__ R, double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= n_targets) return;
  for (int k = lane; k < S; k += 64) {
    double acc = 0.0;
    for (uint32_t e = off[t]; e < off[t + 1]; ++e) acc += (double)R[(size_t)pos_of_pair[e] * S + k];
    loc[(size_t)t * S + k] = acc;
  }
}

// ------------------------------------------------------- host drivers ----

template <int P>
static int setup_impl(galax_fmm* h) {
  constexpr int K = 16 * row_blocks(P);
  hipStream_t s = h->stream;
  const uint32_t nk = (uint32_t)h->n_keys, ne = (uint32_t)h->n_live_pairs;
  h->mx_ntiles = 0;
  if (ne == 0) {
    h->mx_valid = true;
    return 0;
  }
  GX_CHECK(dbuf_reserve(h->ws.a, (size_t)(4 * nk + 8) * sizeof(uint32_t), s));
  uint32_t* cnt = h->ws.a.as<uint32_t>();
  uint32_t* koffp = cnt + nk + 1;
  uint32_t* toff = koffp + nk + 1;
  uint32_t* cursor = toff + nk + 1;
  GX_HIP(hipMemsetAsync(cnt, 0, (size_t)(nk + 1) * 4, s));
  GX_HIP(hipMemsetAsync(cursor, 0, (size_t)(nk + 1) * 4, s));
  k_mx_count<<<nblk(ne), 256, 0, s>>>(h->key_of_pair.as<uint32_t>(), ne, cnt);
  k_mx_pad<<<nblk(nk), 256, 0, s>>>(cnt, nk, koffp, toff);
  GX_HIP(hipGetLastError());
  uint32_t npad = 0, ntiles = 0;
  GX_CHECK(scan_u32(h, koffp, koffp, nk, false, false, &npad));
  GX_CHECK(scan_u32(h, toff, toff, nk, false, false, &ntiles));
  h->mx_npad = npad;
  h->mx_ntiles = ntiles;
  GX_CHECK(dbuf_reserve(h->mx_pos, (size_t)ne * 4, s));
  GX_CHECK(dbuf_reserve(h->mx_src, (size_t)npad * 4, s));
  GX_CHECK(dbuf_reserve(h->mx_tiles, (size_t)ntiles * 12, s));
  k_mx_place<<<nblk(ne), 256, 0, s>>>(h->key_of_pair.as<uint32_t>(), h->tree.m2l_dat.as<uint32_t>(), ne, koffp,
                                      cursor, h->mx_pos.as<uint32_t>(), h->mx_src.as<uint32_t>());
  k_mx_tiles<<<nk, 64, 0, s>>>(cnt, koffp, toff, nk, h->mx_tiles.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(dbuf_reserve(h->mx_that, (size_t)nk * K * K * 4, s));
  GX_CHECK(dbuf_reserve(h->mx_sig, (size_t)nk * (P + 2) * 8, s));
  k_mx_operators<P><<<nk, 256, 0, s>>>(h->gtab.as<double>(), h->key_dxy.as<double>(), nk,
                                       h->cfg.softening * h->cfg.softening, h->mx_that.as<float>(),
                                       h->mx_sig.as<double>());
  GX_HIP(hipGetLastError());
  GX_CHECK(dbuf_reserve(h->mx_R, (size_t)npad * num_moments(P) * 4, s));
  GX_HIP(hipStreamSynchronize(s));
  h->mx_valid = true;
  return 0;
}

template <int P>
static int run_impl(galax_fmm* h, const double* mp, double* loc, bool bf16x3) {
  hipStream_t s = h->stream;
  if (h->mx_ntiles) {
    if (bf16x3)
      k_m2l_mfma_bf16x3<P><<<nblk(h->mx_ntiles, 4), 256, 0, s>>>(h->mx_tiles.as<uint32_t>(),
                                                                  (uint32_t)h->mx_ntiles, h->mx_src.as<uint32_t>(),
                                                                  h->mx_that.as<float>(), h->mx_sig.as<double>(), mp,
                                                                  h->mx_R.as<float>());
    else
      k_m2l_mfma_f32<P><<<nblk(h->mx_ntiles, 4), 256, 0, s>>>(h->mx_tiles.as<uint32_t>(), (uint32_t)h->mx_ntiles,
                                                               h->mx_src.as<uint32_t>(), h->mx_that.as<float>(),
                                                               h->mx_sig.as<double>(), mp, h->mx_R.as<float>());
    GX_HIP(hipGetLastError());
  }
  uint32_t nt = (uint32_t)h->n_live_targets;
  if (nt) {
    k_m2l_gather<P><<<nblk(nt, 4), 256, 0, s>>>(h->tree.m2l_off.as<uint32_t>(), h->mx_pos.as<uint32_t>(), nt,
                                                h->mx_R.as<float>(), loc);
    GX_HIP(hipGetLastError());
  }
  return 0;
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
  const bool bf = h->cfg.precision == GALAX_PREC_MIXED_BF16X3;
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return run_impl<PP>(h, mp, loc, bf);
    GX_P_CASES(GX_CASE)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

}  // namespace gx
