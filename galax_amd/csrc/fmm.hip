// fmm.hip — the FMM passes of compute_fmm_force (galax-core/src/lib.rs:1466-1499)
// on the device, FP64:
//   sanitize  :1475-1482 (in the P2M kernel)  P2M :848-887  M2M :890-938
//   M2L       :1199-1259 (derivatives :1040-1151, plan :1153-1196)
//   L2L       :1262-1302     L2P  :1308-1355     P2P :1395-1462
// Functions marked "exact" run without FMA contraction and in the reference's
// summation order, so their results match galax-core bit for bit given the
// same inputs; P2P sums in a gather order and uses FMA.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "common.h"

namespace gx {

// ----------------------------------------------------------- tables ----

struct Tabs {
  double fact[MAX_P + 1];
  double binom[MAX_P + 1][MAX_P + 1];
  double rfact2[MAX_P + 1][MAX_P + 1];  // RN(1 / (i! j!)), for div_exact
};

__constant__ Tabs c_tabs;

static Tabs host_tabs() {
  Tabs t;
  std::memset(&t, 0, sizeof(t));
  // galax-core/src/lib.rs:15-39
  for (int n = 0; n <= MAX_P; ++n) {
    t.binom[n][0] = 1.0;
    t.binom[n][n] = 1.0;
    for (int k = 1; k < n; ++k) t.binom[n][k] = t.binom[n - 1][k - 1] + t.binom[n - 1][k];
  }
  t.fact[0] = 1.0;
  for (int i = 1; i <= MAX_P; ++i) t.fact[i] = t.fact[i - 1] * (double)i;
  for (int i = 0; i <= MAX_P; ++i)
    for (int j = 0; j <= MAX_P; ++j) t.rfact2[i][j] = 1.0 / (t.fact[i] * t.fact[j]);
  return t;
}

// a / d rounded to nearest, from r = RN(1 / d): q = RN(a r), the residual
// a - q d is exact under FMA, and RN(q + (a - q d) r) = RN(a / d) (Markstein's
// theorem; it holds for normal results, which is every use here). Three FP64
// operations in place of a full division.
__device__ __forceinline__ double div_exact(double a, double d, double r) {
  const double q = a * r;
  return fma(fma(-q, d, a), r, q);
}

static int upload_tabs(hipStream_t s) {
  static thread_local int done_dev = -1;
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (done_dev == dev) return 0;
  Tabs t = host_tabs();
  GX_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tabs), &t, sizeof(t), 0, hipMemcpyHostToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  done_dev = dev;
  return 0;
}

template <int P>
__device__ __forceinline__ void coef_ij(int k, int& i, int& j) {
  int n = 0;
#pragma unroll
  for (int q = 0; q <= P; ++q)
    if (k >= (q + 1) * (q + 2) / 2) n = q + 1;
  j = k - n * (n + 1) / 2;
  i = n - j;
}

static inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// --------------------------------------------------------------- P2M ----

// exact: one wave per leaf, bodies in order (galax-core/src/lib.rs:854-885;
// coeff = mass*dx^i*dy^j/(i!*j!), non-finite -> 0). The kernel also applies
// compute_fmm_force's sanitize to every body (each body is in one leaf),
// writing fixed values back and counting them in *changed; it zeroes *next,
// the counter of the following evaluation (two counters in turn). The wave stages a chunk of
// the leaf's bodies in LDS as the tables m dx^0..m dx^P and dy^0..dy^P
// (dx = x - com, powers by repeated products; m dx^i is the first product of
// the reference's m dx^i dy^j, so every lane's term is the value it would
// compute), then each lane sums its coefficient over the chunk in body order.
// With S <= 32 the two half-waves take two leaves.
template <int P>
__global__ void __launch_bounds__(256) k_p2m(double* __restrict__ x, double* __restrict__ y, double* __restrict__ m,
                                             const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                             const double* __restrict__ comx, const double* __restrict__ comy,
                                             uint32_t n_leaves, double* __restrict__ mp,
                                             uint32_t* __restrict__ changed, uint32_t* __restrict__ next,
                                             float4* __restrict__ rel, float* __restrict__ mp32,
                                             const int8_t* __restrict__ nexp) {
#pragma clang fp contract(off)
  constexpr int S = num_moments(P);
  constexpr int LPW = S <= 32 ? 2 : 1;  // leaves per wave
  constexpr int LW = 64 / LPW;          // lanes per leaf = staging chunk
  __shared__ double smx[4][64][P + 1], spy[4][64][P + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = lane / LW, sl = lane % LW;
  if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0u;
  const uint32_t leaf = (blockIdx.x * 4 + w) * LPW + sub;
  const bool live = leaf < n_leaves;
  const uint32_t s = live ? bs[leaf] : 0u, e = live ? be[leaf] : 0u;
  const double cx = live ? comx[leaf] : 0.0, cy = live ? comy[leaf] : 0.0;
  // chunks run in lockstep over the block (barriers): its longest leaf sets the count
  __shared__ uint32_t blen[4];
  uint32_t len = e - s;
  if (LPW == 2) len = max(len, (uint32_t)__shfl_xor((int)len, 32, 64));
  if (lane == 0) blen[w] = len;
  __syncthreads();
  len = max(max(blen[0], blen[1]), max(blen[2], blen[3]));
  const int base = sub * LW;
  for (int k0 = 0; k0 < S; k0 += LW) {
    const int k = k0 + sl;
    int i = 0, j = 0;
    if (k < S) coef_ij<P>(k, i, j);
    const double div = c_tabs.fact[i] * c_tabs.fact[j], rdiv = c_tabs.rfact2[i][j];
    double acc = 0.0;
    for (uint32_t c = 0; c < len; c += LW) {
      __syncthreads();
      const uint32_t b = s + c + sl;
      if (b < e) {
        // sanitize (galax-core/src/lib.rs:1475-1482), written back once
        double xb = x[b], yb = y[b], mb = m[b];
        bool fix = false;
        if (fabs(xb) > 1e100 || fabs(yb) > 1e100 || !isfinite(xb) || !isfinite(yb)) {
          xb = 0.0;
          yb = 0.0;
          fix = true;
        }
        if (!isfinite(mb) || mb <= 0.0) {
          mb = 1.0;
          fix = true;
        }
        if (fix && k0 == 0) {
          x[b] = xb;
          y[b] = yb;
          m[b] = mb;
          atomicAdd(changed, 1u);
        }
        const double dx = xb - cx, dy = yb - cy;
        // the symmetric FP32 near field's sources: leaf-relative coordinates
        if (rel && k0 == 0) rel[b] = make_float4((float)dx, (float)dy, (float)mb, 0.f);
        double px = 1.0, py = 1.0;
        smx[w][base + sl][0] = mb * px;
        spy[w][base + sl][0] = 1.0;
#pragma unroll
        for (int r = 1; r <= P; ++r) {
          px = px * dx;
          py = py * dy;
          smx[w][base + sl][r] = mb * px;
          spy[w][base + sl][r] = py;
        }
      }
      __syncthreads();
      const uint32_t n = e > s + c ? min((uint32_t)LW, e - s - c) : 0u;
#pragma unroll 4
      for (uint32_t q = 0; q < n; ++q) {
        const double t = smx[w][base + q][i] * spy[w][base + q][j];
        acc += div_exact(t, div, rdiv);
      }
    }
    if (live && k < S) mp[(size_t)leaf * S + k] = isfinite(acc) ? acc : 0.0;
    if (mp32 && live) {  // the mixed forms' M2L rows: M_ij 2^(-e (i+j)), zero padded
      constexpr int K32 = 16 * ((S + 15) / 16);
      if (k < S)
        mp32[(size_t)leaf * K32 + k] = (float)ldexp(isfinite(acc) ? acc : 0.0, -(int)nexp[leaf] * (i + j));
      else if (k < K32)
        mp32[(size_t)leaf * K32 + k] = 0.f;
    }
  }
}

// --------------------------------------------------------------- M2M ----

// exact: parents in [lo, hi); children in order, each child's shifted sum
// added to the parent (galax-core/src/lib.rs:894-932). One wave per parent,
// or per two parents when S <= 32 (half-waves); lane = coefficient. For each
// child the wave stages in LDS the child's multipole and the power tables
// dx^a, dy^a (by repeated products, as each lane would compute them); the
// factorial products and their reciprocals sit in LDS for the block.
template <int P>
__global__ void __launch_bounds__(256) k_m2m(const uint32_t* __restrict__ cf, const uint8_t* __restrict__ cc,
                                             const double* __restrict__ comx, const double* __restrict__ comy,
                                             uint32_t lo, uint32_t hi, double* __restrict__ mp,
                                             float* __restrict__ mp32, const int8_t* __restrict__ nexp) {
#pragma clang fp contract(off)
  constexpr int S = num_moments(P);
  constexpr int PPW = S <= 32 ? 2 : 1;  // parents per wave
  constexpr int LW = 64 / PPW;          // lanes per parent
  constexpr int Q = P + 1;
  __shared__ double smc[4][PPW][S], sdx[4][PPW][Q], sdy[4][PPW][Q], sd[Q][Q], sr[Q][Q];
  __shared__ uint32_t bmax[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = lane / LW, sl = lane % LW;
  for (int q = threadIdx.x; q < Q * Q; q += 256) {
    sd[q / Q][q % Q] = c_tabs.fact[q / Q] * c_tabs.fact[q % Q];
    sr[q / Q][q % Q] = c_tabs.rfact2[q / Q][q % Q];
  }
  const uint32_t p = lo + (blockIdx.x * 4 + w) * PPW + sub;
  const bool live = p < hi;
  const uint32_t c0 = live ? cf[p] : 0u, nc = live ? cc[p] : 0u;
  const double px = live ? comx[p] : 0.0, py = live ? comy[p] : 0.0;
  // children run in lockstep over the block: its largest child count sets the loop
  uint32_t ncm = nc;
  if (PPW == 2) ncm = max(ncm, (uint32_t)__shfl_xor((int)ncm, 32, 64));
  if (lane == 0) bmax[w] = ncm;
  __syncthreads();
  ncm = max(max(bmax[0], bmax[1]), max(bmax[2], bmax[3]));
  double acc[(S + LW - 1) / LW];
#pragma unroll
  for (int r = 0; r < (S + LW - 1) / LW; ++r) acc[r] = 0.0;
  for (uint32_t ci = 0; ci < ncm; ++ci) {
    const bool has = ci < nc;
    __syncthreads();
    if (has) {
      const uint32_t c = c0 + ci;
      for (int k = sl; k < S; k += LW) smc[w][sub][k] = mp[(size_t)c * S + k];
      if (sl < Q) {
        const double dx = comx[c] - px, dy = comy[c] - py;
        double ex = 1.0, ey = 1.0;
        for (int r = 0; r < sl; ++r) {
          ex = ex * dx;
          ey = ey * dy;
        }
        sdx[w][sub][sl] = ex;
        sdy[w][sub][sl] = ey;
      }
    }
    __syncthreads();
    if (has) {
      const double* mc = smc[w][sub];
      const double* ex = sdx[w][sub];
      const double* ey = sdy[w][sub];
      // a child at the parent's COM (a single child usually is): every term
      // but i = oi, j = oj is +-0, which leaves the running sum at +0, so the
      // sum is 0 + M_c(oi, oj) exactly as the full loop gives it
      const bool at_com = Q == 1 || (ex[1] == 0.0 && ey[1] == 0.0);
#pragma unroll
      for (int r = 0; r < (S + LW - 1) / LW; ++r) {
        const int k = sl + r * LW;
        if (k < S) {
          int oi, oj;
          coef_ij<P>(k, oi, oj);
          double sum = 0.0;
          if (at_com) {
            sum = 0.0 + mc[k];
          } else {
            for (int i = 0; i <= oi; ++i)
              for (int j = 0; j <= oj; ++j)
                sum += div_exact(mc[moment_index(i, j)] * ex[oi - i] * ey[oj - j], sd[oi - i][oj - j],
                                 sr[oi - i][oj - j]);
          }
          acc[r] += sum;
        }
      }
    }
  }
  if (live) {
#pragma unroll
    for (int r = 0; r < (S + LW - 1) / LW; ++r) {
      const int k = sl + r * LW;
      if (k < S) mp[(size_t)p * S + k] = isfinite(acc[r]) ? acc[r] : 0.0;  // non-finite -> 0 (galax-core/src/lib.rs:934-936)
      if (mp32) {  // the mixed forms' M2L rows: M_ij 2^(-e (i+j)), zero padded
        constexpr int K32 = 16 * ((S + 15) / 16);
        if (k < S) {
          int oi, oj;
          coef_ij<P>(k, oi, oj);
          mp32[(size_t)p * K32 + k] = (float)ldexp(isfinite(acc[r]) ? acc[r] : 0.0, -(int)nexp[p] * (oi + oj));
        } else if (k < K32) {
          mp32[(size_t)p * K32 + k] = 0.f;
        }
      }
    }
  }
}

// ----------------------------------------------- kernel derivatives ----

// exact: G^(k,l) for one (dx, dy) by evaluating the numerator polynomials
// N_{k,l} (computed on the host, kernel_polys) over R^{1+2(k+l)}, each term
// c * x^a * y^b with powi by binary exponentiation as LLVM lowers Rust's
// f64::powi (galax-core/src/lib.rs:1132-1148). One lane per (k, l).
__device__ __forceinline__ double powi_dev(double a, int b) {
#pragma clang fp contract(off)
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// polys layout: [(k*stride + l) * stride*stride + a*stride + b], stride = 2p+1
__global__ void k_derivs_poly(const double* __restrict__ polys, int p, const double* __restrict__ dxy, uint32_t nk,
                              double eps, double* __restrict__ gtab) {
#pragma clang fp contract(off)
  const int stride = 2 * p + 1, GS = stride * stride;
  uint32_t key = blockIdx.x;
  if (key >= nk) return;
  const double x = dxy[2 * key], y = dxy[2 * key + 1];
  const double r2 = x * x + y * y + eps * eps;
  const double r = sqrt(r2);
  for (int t = threadIdx.x; t < GS; t += blockDim.x) {
    int k = t / stride, l = t % stride;
    double out = 0.0;
    if (k + l <= 2 * p) {
      const double* c = polys + (size_t)(k * stride + l) * GS;
      double val = 0.0;
      for (int a = 0; a < stride; ++a)
        for (int b = 0; b < stride - a; ++b) {
          double cc = c[a * stride + b];
          if (cc != 0.0) val += cc * powi_dev(x, a) * powi_dev(y, b);
        }
      // inv_r[1] = 1/r; inv_r[n] = inv_r[n-2] / r2 (galax-core/src/lib.rs:1053-1057)
      double ir = 1.0 / r;
      for (int q = 0; q < k + l; ++q) ir = ir / r2;
      out = val * ir;
      gtab[(size_t)key * GS + deriv_index(k, l, p)] = out;
    }
  }
}

// ----------------------------------------------- M2L key cache (setup) ----
// The reference caches derivative tables in a map keyed by
// (dx as i64, dy as i64) of the COM offset target - source; with one thread the
// first pair in (target id, list position) order fills each entry and every
// later pair with that key reuses it (galax-core/src/lib.rs:1209-1247). The
// tree is static, so the keys and tables are computed once per tree.

__device__ __forceinline__ int64_t rust_f64_as_i64(double v) {
  if (isnan(v)) return 0;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ __forceinline__ bool pack_key(double dx, double dy, uint64_t* out) {
  int64_t kx = rust_f64_as_i64(dx), ky = rust_f64_as_i64(dy);
  const int64_t lim = 2147483647ll;
  if (kx <= -lim || kx >= lim || ky <= -lim || ky >= lim) return false;
  *out = ((uint64_t)(uint32_t)(kx + 2147483648ll) << 32) | (uint64_t)(uint32_t)(ky + 2147483648ll);
  return true;
}

// target of each live pair (CSR row index)
__global__ void k_pair_targets(const uint32_t* __restrict__ off, uint32_t n_targets, uint32_t* __restrict__ tgt_of_pair) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_targets) return;
  for (uint32_t e = off[t]; e < off[t + 1]; ++e) tgt_of_pair[e] = t;
}

// Global table: insert the key (if new) and lower its representative to e.
// The representative of a key is its smallest pair index (first touch in
// list order). A slot already holding the key and an index <= e needs no
// atomic: slots only ever take a key once and indices only decrease, so a
// read that shows both is final. The probes are ordinary cached loads; a
// stale read is safe (an empty slot sends the pair to the CAS, which returns
// the key a concurrent insert left there; a stale index is only larger).
__device__ __forceinline__ void keys_global_insert(unsigned long long* hkeys, uint32_t* hrep, uint64_t mask,
                                                   uint64_t key, uint32_t e) {
  uint64_t slot = mix64(key) & mask;
  for (;;) {
    const unsigned long long cur = hkeys[slot];
    if (cur == key) {
      if (hrep[slot] > e) atomicMin(&hrep[slot], e);
      return;
    }
    if (cur == EMPTY_KEY) {
      const unsigned long long prev = atomicCAS(&hkeys[slot], (unsigned long long)EMPTY_KEY, (unsigned long long)key);
      if (prev == EMPTY_KEY || prev == key) {
        atomicMin(&hrep[slot], e);
        return;
      }
    }
    slot = (slot + 1) & mask;
  }
}

// A few thousand keys are shared by millions of pairs, and device-scope
// atomics on the same few lines serialize in memory (an XCD's L2 keeps its
// stale copy of a line that another XCD's atomic changed, so the read-first
// filter above does not help across XCDs). Each block therefore first
// reduces its KI_PAIRS consecutive pairs to their distinct keys with the
// smallest index each, in an LDS hash, and only those go to the global table.
// A pair whose key finds no LDS slot within KI_PROBES probes goes directly.
constexpr uint32_t KI_PAIRS = 32768, KI_SLOTS = 4096, KI_PROBES = 64;

__global__ void __launch_bounds__(1024) k_keys_insert(const uint32_t* __restrict__ dat,
                                                      const uint32_t* __restrict__ tgt_of_pair, uint32_t n_pairs,
                                                      const double* __restrict__ comx, const double* __restrict__ comy,
                                                      unsigned long long* hkeys, uint32_t* hrep, uint64_t mask,
                                                      uint32_t* __restrict__ err) {
  __shared__ unsigned long long lk[KI_SLOTS];
  __shared__ uint32_t le[KI_SLOTS];
  for (uint32_t i = threadIdx.x; i < KI_SLOTS; i += blockDim.x) {
    lk[i] = EMPTY_KEY;
    le[i] = 0xFFFFFFFFu;
  }
  __syncthreads();
  const uint32_t e0 = blockIdx.x * KI_PAIRS, e1 = min(e0 + KI_PAIRS, n_pairs);
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const uint32_t t = tgt_of_pair[e], s = dat[e];
    uint64_t key;
    if (!pack_key(comx[t] - comx[s], comy[t] - comy[s], &key)) {
      atomicOr(err, 1u);
      continue;
    }
    uint32_t ls = (uint32_t)mix64(key) & (KI_SLOTS - 1);
    bool done = false;
    for (uint32_t q = 0; q < KI_PROBES && !done; ++q) {
      unsigned long long cur = lk[ls];
      if (cur == EMPTY_KEY) cur = atomicCAS(&lk[ls], (unsigned long long)EMPTY_KEY, (unsigned long long)key);
      if (cur == EMPTY_KEY || cur == key) {
        atomicMin(&le[ls], e);
        done = true;
      } else {
        ls = (ls + 1) & (KI_SLOTS - 1);
      }
    }
    if (!done) keys_global_insert(hkeys, hrep, mask, key, e);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < KI_SLOTS; i += blockDim.x)
    if (lk[i] != EMPTY_KEY) keys_global_insert(hkeys, hrep, mask, lk[i], le[i]);
}

__global__ void k_keys_flag(const unsigned long long* __restrict__ hkeys, uint64_t nslots, uint32_t* __restrict__ f) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) f[i] = hkeys[i] != EMPTY_KEY ? 1u : 0u;
}

// compact occupied slots; representative offsets (dx, dy) of each key
__global__ void k_keys_compact(const unsigned long long* __restrict__ hkeys, const uint32_t* __restrict__ hrep,
                               uint64_t nslots, const uint32_t* __restrict__ kid, uint32_t* __restrict__ slot_kid,
                               const uint32_t* __restrict__ dat, const uint32_t* __restrict__ tgt_of_pair,
                               const double* __restrict__ comx, const double* __restrict__ comy,
                               double* __restrict__ dxy) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  if (hkeys[i] == EMPTY_KEY) return;
  uint32_t id = kid[i];
  slot_kid[i] = id;
  uint32_t e = hrep[i];
  uint32_t t = tgt_of_pair[e], s = dat[e];
  dxy[2 * id] = comx[t] - comx[s];
  dxy[2 * id + 1] = comy[t] - comy[s];
}

__global__ void k_keys_lookup(const uint32_t* __restrict__ dat, const uint32_t* __restrict__ tgt_of_pair,
                              uint32_t n_pairs, const double* __restrict__ comx, const double* __restrict__ comy,
                              const unsigned long long* __restrict__ hkeys, const uint32_t* __restrict__ slot_kid,
                              uint64_t mask, uint32_t* __restrict__ key_of_pair) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pairs) return;
  uint32_t t = tgt_of_pair[e], s = dat[e];
  uint64_t key;
  if (!pack_key(comx[t] - comx[s], comy[t] - comy[s], &key)) return;
  uint64_t slot = mix64(key) & mask;
  while (hkeys[slot] != key) slot = (slot + 1) & mask;
  key_of_pair[e] = slot_kid[slot];
}

// --------------------------------------------------------------- M2L ----

// exact: one wave per target, one lane per local coefficient (I, J); pairs in
// list order, source terms in plan order (i outer, j inner), zero multipole
// terms skipped, non-finite multipoles read as 0 (the post-M2M scrub,
// galax-core/src/lib.rs:935-937).
template <int P>
__global__ void __launch_bounds__(256) k_m2l_ref(const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                                                 const uint32_t* __restrict__ key_of_pair,
                                                 const double* __restrict__ gtab, uint32_t n_targets,
                                                 const double* __restrict__ mp, double* __restrict__ loc) {
#pragma clang fp contract(off)
  constexpr int S = num_moments(P);
  constexpr int ST = 2 * P + 1, GS = ST * ST;
  const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= n_targets) return;
  for (int k = lane; k < S; k += 64) {
    int I, J;
    coef_ij<P>(k, I, J);
    const double lf = 1.0 / (c_tabs.fact[I] * c_tabs.fact[J]);
    double acc = 0.0;
    for (uint32_t e = off[t]; e < off[t + 1]; ++e) {
      const uint32_t s = dat[e];
      const double* ms = mp + (size_t)s * S;
      const double* G = gtab + (size_t)key_of_pair[e] * GS;
#pragma unroll
      for (int i = 0; i <= P; ++i) {
#pragma unroll
        for (int j = 0; j <= P - i; ++j) {
          double mv = ms[moment_index(i, j)];
          if (!isfinite(mv)) mv = 0.0;
          if (mv == 0.0) continue;
          const double w = ((i + j) % 2 == 0) ? lf : -lf;
          acc += mv * w * G[(i + I + j + J) * ST + (j + J)];
        }
      }
    }
    loc[(size_t)t * S + k] = acc;
  }
}

// Exact per-pair derivatives of G = -1/sqrt(x^2+y^2+eps^2) by the recurrence
// obtained from R^2 dphi/dx = -x phi (phi = 1/R) with Leibniz' rule:
//   R^2 D(k,l) = -(2k-1) x D(k-1,l) - (k-1)^2 D(k-2,l) - 2l y D(k,l-1) - l(l-1) D(k,l-2)   (k >= 1)
//   R^2 D(0,l) = -(2l-1) y D(0,l-1) - (l-1)^2 D(0,l-2)
// with D = d^k/dx^k d^l/dy^l phi and G^(k,l) = -D(k,l). Order n needs orders
// n-1 and n-2 only, so a wave computes one order per step.
template <int P>
__device__ __forceinline__ void derivs_recurrence(double x, double y, double eps2, double* G /* LDS, ST*ST */,
                                                  int lane) {
  constexpr int ST = 2 * P + 1;
  const double r2 = x * x + y * y + eps2;
  const double ir2 = 1.0 / r2;
  if (lane == 0) G[0] = 1.0 / sqrt(r2);  // D(0,0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int n = 1; n <= 2 * P; ++n) {
    for (int k = lane; k <= n; k += 64) {
      int l = n - k;
      double v;
      if (k >= 1) {
        v = -(2.0 * k - 1.0) * x * G[(n - 1) * ST + l];
        if (k >= 2) v -= (double)(k - 1) * (k - 1) * G[(n - 2) * ST + l];
        if (l >= 1) v -= 2.0 * l * y * G[(n - 1) * ST + (l - 1)];
        if (l >= 2) v -= (double)l * (l - 1) * G[(n - 2) * ST + (l - 2)];
      } else {
        v = -(2.0 * l - 1.0) * y * G[(n - 1) * ST + (l - 1)];
        if (l >= 2) v -= (double)(l - 1) * (l - 1) * G[(n - 2) * ST + (l - 2)];
      }
      G[n * ST + l] = v * ir2;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

// CORRECTED variant: exact derivatives per pair (D stored; G = -D).
template <int P>
__global__ void __launch_bounds__(256) k_m2l_exact(const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                                                   uint32_t n_targets, const double* __restrict__ comx,
                                                   const double* __restrict__ comy, double eps2,
                                                   const double* __restrict__ mp, double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  constexpr int ST = 2 * P + 1;
  __shared__ double sG[4][ST * ST];
  const int w = threadIdx.x >> 6;
  const uint32_t t = blockIdx.x * 4 + w;
  const int lane = threadIdx.x & 63;
  if (t >= n_targets) return;
  double* G = sG[w];
  double acc[(S + 63) / 64];
#pragma unroll
  for (int q = 0; q < (S + 63) / 64; ++q) acc[q] = 0.0;
  const double tx = comx[t], ty = comy[t];
  for (uint32_t e = off[t]; e < off[t + 1]; ++e) {
    const uint32_t s = dat[e];
    derivs_recurrence<P>(tx - comx[s], ty - comy[s], eps2, G, lane);
    const double* ms = mp + (size_t)s * S;
#pragma unroll
    for (int q = 0; q < (S + 63) / 64; ++q) {
      int k = lane + 64 * q;
      if (k < S) {
        int I, J;
        coef_ij<P>(k, I, J);
        const double lf = 1.0 / (c_tabs.fact[I] * c_tabs.fact[J]);
        double a = 0.0;
#pragma unroll
        for (int i = 0; i <= P; ++i)
#pragma unroll
          for (int j = 0; j <= P - i; ++j) {
            double mv = ms[moment_index(i, j)];
            if (!isfinite(mv)) mv = 0.0;
            // w * G = sign/(I!J!) * (-D)
            const double sg = ((i + j) % 2 == 0) ? -1.0 : 1.0;
            a += mv * G[(i + I + j + J) * ST + (j + J)] * sg;
          }
        acc[q] += a * lf;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
#pragma unroll
  for (int q = 0; q < (S + 63) / 64; ++q) {
    int k = lane + 64 * q;
    if (k < S) loc[(size_t)t * S + k] = acc[q];
  }
}

// CORRECTED variant, p <= 8: one wave per target, one lane per pair. Each lane
// runs its pair's derivative recurrence order by order (the recurrence of
// derivs_recurrence) and scatters every new entry D(k, l) straight into the S
// local coefficients it feeds (m2l_pair_scatter),
//   L(I, J) += s(i, j) M(i, j) D(i + I, j + J),  i = k - I, j = l - J,
// so only three orders of the table are live and the moments, the locals and
// that window stay in registers. The wave then sums the lanes' locals with a
// halving exchange (lane c ends with coefficient c, a fixed order) and adds
// them over batches of 64 pairs. A pair costs about 4(2p+1)^2 + 2 S^2
// operations with no LDS traffic and no serial steps.
template <int P>
__global__ void __launch_bounds__(64) k_m2l_exact_lp(const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat, const uint32_t* __restrict__ order,
                                                     const double* __restrict__ comx, const double* __restrict__ comy,
                                                     double eps2, const double* __restrict__ mp,
                                                     double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  const uint32_t t = order[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t e0 = off[t], e1 = off[t + 1];
  const double tx = comx[t], ty = comy[t];
  double acc = 0.0;  // coefficient `lane`
  for (uint32_t eb = e0; eb < e1; eb += 64) {  // uniform
    const uint32_t e = eb + (uint32_t)lane;
    double L[S];
#pragma unroll
    for (int k = 0; k < S; ++k) L[k] = 0.0;
    if (e < e1) {
      const uint32_t sidx = dat[e];
      const double x = tx - comx[sidx], y = ty - comy[sidx];
      const double r2 = x * x + y * y + eps2;
      const double* ms = mp + (size_t)sidx * S;
      double smv[S];
#pragma unroll
      for (int q = 0; q <= P; ++q)
#pragma unroll
        for (int j = 0; j <= q; ++j) {
          double mv = ms[moment_index(q - j, j)];
          if (!isfinite(mv)) mv = 0.0;
          // w * G = sign/(I!J!) * (-D); the 1/(I!J!) is applied once at the end
          smv[moment_index(q - j, j)] = (q % 2 == 0) ? -mv : mv;
        }
      m2l_pair_scatter<P, double>(x, y, 1.0 / r2, 1.0 / sqrt(r2), smv, L);
    }
    acc += wave_sum_scatter<S>(L, lane);
  }
  if (lane < S) {
    int I, J;
    coef_ij<P>(lane, I, J);
    loc[(size_t)t * S + lane] = acc * (1.0 / (c_tabs.fact[I] * c_tabs.fact[J]));
  }
}

// The same in FP32 for the mixed forms (p <= 8). Each pair is rescaled by
// 2^a, a power of two near its distance: x' = x 2^-a, so the table is
// D' = D 2^{a(n+1)} (the recurrence is homogeneous), the multipole enters as
// M' = M 2^{-a(i+j)} and the pair's locals leave as L' 2^{-a(I+J+1)}, all
// exact scalings; every FP32 quantity is then of moderate size. The lanes'
// locals are summed in FP64.
template <int P>
__global__ void __launch_bounds__(64) k_m2l_exact_lp32(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ dat, const uint32_t* __restrict__ order,
                                                       const double* __restrict__ comx,
                                                       const double* __restrict__ comy, double eps2,
                                                       const double* __restrict__ mp, double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  const uint32_t t = order[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t e0 = off[t], e1 = off[t + 1];
  const double tx = comx[t], ty = comy[t];
  double acc = 0.0;  // coefficient `lane`
  for (uint32_t eb = e0; eb < e1; eb += 64) {  // uniform
    const uint32_t e = eb + (uint32_t)lane;
    float L[S];
#pragma unroll
    for (int k = 0; k < S; ++k) L[k] = 0.f;
    int a = 0;
    if (e < e1) {
      const uint32_t sidx = dat[e];
      const double x = tx - comx[sidx], y = ty - comy[sidx];
      const double r2 = x * x + y * y + eps2;
      // 2^a within a factor 2 of the distance; a zero or non-finite distance
      // keeps a = 0 and gives non-finite locals, as the FP64 kernel does
      if (r2 > 0.0 && isfinite(r2)) a = ilogb(r2) / 2;
      const double r2s = ldexp(r2, -2 * a);
      const double* ms = mp + (size_t)sidx * S;
      float smv[S];
#pragma unroll
      for (int q = 0; q <= P; ++q)
#pragma unroll
        for (int j = 0; j <= q; ++j) {
          double mv = ms[moment_index(q - j, j)];
          if (!isfinite(mv)) mv = 0.0;
          smv[moment_index(q - j, j)] = (float)ldexp((q % 2 == 0) ? -mv : mv, -a * q);
        }
      m2l_pair_scatter<P, float>((float)ldexp(x, -a), (float)ldexp(y, -a), (float)(1.0 / r2s),
                                 (float)(1.0 / sqrt(r2s)), smv, L);
    }
    double Ld[S];
#pragma unroll
    for (int q = 0; q <= P; ++q)
#pragma unroll
      for (int j = 0; j <= q; ++j) Ld[moment_index(q - j, j)] = ldexp((double)L[moment_index(q - j, j)], -a * (q + 1));
    acc += wave_sum_scatter<S>(Ld, lane);
  }
  if (lane < S) {
    int I, J;
    coef_ij<P>(lane, I, J);
    loc[(size_t)t * S + lane] = acc * (1.0 / (c_tabs.fact[I] * c_tabs.fact[J]));
  }
}

// --------------------------------------------------------------- L2L ----

// exact: L_child[oi,oj] += sum_{i>=oi, j>=oj, i+j<=p} L_par[i,j] C(i,oi) C(j,oj) dx^(i-oi) dy^(j-oj)
// (galax-core/src/lib.rs:1274-1299). l2l_sum is the sum for one coefficient of
// one child at offset (dx, dy) from its parent, in the reference's order.
template <int P>
__device__ __forceinline__ double l2l_sum(const double* __restrict__ lp, const double* __restrict__ bt, int ld,
                                          int oi, int oj, double dx, double dy) {
#pragma clang fp contract(off)
  // rolled loops keep the register count low where this runs inside the
  // L2P+P2P kernel; px, py are dx^(i-oi), dy^(j-oj) by repeated products;
  // bt[n * ld + k] is the binomial C(n, k)
  double sum = 0.0, px = 1.0;
#pragma unroll 1
  for (int i = oi; i <= P; ++i) {
    const double bi = bt[i * ld + oi];
    double py = 1.0;
#pragma unroll 1
    for (int j = oj; j <= P - i; ++j) {
      sum += lp[moment_index(i, j)] * bi * bt[j * ld + oj] * px * py;
      py = py * dy;
    }
    px = px * dx;
  }
  return sum;
}

// children in [lo, hi), one thread per (child, coefficient)
template <int P>
__global__ void k_l2l(const uint32_t* __restrict__ par, const double* __restrict__ comx,
                      const double* __restrict__ comy, uint32_t lo, uint32_t hi, double* __restrict__ loc) {
#pragma clang fp contract(off)
  constexpr int S = num_moments(P);
  size_t tt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= (size_t)(hi - lo) * S) return;
  uint32_t c = lo + (uint32_t)(tt / S);
  int k = (int)(tt % S);
  uint32_t p = par[c];
  if (p == NONE) return;
  int oi, oj;
  coef_ij<P>(k, oi, oj);
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  loc[(size_t)c * S + k] += l2l_sum<P>(loc + (size_t)p * S, &c_tabs.binom[0][0], MAX_P + 1, oi, oj, dx, dy);
}

// CORRECTED variant, p <= 8: the all-level M2M and L2L, one thread per node.
// The thread keeps the source expansion and the offset's powers in
// registers and applies the shift with its binomial and factorial constants
// known at compile time: no LDS, no barriers, no table loads, about 2 S^2 / 4
// operations per shift. The sums are the same as k_m2m's and l2l_sum's, in
// another order and with multiplications by 1/a! instead of divisions by a!.

// parents in [lo, hi): M_p(oi, oj) = sum over children c and i <= oi, j <= oj
// of M_c(i, j) dx^(oi-i) dy^(oj-j) / ((oi-i)! (oj-j)!), (dx, dy) = com_c - com_p
// Blocks of 64 threads spread a level's waves over every CU; the register
// budget (up to 256 per lane) holds the child's expansion, the sums and the
// powers without spills (at the default budget of 128, p = 8 spilled 1 KB per
// lane: 240 -> 166 us per evaluation at 2^20 Plummer bodies).
template <int P>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_m2m_nodes(const uint32_t* __restrict__ cf, const uint8_t* __restrict__ cc,
                                                   const double* __restrict__ comx, const double* __restrict__ comy,
                                                   uint32_t lo, uint32_t hi, double* __restrict__ mp) {
  constexpr int S = num_moments(P);
  const uint32_t p = lo + blockIdx.x * 64u + threadIdx.x;
  if (p >= hi) return;
  const uint32_t c0 = cf[p], nc = cc[p];
  const double px = comx[p], py = comy[p];
  double acc[S];
#pragma unroll
  for (int k = 0; k < S; ++k) acc[k] = 0.0;
  for (uint32_t ci = 0; ci < nc; ++ci) {
    const uint32_t c = c0 + ci;
    const double dx = comx[c] - px, dy = comy[c] - py;
    double tx[P + 1], ty[P + 1];  // dx^a / a!, dy^b / b!
    tx[0] = ty[0] = 1.0;
#pragma unroll
    for (int a = 1; a <= P; ++a) {
      tx[a] = tx[a - 1] * dx;
      ty[a] = ty[a - 1] * dy;
    }
    static_for<2, P + 1>([&](auto ac) {
      constexpr int a = decltype(ac)::value;
      tx[a] *= 1.0 / ct_fact(a);
      ty[a] *= 1.0 / ct_fact(a);
    });
    const double* mc = mp + (size_t)c * S;
    double m[S];
#pragma unroll
    for (int k = 0; k < S; ++k) m[k] = mc[k];
    static_for<0, P + 1>([&](auto nc_) {
      constexpr int n = decltype(nc_)::value;
      static_for<0, n + 1>([&](auto jc) {
        constexpr int oj = decltype(jc)::value, oi = n - oj;
        double sum = acc[moment_index(oi, oj)];
        static_for<0, oi + 1>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          static_for<0, oj + 1>([&](auto jjc) {
            constexpr int j = decltype(jjc)::value;
            sum = fma(m[moment_index(i, j)], tx[oi - i] * ty[oj - j], sum);
          });
        });
        acc[moment_index(oi, oj)] = sum;
      });
    });
  }
  double* out = mp + (size_t)p * S;
#pragma unroll
  for (int k = 0; k < S; ++k) out[k] = isfinite(acc[k]) ? acc[k] : 0.0;  // non-finite -> 0 (galax-core/src/lib.rs:934-936)
}

// Corrected variant, levels whose parents each have one child (Morton-chunk
// trees have runs of them: at 2^20 Plummer bodies levels 1-7 hold one parent
// per leaf). A thread per (node, coefficient). A child at its parent's COM,
// the usual case (the same bodies), shifts by zero: the parent's multipole is
// the child's and the child's local gains the parent's, which is what the
// full shifts of k_m2m_nodes / k_l2l_nodes give bit for bit (every other term
// is +-0). An off-COM child takes its coefficient's full shift.
template <int P>
__global__ void __launch_bounds__(256) k_m2m_single(const uint32_t* __restrict__ cf, const double* __restrict__ comx,
                                                    const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                    double* __restrict__ mp) {
  constexpr int S = num_moments(P);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t p = lo + (uint32_t)(t / S);
  const int k = (int)(t % S);
  if (p >= hi) return;
  const uint32_t c = cf[p];
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  const double* mc = mp + (size_t)c * S;
  double v;
  if (dx == 0.0 && dy == 0.0) {
    v = mc[k];
  } else {
    // M_p(oi, oj) = sum_{a <= oi, b <= oj} M_c(oi - a, oj - b) dx^a / a! dy^b / b!
    int oi, oj;
    coef_ij<P>(k, oi, oj);
    v = 0.0;
    double ta = 1.0;
    for (int a = 0; a <= oi; ++a) {
      if (a > 0) ta *= dx / a;
      double tb = 1.0;
      for (int b = 0; b <= oj; ++b) {
        if (b > 0) tb *= dy / b;
        v = fma(mc[moment_index(oi - a, oj - b)], ta * tb, v);
      }
    }
  }
  mp[(size_t)p * S + k] = isfinite(v) ? v : 0.0;  // non-finite -> 0 (galax-core/src/lib.rs:934-936)
}

// children in [lo, hi), each its parent's only child
template <int P>
__global__ void __launch_bounds__(256) k_l2l_single(const uint32_t* __restrict__ par, const double* __restrict__ comx,
                                                    const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                    double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = lo + (uint32_t)(t / S);
  const int k = (int)(t % S);
  if (c >= hi) return;
  const uint32_t p = par[c];
  if (p == NONE) return;
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  const double* lp = loc + (size_t)p * S;
  double v;
  if (dx == 0.0 && dy == 0.0) {
    v = lp[k];
  } else {
    // L_c(oi, oj) += sum over i >= oi, j >= oj of L_p(i, j) C(i, oi) C(j, oj) dx^(i-oi) dy^(j-oj)
    int oi, oj;
    coef_ij<P>(k, oi, oj);
    v = 0.0;
    double ea = 1.0;
    for (int i = oi; i <= P - oj; ++i) {
      double eb = 1.0;
      for (int j = oj; i + j <= P; ++j) {
        const double bc = c_tabs.fact[i] / (c_tabs.fact[oi] * c_tabs.fact[i - oi]) * c_tabs.fact[j] /
                          (c_tabs.fact[oj] * c_tabs.fact[j - oj]);
        v = fma(lp[moment_index(i, j)] * bc, ea * eb, v);
        eb *= dy;
      }
      ea *= dx;
    }
  }
  loc[(size_t)c * S + k] += v;
}

// Chains of one-child levels in one launch (corrected variant): levels
// where the parent lo + i of each level has the child lo - n + i (n nodes per
// level; checked once per tree by k_chain_check). A block holds 256 / S
// chains, a thread per (chain, coefficient); the expansion of the chain's
// current node stays in LDS between levels, and each level's terms are those
// of k_m2m_single / k_l2l_single (a copy for a child at its parent's COM).
__global__ void k_chain_check(const uint32_t* __restrict__ cf, uint32_t lo, uint32_t n, uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && cf[lo + i] != lo - n + i) atomicOr(bad, 1u);
}

// M2M up nlev levels: parents lo0 + l n + i, l = 0 .. nlev - 1
template <int P>
__global__ void __launch_bounds__(256) k_m2m_chain(const double* __restrict__ comx, const double* __restrict__ comy,
                                                   uint32_t lo0, uint32_t n, uint32_t nlev, double* __restrict__ mp) {
  constexpr int S = num_moments(P), NB = 256 / S;
  __shared__ double sm[NB][S];
  const int ns = (int)threadIdx.x / S, k = (int)threadIdx.x % S;
  const uint32_t i = blockIdx.x * NB + (uint32_t)ns;
  const bool live = ns < NB && i < n;  // every thread reaches the barriers
  int oi = 0, oj = 0;
  if (live) {
    coef_ij<P>(k, oi, oj);
    sm[ns][k] = mp[(size_t)(lo0 - n + i) * S + k];
  }
  __syncthreads();
  for (uint32_t l = 0; l < nlev; ++l) {
    const uint32_t p = lo0 + l * n + i, c = p - n;
    double v = 0.0;
    if (live) {
      const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
      if (dx == 0.0 && dy == 0.0) {
        v = sm[ns][k];
      } else {  // M_p(oi, oj) = sum_{a <= oi, b <= oj} M_c(oi - a, oj - b) dx^a / a! dy^b / b!
        double ta = 1.0;
        for (int a = 0; a <= oi; ++a) {
          if (a > 0) ta *= dx / a;
          double tb = 1.0;
          for (int b = 0; b <= oj; ++b) {
            if (b > 0) tb *= dy / b;
            v = fma(sm[ns][moment_index(oi - a, oj - b)], ta * tb, v);
          }
        }
      }
      v = isfinite(v) ? v : 0.0;  // non-finite -> 0 (galax-core/src/lib.rs:934-936)
    }
    __syncthreads();
    if (live) {
      sm[ns][k] = v;
      mp[(size_t)p * S + k] = v;
    }
    __syncthreads();
  }
}

// L2L down nlev levels: children hi0 - l n + i - n of parents hi0 - l n + i,
// l = 0 .. nlev - 1 (hi0: the first level of the top parents)
template <int P>
__global__ void __launch_bounds__(256) k_l2l_chain(const double* __restrict__ comx, const double* __restrict__ comy,
                                                   uint32_t hi0, uint32_t n, uint32_t nlev, double* __restrict__ loc) {
  constexpr int S = num_moments(P), NB = 256 / S;
  __shared__ double sl[NB][S];
  const int ns = (int)threadIdx.x / S, k = (int)threadIdx.x % S;
  const uint32_t i = blockIdx.x * NB + (uint32_t)ns;
  const bool live = ns < NB && i < n;  // every thread reaches the barriers
  int oi = 0, oj = 0;
  if (live) {
    coef_ij<P>(k, oi, oj);
    sl[ns][k] = loc[(size_t)(hi0 + i) * S + k];
  }
  __syncthreads();
  for (uint32_t l = 0; l < nlev; ++l) {
    const uint32_t p = hi0 - l * n + i, c = p - n;
    double v = 0.0;
    if (live) {
      const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
      double w;
      if (dx == 0.0 && dy == 0.0) {
        w = sl[ns][k];
      } else {  // sum over i >= oi, j >= oj of L_p(i, j) C(i, oi) C(j, oj) dx^(i-oi) dy^(j-oj)
        w = 0.0;
        double ea = 1.0;
        for (int a = oi; a <= P - oj; ++a) {
          double eb = 1.0;
          for (int b = oj; a + b <= P; ++b) {
            const double bc = c_tabs.fact[a] / (c_tabs.fact[oi] * c_tabs.fact[a - oi]) * c_tabs.fact[b] /
                              (c_tabs.fact[oj] * c_tabs.fact[b - oj]);
            w = fma(sl[ns][moment_index(a, b)] * bc, ea * eb, w);
            eb *= dy;
          }
          ea *= dx;
        }
      }
      v = loc[(size_t)c * S + k] + w;
    }
    __syncthreads();
    if (live) {
      sl[ns][k] = v;
      loc[(size_t)c * S + k] = v;
    }
    __syncthreads();
  }
}

// children in [lo, hi): L_c(oi, oj) += sum over i >= oi, j >= oj, i + j <= p
// of L_p(i, j) C(i, oi) C(j, oj) dx^(i-oi) dy^(j-oj), (dx, dy) = com_c - com_p
template <int P>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_l2l_nodes(const uint32_t* __restrict__ par, const double* __restrict__ comx,
                                                   const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                   double* __restrict__ loc) {
  constexpr int S = num_moments(P);
  const uint32_t c = lo + blockIdx.x * 64u + threadIdx.x;
  if (c >= hi) return;
  const uint32_t p = par[c];
  if (p == NONE) return;
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  double ex[P + 1], ey[P + 1];
  ex[0] = ey[0] = 1.0;
#pragma unroll
  for (int a = 1; a <= P; ++a) {
    ex[a] = ex[a - 1] * dx;
    ey[a] = ey[a - 1] * dy;
  }
  const double* lp = loc + (size_t)p * S;
  double l[S];
#pragma unroll
  for (int k = 0; k < S; ++k) l[k] = lp[k];
  double* lc = loc + (size_t)c * S;
  static_for<0, P + 1>([&](auto nc_) {
    constexpr int n = decltype(nc_)::value;
    static_for<0, n + 1>([&](auto jc) {
      constexpr int oj = decltype(jc)::value, oi = n - oj;
      double sum = 0.0;
      static_for<oi, P + 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        static_for<oj, P - i + 1>([&](auto jjc) {
          constexpr int j = decltype(jjc)::value;
          constexpr double bc = ct_binom(i, oi) * ct_binom(j, oj);
          sum = fma(l[moment_index(i, j)] * bc, ex[i - oi] * ey[j - oj], sum);
        });
      });
      lc[moment_index(oi, oj)] += sum;
    });
  });
}

// ------------------------------------------------------ L2P + P2P ----

// exact far field of one body (galax-core/src/lib.rs:1321-1345). ROWS: the
// rows i of L are walked by a rolled loop (the powers of dux by repeated
// products, as in the table), so that at most one row of coefficients is in
// registers at a time; same terms in the same order.
template <int P, bool ROWS = false>
__device__ __forceinline__ void l2p_body(const double* __restrict__ L, double dux, double duy, double* fx,
                                         double* fy, double* fphi) {
#pragma clang fp contract(off)
  double xp[P + 1], yp[P + 1];
  xp[0] = 1.0;
  yp[0] = 1.0;
#pragma unroll
  for (int q = 1; q <= P; ++q) {
    xp[q] = xp[q - 1] * dux;
    yp[q] = yp[q - 1] * duy;
  }
  double ax = 0.0, ay = 0.0, ph = 0.0;
  if (ROWS) {
    double xpi = 1.0, xpm = 1.0;  // dux^i, dux^(i-1)
#pragma unroll 1
    for (int i = 0; i <= P; ++i) {
#pragma unroll
      for (int j = 0; j <= P; ++j) {
        if (j <= P - i) {
          double c = L[moment_index(i, j)];
          if (i >= 1) ax += c * (double)i * xpm * yp[j];
          if (j >= 1) ay += c * (double)j * xpi * yp[j - 1];
          ph += c * xpi * yp[j];
        }
      }
      xpm = xpi;
      xpi = xpi * dux;
    }
  } else {
#pragma unroll
    for (int i = 0; i <= P; ++i)
#pragma unroll
      for (int j = 0; j <= P - i; ++j) {
        double c = L[moment_index(i, j)];
        if (i >= 1) ax += c * (double)i * xp[i - 1] * yp[j];
        if (j >= 1) ay += c * (double)j * xp[i] * yp[j - 1];
        ph += c * xp[i] * yp[j];
      }
  }
  *fx = ax;
  *fy = ay;
  *fphi = ph;
}

// Far field of one body by nested Horner schemes (FP64): for each row i,
// r_i(y) = sum_j L(i, j) y^j and its y-derivative, then over x the value
// sum_i r_i x^i and both derivatives; 2 S + 3 (P + 1) FMAs instead of the
// reference-ordered sum of l2p_body (~470 VALU instructions at p = 6). The
// FP32 forms' near field uses it; the FP64 forms keep l2p_body's order.
template <int P>
__device__ __forceinline__ void l2p_horner(const double* __restrict__ L, double x, double y, double* fx, double* fy,
                                           double* fphi) {
  double r[P + 1], dr[P + 1];
  static_for<0, P + 1>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    double v = L[moment_index(i, P - i)], d = 0.0;
    static_for<1, P - i + 1>([&](auto jc) {
      constexpr int j = P - i - decltype(jc)::value;
      d = fma(d, y, v);
      v = fma(v, y, L[moment_index(i, j)]);
    });
    r[i] = v;
    dr[i] = d;
  });
  double ph = r[P], gx = 0.0, gy = dr[P];
#pragma unroll
  for (int i = P - 1; i >= 0; --i) {
    gx = fma(gx, x, ph);
    ph = fma(ph, x, r[i]);
    gy = fma(gy, x, dr[i]);
  }
  *fx = gx;
  *fy = gy;
  *fphi = ph;
}

// Near-field sum of one target body over the staged sources half, half + 2,
// ... < n (n even: the chunk is padded with a far zero-mass source). The loop
// index is the same for both half-waves, so it lives in scalar registers.
// MASK: the chunk holds the target's own leaf; slot `self` is the body itself
// and contributes nothing (its inverse distance is forced to 0).
template <typename T, bool PHI, bool MASK>
__device__ __forceinline__ void near_sum(const typename Near<T>::V4* __restrict__ src, int half, int n, int self,
                                         T xr, T yr, T eps2, T& accx, T& accy, T& accp) {
  const typename Near<T>::V4* sh = src + half;
  constexpr int U = sizeof(T) == 8 ? 2 : 4;
#pragma unroll U
  for (int jj = 0; jj < n; jj += 2) {
    const typename Near<T>::V4 sj = sh[jj];
    const T dx = sj.x - xr, dy = sj.y - yr;
    const T r2 = fma(dx, dx, fma(dy, dy, eps2));
    T inv = Near<T>::rsq(r2);
    if (MASK) inv = jj + half == self ? (T)0 : inv;
    const T f = sj.z * (inv * inv * inv);
    accx = fma(f, dx, accx);
    accy = fma(f, dy, accy);
    if (PHI) accp = fma(-sj.z, inv, accp);
  }
}

// One 64-lane wave per work item (a target leaf and at most P2P_ITEM_LEAVES of
// its gather list). The item's sources are staged through LDS in chunks of
// P2P_CHUNK bodies as (x, y) relative to the target leaf's COM plus m, in T
// precision. Lanes 0-31 and 32-63 hold the same 32 target bodies and take the
// even / odd sources of a chunk; the halves are combined by one xor-shuffle.
// Only the chunks that hold the target leaf itself (found among the item's
// source leaves by ballot) run with the self-exclusion mask. The first item of
// a target also adds the far field in FP64: the leaf's local expansion at each
// body and, with fuse_l2l, the parent's local about the parent's COM (in place
// of the last L2L level). Targets with one item write the result; otherwise
// each item writes a partial to its slot.
constexpr int P2P_CHUNK = 256;
template <int P, typename T, bool PHI>
__global__ void __launch_bounds__(64) k_l2p_p2p(const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ m, const uint32_t* __restrict__ bs,
                                                const uint32_t* __restrict__ be, const double* __restrict__ comx,
                                                const double* __restrict__ comy, const double* __restrict__ loc,
                                                const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gdat,
                                                const uint32_t* __restrict__ itgt, const uint32_t* __restrict__ ig0,
                                                const uint32_t* __restrict__ ig1, const uint32_t* __restrict__ islot,
                                                uint32_t n_max, double eps2d, double* __restrict__ ax,
                                                double* __restrict__ ay, double* __restrict__ phi,
                                                double* __restrict__ part, const uint32_t* __restrict__ par,
                                                int fuse_l2l) {
  constexpr int S = num_moments(P);
  typedef typename Near<T>::V4 V4;
  __shared__ double sloc[S], spar[S];
  static_assert(P2P_ITEM_LEAVES <= 64, "one lane per source leaf of an item");
  __shared__ V4 src[P2P_CHUNK + 1];
  const uint32_t item = xcd_runs(blockIdx.x, gridDim.x);
  const uint32_t leaf = itgt[item], g0 = ig0[item], g1 = ig1[item], slot = islot[item];
  const bool first = g0 == goff[leaf];
  const int lane = threadIdx.x;
  const int tl = lane & 31, half = lane >> 5;
  const uint32_t ts = bs[leaf], te = be[leaf];
  const double ox = comx[leaf], oy = comy[leaf];
  const T eps2 = (T)eps2d;
  // source leaves of the item: start and offset in the item's source order
  const int nlv = (int)(g1 - g0);
  uint32_t cnt = 0, ss = 0;
  bool is_self = false;
  if (lane < nlv) {
    const uint32_t sl = gdat[g0 + lane];
    ss = bs[sl];
    cnt = be[sl] - ss;
    is_self = sl == leaf;
  }
  uint32_t inc = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(inc, d, 64);
    if (lane >= d) inc += v;
  }
  const uint32_t offs = inc - cnt;  // lane j < nlv: source-order offset of leaf j (its first body: ss)
  const uint32_t total = __builtin_amdgcn_readlane(inc, 63);  // uniform: scalar loop bounds below
  // source-order range [self_lo, self_hi) of the target leaf's own bodies
  const unsigned long long sm_ = __ballot(is_self);
  uint32_t self_lo = 0, self_hi = 0;
  if (sm_) {
    self_lo = __shfl(inc - cnt, __ffsll((long long)sm_) - 1, 64);
    self_hi = self_lo + (te - ts);
  }
  // far field (first item of the leaf only): the leaf's local, staged in LDS,
  // evaluated at each body by lanes 0-31. With fuse_l2l the parent's local
  // is staged too and lanes 32-63 evaluate it about the parent's COM: L2L is
  // an exact re-expansion of the parent's polynomial, so the sum of the two
  // equals L2P of the shifted local (up to rounding), and the separate L2L
  // of the last level is not needed.
  uint32_t pa = NONE;
  double pcx = 0.0, pcy = 0.0;
  if (first) {  // uniform over the block
    pa = fuse_l2l ? par[leaf] : NONE;
    if (pa != NONE) {
      pcx = comx[pa];
      pcy = comy[pa];
    }
    for (int k = lane; k < S; k += 64) {
      sloc[k] = loc[(size_t)leaf * S + k];
      if (pa != NONE) spar[k] = loc[(size_t)pa * S + k];
    }
    __syncthreads();
  }
  const bool far_lane = first && (half == 0 || pa != NONE);
  const double fcx = half ? pcx : ox, fcy = half ? pcy : oy;
  const double* far_loc = half ? spar : sloc;
  for (uint32_t tb = ts; tb < te; tb += 32) {
    const uint32_t i = tb + tl;
    const bool valid = i < te;
    const double xi = valid ? x[i] : ox, yi = valid ? y[i] : oy;
    const T xr = (T)(xi - ox), yr = (T)(yi - oy);
    double fx = 0.0, fy = 0.0, fp = 0.0;
    if (far_lane && valid) l2p_body<P, true>(far_loc, xi - fcx, yi - fcy, &fx, &fy, &fp);
    T accx = 0, accy = 0, accp = 0;
    uint32_t jcur = 0;
    for (uint32_t c0 = 0; c0 < total; c0 += P2P_CHUNK) {
      const int n = (int)min((uint32_t)P2P_CHUNK, total - c0);
      __syncthreads();
      for (int w0 = 0; w0 < n; w0 += 64) {  // uniform: 64-element windows
        const uint32_t g = item_source(c0 + (uint32_t)w0, lane, nlv, offs, ss, jcur);
        const int pp = w0 + lane;
        if (pp < n) {
          V4 v;
          v.x = (T)(x[g] - ox);
          v.y = (T)(y[g] - oy);
          v.z = (T)m[g];
          v.w = (T)0;
          src[pp] = v;
        }
      }
      if (lane == 0 && (n & 1)) {  // pad to even: zero mass far away (r^2 overflows, 1/r = 0)
        V4 v;
        v.x = (T)3e19;
        v.y = (T)3e19;
        v.z = (T)0;
        v.w = (T)0;
        src[n] = v;
      }
      __syncthreads();
      // every lane runs the sum (lanes past the leaf's end sit at its COM and
      // their sums are dropped), so the loop needs no exec mask
      const int ne = (n + 1) & ~1;
      if (c0 < self_hi && c0 + (uint32_t)n > self_lo)
        near_sum<T, PHI, true>(src, half, ne, (int)(self_lo + (i - ts)) - (int)c0, xr, yr, eps2, accx, accy,
                               accp);
      else
        near_sum<T, PHI, false>(src, half, ne, 0, xr, yr, eps2, accx, accy, accp);
    }
    accx += __shfl_xor(accx, 32, 64);
    accy += __shfl_xor(accy, 32, 64);
    if (PHI) accp += __shfl_xor(accp, 32, 64);
    if (pa != NONE) {  // uniform: add the parent's far field from lanes 32-63
      fx += __shfl_xor(fx, 32, 64);
      fy += __shfl_xor(fy, 32, 64);
      if (PHI) fp += __shfl_xor(fp, 32, 64);
    }
    if (valid && half == 0) {
      const double rx = (0.0 - fx) + (double)accx;
      const double ry = (0.0 - fy) + (double)accy;
      const double rp = fp + (double)accp;
      if (slot == NONE) {
        ax[i] = rx;
        ay[i] = ry;
        if (PHI) phi[i] = rp;
      } else {
        double* pp = part + ((size_t)slot * n_max + (i - ts)) * 3;
        pp[0] = rx;
        pp[1] = ry;
        pp[2] = rp;
      }
    }
  }
}

// Targets split over several items: sum their partials. 256 threads = 8
// stripes x 32 bodies; stripe s sums items s, s+8, ... in order, then the 8
// stripe sums are added in stripe order (a fixed order: deterministic).
__global__ void __launch_bounds__(256) k_p2p_combine(const uint32_t* __restrict__ multi, uint32_t n_multi,
                                                     const uint32_t* __restrict__ goff,
                                                     const uint32_t* __restrict__ part_off,
                                                     const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                                     uint32_t n_max, const double* __restrict__ part,
                                                     double* __restrict__ ax, double* __restrict__ ay,
                                                     double* __restrict__ phi) {
  __shared__ double red[8][32][3];
  const uint32_t q = blockIdx.x;
  const uint32_t t = multi[q];
  const uint32_t nit = (goff[t + 1] - goff[t] + P2P_ITEM_LEAVES - 1) / P2P_ITEM_LEAVES;
  const uint32_t po = part_off[t], b0 = bs[t], b1 = be[t];
  const int bl = threadIdx.x & 31, stripe = threadIdx.x >> 5;
  for (uint32_t tb = b0; tb < b1; tb += 32) {
    const uint32_t b = tb + bl;
    double sx = 0.0, sy = 0.0, sp = 0.0;
    if (b < b1) {
      for (uint32_t k = stripe; k < nit; k += 8) {
        const double* pp = part + ((size_t)(po + k) * n_max + (b - b0)) * 3;
        sx += pp[0];
        sy += pp[1];
        sp += pp[2];
      }
    }
    red[stripe][bl][0] = sx;
    red[stripe][bl][1] = sy;
    red[stripe][bl][2] = sp;
    __syncthreads();
    if (stripe == 0 && b < b1) {
      double rx = 0.0, ry = 0.0, rp = 0.0;
      for (int s8 = 0; s8 < 8; ++s8) {
        rx += red[s8][bl][0];
        ry += red[s8][bl][1];
        rp += red[s8][bl][2];
      }
      ax[b] = rx;
      ay[b] = ry;
      if (phi) phi[b] = rp;
    }
    __syncthreads();
  }
}

hipError_t p2p_combine(const P2PItems& it, const uint32_t* goff, const uint32_t* bs, const uint32_t* be,
                       uint32_t n_max, const double* part, double* o0, double* o1, double* o2, hipStream_t s) {
  if (!it.n_multi) return hipSuccess;
  k_p2p_combine<<<(unsigned)it.n_multi, 256, 0, s>>>(it.multi.as<uint32_t>(), (uint32_t)it.n_multi, goff,
                                                     it.part_off.as<uint32_t>(), bs, be, n_max, part, o0, o1, o2);
  return hipGetLastError();
}

// ------------------------------------------------ symmetric near field ----

// One wave per work item: leaf t (target side, "i") with its owned partners
// own_dat[u0 .. u1) ("j"), in FP32 on coordinates relative to t's COM (each
// partner's leaf-relative coordinates plus the FP64 COM difference, rounded
// once). A 32 x 32 leaf pair is four 16 x 16 blocks, one per 16-lane row r:
// row r holds i-bodies 16 (r & 1) + q and j-bodies 16 (r >> 1) + q (q = lane
// within the row). In step K the lane pairs its j with the i of the row lane
// rotated by K, so the i operands arrive through DPP and no LDS is used.
// Per unordered pair: one rsq, the force on j (summed in the lane) and on i
// (summed per rotation K in acc_i[K], over every partner of the item; the 16
// sums are rotated back once at the end). Forces on a partner's bodies go to
// the pair's slot; forces on t's bodies, plus on the first item its self
// block (j-side sums only, every ordered pair of t once, the diagonal masked)
// and its far field in FP64 (as in k_l2p_p2p), go to the item's own slot.
// Padding bodies sit at 3e19 with zero mass: r^2 overflows to inf, 1/r = 0.
// Leaves here are the tree's leaves, or halves of them (vleaf: each half's
// leaf, whose local expansion and parent give the far field; NULL: the
// leaves themselves); bs, be, cx, cy: body range and COM of each.
// Five waves per SIMD for the forces alone (96 VGPRs, a few spilled: the
// kernel is VALU-bound and the extra wave hides its loads; 4% faster than four
// waves); the potential's extra sums would spill heavily at five.
template <int P, bool PHI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PHI ? 1 : 5))) k_p2p_sym(
    const float4* __restrict__ rel, const double* __restrict__ x, const double* __restrict__ y,
    const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be, const double* __restrict__ cx,
    const double* __restrict__ cy, const uint32_t* __restrict__ vleaf, const double* __restrict__ comx,
    const double* __restrict__ comy, const double* __restrict__ loc, const uint32_t* __restrict__ par, int fuse_l2l,
    const uint32_t* __restrict__ own_off, const uint32_t* __restrict__ own_dat, const uint32_t* __restrict__ itgt,
    const uint32_t* __restrict__ iu0, const uint32_t* __restrict__ iu1, const uint32_t* __restrict__ iown,
    float eps2, double2* __restrict__ own2, double* __restrict__ ownp, float2* __restrict__ pair2,
    float* __restrict__ pairp) {
  constexpr int S = num_moments(P);
  constexpr float FAR = 3e19f;
  __shared__ double sloc[S], spar[S];
  const uint32_t item = xcd_runs(blockIdx.x, gridDim.x);
  const uint32_t t = itgt[item], u0 = iu0[item], u1 = iu1[item], own = iown[item];
  const bool first = u0 == own_off[t];
  const int lane = threadIdx.x, r = lane >> 4, q = lane & 15;
  const uint32_t ts = bs[t], na = be[t] - ts;
  const double ox = cx[t], oy = cy[t];
  const uint32_t ia = 16u * (uint32_t)(r & 1) + (uint32_t)q;   // i-layout
  const uint32_t jb = 16u * (uint32_t)(r >> 1) + (uint32_t)q;  // j-layout
  float xi = FAR, yi = FAR, mi = 0.f;
  if (ia < na) {
    const float4 v = rel[ts + ia];
    xi = v.x;
    yi = v.y;
    mi = v.z;
  }
  float aix[16], aiy[16], aip[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) aix[k] = aiy[k] = aip[k] = 0.f;
  // self block: every ordered pair (i, j) of t, sums on the j side
  float sx = 0.f, sy = 0.f, sp = 0.f;
  if (first) {
    float xj = FAR, yj = FAR;
    if (jb < na) {
      const float4 v = rel[ts + jb];
      xj = v.x;
      yj = v.y;
    }
    const bool diag = (r & 1) == (r >> 1);  // rows 0 and 3: i == j at K = 0
    static_for<0, 16>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const float dx = row_rot<K>(xi) - xj, dy = row_rot<K>(yi) - yj;
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2));
      float rs = Near<float>::rsq(r2);
      if constexpr (K == 0) rs = diag ? 0.f : rs;
      const float fmi = ((rs * rs) * rs) * row_rot<K>(mi);
      sx = fmaf(fmi, dx, sx);
      sy = fmaf(fmi, dy, sy);
      if (PHI) sp = fmaf(-row_rot<K>(mi), rs, sp);
    });
  }
  // partner data one step ahead: its loads are in flight during the
  // previous partner's arithmetic
  auto fetch = [&](uint32_t u, uint32_t& nb, float4& v, double& bx, double& by) {
    const uint32_t b = own_dat[u];
    const uint32_t b0 = bs[b];
    nb = be[b] - b0;
    bx = cx[b];
    by = cy[b];
    v = make_float4(FAR, FAR, 0.f, 0.f);
    if (jb < nb) v = rel[b0 + jb];
  };
  auto partner = [&](uint32_t u, uint32_t nb, const float4& v, double cxb, double cyb) {
    float xj = FAR, yj = FAR, mj = 0.f;
    if (jb < nb) {
      xj = v.x + (float)(cxb - ox);
      yj = v.y + (float)(cyb - oy);
      mj = v.z;
    }
    float ajx = 0.f, ajy = 0.f, ajp = 0.f;
    static_for<0, 16>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const float dx = row_rot<K>(xi) - xj, dy = row_rot<K>(yi) - yj;  // x_i - x_j
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2));
      const float rs = Near<float>::rsq(r2);
      const float r3 = (rs * rs) * rs;
      const float fmi = r3 * row_rot<K>(mi), fmj = r3 * mj;
      ajx = fmaf(fmi, dx, ajx);  // on j, towards i
      ajy = fmaf(fmi, dy, ajy);
      aix[K] = fmaf(-fmj, dx, aix[K]);  // on i, towards j
      aiy[K] = fmaf(-fmj, dy, aiy[K]);
      if (PHI) {  // (padding: r^2 = inf and 1/r = 0, so no term with r^2 here)
        ajp = fmaf(-row_rot<K>(mi), rs, ajp);
        aip[K] = fmaf(-mj, rs, aip[K]);
      }
    });
    // rows (0, 1) and (2, 3) hold the same j-bodies
    ajx += __shfl_xor(ajx, 16, 64);
    ajy += __shfl_xor(ajy, 16, 64);
    if (PHI) ajp += __shfl_xor(ajp, 16, 64);
    if ((r & 1) == 0 && jb < nb) {
      const size_t o = (size_t)u * SYM_LEAF + jb;
      pair2[o] = make_float2(ajx, ajy);
      if (PHI) pairp[o] = ajp;
    }
  };
  if constexpr (!PHI) {  // one partner at a time: five waves per SIMD hide its loads
    for (uint32_t u = u0; u < u1; ++u) {
      uint32_t nb = 0;
      float4 v;
      double bx, by;
      fetch(u, nb, v, bx, by);
      partner(u, nb, v, bx, by);
    }
  } else {
    // two partners per iteration with their data in two register sets (A, B),
    // each fetched while the other is computed: no copies between iterations
    uint32_t nbA = 0, nbB = 0;
    float4 vA = make_float4(FAR, FAR, 0.f, 0.f), vB = vA;
    double cxA = 0.0, cyA = 0.0, cxB = 0.0, cyB = 0.0;
    if (u0 < u1) fetch(u0, nbA, vA, cxA, cyA);
    for (uint32_t u = u0; u < u1; u += 2) {
      if (u + 1 < u1) fetch(u + 1, nbB, vB, cxB, cyB);
      partner(u, nbA, vA, cxA, cyA);
      if (u + 1 >= u1) break;
      if (u + 2 < u1) fetch(u + 2, nbA, vA, cxA, cyA);
      partner(u + 1, nbB, vB, cxB, cyB);
    }
  }
  // the i side: acc_i[K] of lane q belongs to the i of its rotated lane
  float tix = 0.f, tiy = 0.f, tip = 0.f;
  static_for<0, 16>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    tix += row_rot<16 - K>(aix[K]);
    tiy += row_rot<16 - K>(aiy[K]);
    if (PHI) tip += row_rot<16 - K>(aip[K]);
  });
  tix += __shfl_xor(tix, 32, 64);  // rows r and r ^ 2 hold the same i-bodies
  tiy += __shfl_xor(tiy, 32, 64);
  if (PHI) tip += __shfl_xor(tip, 32, 64);
  // lanes 0-31 now hold body e = lane on the i side
  double fx = 0.0, fy = 0.0, fp = 0.0;
  if (first) {  // uniform over the block
    sx += __shfl_xor(sx, 16, 64);
    sy += __shfl_xor(sy, 16, 64);
    if (PHI) sp += __shfl_xor(sp, 16, 64);
    const int src = lane < 16 ? lane : lane + 16;  // body 16 + q sits in row 2
    sx = __shfl(sx, src, 64);
    sy = __shfl(sy, src, 64);
    if (PHI) sp = __shfl(sp, src, 64);
    const uint32_t lt = vleaf ? vleaf[t] : t;
    const uint32_t pa = fuse_l2l ? par[lt] : NONE;
    for (int k = lane; k < S; k += 64) {
      sloc[k] = loc[(size_t)lt * S + k];
      if (pa != NONE) spar[k] = loc[(size_t)pa * S + k];
    }
    __syncthreads();
    const int half = lane >> 5;
    const uint32_t e = (uint32_t)(lane & 31);
    if ((half == 0 || pa != NONE) && e < na) {
      const double fcx = half ? comx[pa] : ox, fcy = half ? comy[pa] : oy;
      l2p_horner<P>(half ? spar : sloc, x[ts + e] - fcx, y[ts + e] - fcy, &fx, &fy, &fp);
    }
    if (pa != NONE) {
      fx += __shfl_xor(fx, 32, 64);
      fy += __shfl_xor(fy, 32, 64);
      if (PHI) fp += __shfl_xor(fp, 32, 64);
    }
  }
  if (lane < 32 && (uint32_t)lane < na) {
    const size_t o = (size_t)own * SYM_LEAF + (size_t)lane;
    own2[o] = make_double2((0.0 - fx) + (double)(tix + sx), (0.0 - fy) + (double)(tiy + sy));
    if (PHI) ownp[o] = fp + (double)(tip + sp);
  }
}

// A chunk of a leaf's refs (32 lanes, one per body; 8 chunks per block),
// summed in ref order in FP64: into the leaf's output when it is the leaf's
// only chunk, else into a chunk sum.
template <bool PHI>
__global__ void __launch_bounds__(256) k_sym_chunk_sum(uint32_t nch, const uint32_t* __restrict__ ch_leaf,
                                                       const uint32_t* __restrict__ ch_r0,
                                                       const uint32_t* __restrict__ ch_r1,
                                                       const uint32_t* __restrict__ ch_dst,
                                                       const uint32_t* __restrict__ refs,
                                                       const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                                       const double2* __restrict__ own2,
                                                       const double* __restrict__ ownp,
                                                       const float2* __restrict__ pair2,
                                                       const float* __restrict__ pairp, double2* __restrict__ cpart2,
                                                       double* __restrict__ cpartp, double* __restrict__ ax,
                                                       double* __restrict__ ay, double* __restrict__ phi) {
  __shared__ uint32_t sref[8][SYM_CHUNK];
  const uint32_t cl = threadIdx.x >> 5;
  const uint32_t c = blockIdx.x * 8 + cl;
  const uint32_t e = threadIdx.x & 31;
  const bool live = c < nch;
  uint32_t t = 0, b0 = 0, r0 = 0, nr = 0;
  bool body = false;
  if (live) {
    t = ch_leaf[c];
    b0 = bs[t];
    body = e < be[t] - b0;
    r0 = ch_r0[c];
    nr = ch_r1[c] - r0;
  }
  // the chunk's refs (at most SYM_CHUNK), loaded one per lane; its own-slot
  // refs come first (a leaf lists its items before its other pairs)
  const uint32_t myref = (live && e < nr) ? refs[r0 + e] : 0u;
  sref[cl][e] = myref;
  const unsigned long long own_lanes = __ballot((myref & SYM_OWN_TAG) != 0u);
  const uint32_t nown = (uint32_t)__popcll((own_lanes >> (threadIdx.x & 32)) & 0xFFFFFFFFull);
  __syncthreads();
  if (!live || !body) return;
  double sx = 0.0, sy = 0.0, sp = 0.0;
#pragma unroll 2
  for (uint32_t k = 0; k < nown; ++k) {
    const size_t o = (size_t)(sref[cl][k] & ~SYM_OWN_TAG) * SYM_LEAF + e;
    const double2 v = own2[o];
    sx += v.x;
    sy += v.y;
    if (PHI) sp += ownp[o];
  }
#pragma unroll 8
  for (uint32_t k = nown; k < nr; ++k) {
    const size_t o = (size_t)sref[cl][k] * SYM_LEAF + e;
    const float2 v = pair2[o];
    sx += (double)v.x;
    sy += (double)v.y;
    if (PHI) sp += (double)pairp[o];
  }
  const uint32_t dst = ch_dst[c];
  if (dst == NONE) {
    ax[b0 + e] = sx;
    ay[b0 + e] = sy;
    if (PHI) phi[b0 + e] = sp;
  } else {
    cpart2[(size_t)dst * SYM_LEAF + e] = make_double2(sx, sy);
    if (PHI) cpartp[(size_t)dst * SYM_LEAF + e] = sp;
  }
}

// Leaves with several chunks: the chunk sums added in chunk order.
template <bool PHI>
__global__ void __launch_bounds__(256) k_sym_heavy_sum(uint32_t nheavy, const uint32_t* __restrict__ heavy_leaf,
                                                       const uint32_t* __restrict__ heavy_p0,
                                                       const uint32_t* __restrict__ heavy_n,
                                                       const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                                       const double2* __restrict__ cpart2,
                                                       const double* __restrict__ cpartp, double* __restrict__ ax,
                                                       double* __restrict__ ay, double* __restrict__ phi) {
  const uint32_t k = blockIdx.x * 8 + (threadIdx.x >> 5);
  const uint32_t e = threadIdx.x & 31;
  if (k >= nheavy) return;
  const uint32_t t = heavy_leaf[k];
  const uint32_t b0 = bs[t];
  if (e >= be[t] - b0) return;
  double sx = 0.0, sy = 0.0, sp = 0.0;
  const uint32_t p0 = heavy_p0[k], n = heavy_n[k];
#pragma unroll 8
  for (uint32_t c = 0; c < n; ++c) {
    const double2 v = cpart2[(size_t)(p0 + c) * SYM_LEAF + e];
    sx += v.x;
    sy += v.y;
    if (PHI) sp += cpartp[(size_t)(p0 + c) * SYM_LEAF + e];
  }
  ax[b0 + e] = sx;
  ay[b0 + e] = sy;
  if (PHI) phi[b0 + e] = sp;
}

// The symmetric near field's combine: each leaf's slots into (o0, o1[, o2]),
// the two force components and, with o2, the third value (the potential, or
// in the adjoint m-bar). bs, be: the structure's leaves (or halves).
int sym_combine(const SymP2P& sp, const uint32_t* bs, const uint32_t* be, double* o0, double* o1, double* o2,
                hipStream_t s) {
  auto go = [&](auto phc) -> int {
    constexpr bool PH = decltype(phc)::value;
    k_sym_chunk_sum<PH><<<(unsigned)((sp.n_chunks + 7) / 8), 256, 0, s>>>(
        (uint32_t)sp.n_chunks, sp.ch_leaf.as<uint32_t>(), sp.ch_r0.as<uint32_t>(), sp.ch_r1.as<uint32_t>(),
        sp.ch_dst.as<uint32_t>(), sp.refs.as<uint32_t>(), bs, be, sp.own2.as<double2>(), sp.ownp.as<double>(),
        sp.pair2.as<float2>(), sp.pairp.as<float>(), sp.cpart2.as<double2>(), sp.cpartp.as<double>(), o0, o1, o2);
    GX_HIP(hipGetLastError());
    if (sp.n_heavy) {
      k_sym_heavy_sum<PH><<<(unsigned)((sp.n_heavy + 7) / 8), 256, 0, s>>>(
          (uint32_t)sp.n_heavy, sp.heavy_leaf.as<uint32_t>(), sp.heavy_p0.as<uint32_t>(), sp.heavy_n.as<uint32_t>(),
          bs, be, sp.cpart2.as<double2>(), sp.cpartp.as<double>(), o0, o1, o2);
      GX_HIP(hipGetLastError());
    }
    return 0;
  };
  return o2 ? go(std::true_type{}) : go(std::false_type{});
}

// ------------------------------------------------------ direct sum ----

// ------------------------------------------------------ orchestration ----

#define GX_P_CASES(F) \
  F(0) F(1) F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15) F(16)

// Host restatement of compute_kernel_derivs' polynomial construction
// (galax-core/src/lib.rs:1059-1130); same operations in the same order.
std::vector<double> kernel_polys(double eps, int p) {
  const int order = 2 * p, stride = order + 1, GS = stride * stride;
  std::vector<double> poly((size_t)GS * GS, 0.0);
  auto P_ = [&](int k, int l) { return poly.data() + (size_t)(k * stride + l) * GS; };
  const double eps2 = eps * eps;
  P_(0, 0)[0] = -1.0;
  std::vector<double> coeff(GS);
  for (int n = 1; n <= order; ++n) {
    for (int k = 0; k <= n; ++k) {
      int l = n - k;
      std::fill(coeff.begin(), coeff.end(), 0.0);
      if (k > 0) {
        const double* prev = P_(k - 1, l);
        int mm = 1 + 2 * (k - 1) + 2 * l;
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = prev[a * stride + b];
            if (c == 0.0) continue;
            if (a > 0) {
              double dc = c * (double)a;
              if (a + 1 < stride) coeff[(a + 1) * stride + b] += dc;
              if (b + 2 < stride) coeff[(a - 1) * stride + b + 2] += dc;
              coeff[(a - 1) * stride + b] += dc * eps2;
            }
          }
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = prev[a * stride + b];
            if (c == 0.0) continue;
            if (a + 1 < stride) coeff[(a + 1) * stride + b] -= (double)mm * c;
          }
      } else if (l > 0) {
        const double* prev = P_(k, l - 1);
        int mm = 1 + 2 * k + 2 * (l - 1);
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = prev[a * stride + b];
            if (c == 0.0) continue;
            if (b > 0) {
              double dc = c * (double)b;
              if (a + 2 < stride) coeff[(a + 2) * stride + b - 1] += dc;
              if (b + 1 < stride) coeff[a * stride + b + 1] += dc;
              coeff[a * stride + b - 1] += dc * eps2;
            }
          }
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = prev[a * stride + b];
            if (c == 0.0) continue;
            if (b + 1 < stride) coeff[a * stride + b + 1] -= (double)mm * c;
          }
      }
      std::copy(coeff.begin(), coeff.end(), P_(k, l));
    }
  }
  return poly;
}

// Builds the key table of the live M2L pairs [0, m2l_off[n_targets]).
static int m2l_keys_setup(galax_fmm* h, uint32_t n_targets) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const int p = (int)h->cfg.p;
  std::vector<uint32_t> offs(2);
  GX_HIP(hipMemcpyAsync(offs.data() + 1, t.m2l_off.as<uint32_t>() + n_targets, 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  uint32_t n_pairs = offs[1];
  h->n_live_pairs = n_pairs;
  h->n_live_targets = n_targets;
  h->n_keys = 0;
  const int GS = (2 * p + 1) * (2 * p + 1);
  if (n_pairs == 0) {
    h->keys_valid = true;
    return 0;
  }
  uint64_t nslots = 64;
  while (nslots < 2ull * n_pairs) nslots <<= 1;
  uint64_t mask = nslots - 1;
  GX_CHECK(dbuf_reserve(h->hash, nslots * (8 + 4 + 4 + 4), s));
  unsigned long long* hkeys = h->hash.as<unsigned long long>();
  uint32_t* hrep = reinterpret_cast<uint32_t*>(hkeys + nslots);
  uint32_t* kid = hrep + nslots;
  uint32_t* slot_kid = kid + nslots;
  GX_HIP(hipMemsetAsync(hkeys, 0xFF, nslots * 8, s));
  GX_HIP(hipMemsetAsync(hrep, 0xFF, nslots * 4, s));
  GX_CHECK(dbuf_reserve(h->tmp_u32, (size_t)n_pairs * 4, s));
  GX_CHECK(dbuf_reserve(h->key_of_pair, (size_t)n_pairs * 4, s));
  GX_CHECK(dbuf_reserve(h->flag, 64, s));
  uint32_t* err = h->flag.as<uint32_t>();
  GX_HIP(hipMemsetAsync(err, 0, 4, s));
  const double* comx = t.com_x.as<double>();
  const double* comy = t.com_y.as<double>();
  k_pair_targets<<<nblk(n_targets), 256, 0, s>>>(t.m2l_off.as<uint32_t>(), n_targets, h->tmp_u32.as<uint32_t>());
  k_keys_insert<<<(unsigned)((n_pairs + KI_PAIRS - 1) / KI_PAIRS), 1024, 0, s>>>(
      t.m2l_dat.as<uint32_t>(), h->tmp_u32.as<uint32_t>(), n_pairs, comx, comy, hkeys, hrep, mask, err);
  GX_HIP(hipGetLastError());
  uint32_t herr = 0;
  GX_HIP(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  if (herr)
    return set_error(GALAX_ERR_UNSUPPORTED, "M2L cache key (dx as i64, dy as i64) outside +-2^31 is not supported");
  k_keys_flag<<<nblk(nslots), 256, 0, s>>>(hkeys, nslots, kid);
  GX_HIP(hipGetLastError());
  uint32_t nk = 0;
  GX_CHECK(scan_u32(h, kid, kid, nslots, false, false, &nk));
  h->n_keys = nk;
  GX_CHECK(dbuf_reserve(h->key_dxy, (size_t)nk * 16 + 16, s));
  double* dxy = h->key_dxy.as<double>();
  k_keys_compact<<<nblk(nslots), 256, 0, s>>>(hkeys, hrep, nslots, kid, slot_kid, t.m2l_dat.as<uint32_t>(),
                                              h->tmp_u32.as<uint32_t>(), comx, comy, dxy);
  GX_HIP(hipGetLastError());
  k_keys_lookup<<<nblk(n_pairs), 256, 0, s>>>(t.m2l_dat.as<uint32_t>(), h->tmp_u32.as<uint32_t>(), n_pairs, comx,
                                              comy, hkeys, slot_kid, mask, h->key_of_pair.as<uint32_t>());
  GX_HIP(hipGetLastError());
  // derivative tables of the representatives, with the reference's polynomials
  std::vector<double> polys = kernel_polys(h->cfg.softening, p);
  GX_CHECK(dbuf_reserve(h->ws.f, polys.size() * 8, s));
  GX_HIP(hipMemcpyAsync(h->ws.f.p, polys.data(), polys.size() * 8, hipMemcpyHostToDevice, s));
  GX_CHECK(dbuf_reserve(h->gtab, (size_t)nk * GS * 8, s));
  GX_HIP(hipMemsetAsync(h->gtab.p, 0, (size_t)nk * GS * 8, s));
  k_derivs_poly<<<nk, 64, 0, s>>>(h->ws.f.as<double>(), p, dxy, nk, h->cfg.softening, h->gtab.as<double>());
  GX_HIP(hipGetLastError());
  GX_HIP(hipStreamSynchronize(s));
  h->keys_valid = true;
  return 0;
}

// per level l >= 1 of the tree: do its parents form one-child chains
// aligned with the level below (k_m2m_chain / k_l2l_chain)? Once per tree.
static int chain_check(galax_fmm* h) {
  DevTree& t = h->tree;
  if (t.chain_ok.size() == (size_t)t.n_levels + 1) return 0;
  hipStream_t s = h->stream;
  std::vector<uint8_t> ok(t.n_levels + 1, 0);
  std::vector<uint32_t> bad(t.n_levels + 1, 0);
  ScopedDBuf flags;
  GX_CHECK(dbuf_reserve(flags, (t.n_levels + 1) * 4, s));
  GX_HIP(hipMemsetAsync(flags.p, 0, (t.n_levels + 1) * 4, s));
  for (uint32_t l = 1; l < t.n_levels; ++l) {
    const uint32_t lo = t.level_offset[l], n = t.level_offset[l + 1] - lo;
    if (n == 0 || n != lo - t.level_offset[l - 1]) continue;
    ok[l] = 1;
    k_chain_check<<<nblk(n), 256, 0, s>>>(t.child_first.as<uint32_t>(), lo, n, flags.as<uint32_t>() + l);
    GX_HIP(hipGetLastError());
  }
  GX_HIP(hipMemcpyAsync(bad.data(), flags.p, (t.n_levels + 1) * 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  for (uint32_t l = 0; l <= t.n_levels; ++l) ok[l] = ok[l] && !bad[l];
  t.chain_ok.swap(ok);
  return 0;
}

template <int P>
static int run_passes(galax_fmm* h, double* d_x, double* d_y, double* d_m, double* d_ax, double* d_ay,
                      double* d_phi, bool near) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  constexpr int S = num_moments(P);
  const bool corrected = h->cfg.algo == GALAX_ALGO_CORRECTED;
  const size_t nn = t.n_nodes, nl = t.n_leaves;
  const double eps = h->cfg.softening;
  const double* comx = t.com_x.as<double>();
  const double* comy = t.com_y.as<double>();
  GX_CHECK(dbuf_reserve(h->mp, nn * S * 8, s));
  GX_CHECK(dbuf_reserve(h->loc, nn * S * 8, s));
  double* mp = h->mp.as<double>();
  double* loc = h->loc.as<double>();
  // Levels whose expansions reach a body. REFERENCE: the ascending-id L2L
  // (galax-core/src/lib.rs:1266) pushes each node's local into its children
  // before its own parent has pushed into it, so a leaf receives its own M2L
  // and its parent's M2L only; multipoles above level 1 feed only M2L at
  // levels >= 2, whose locals never reach a body. Those passes are skipped.
  const uint32_t top = corrected ? t.n_levels : std::min<uint32_t>(t.n_levels, 2);
  const uint32_t live_hi = t.level_offset[top];
  hipEvent_t* ev = h->ev;
  const uint32_t* goff = corrected ? h->c_gat_off.as<uint32_t>() : t.gat_off.as<uint32_t>();
  const uint32_t* gdat = corrected ? h->c_gat_dat.as<uint32_t>() : t.gat_dat.as<uint32_t>();
  const bool fp32_near =
      h->cfg.precision == GALAX_PREC_MIXED_FP32 || h->cfg.precision == GALAX_PREC_MIXED_BF16X3;
  // the FP32 forms with leaves of at most SYM_LEAF bodies evaluate each
  // unordered leaf pair once (k_p2p_sym); P2M then also writes the bodies'
  // leaf-relative FP32 coordinates
  SymP2P* sym = nullptr;
  if (near && fp32_near && h->cfg.n_max <= 2 * SYM_LEAF && !h->near_gather && nl > 0) {
    SymP2P& sp = corrected ? h->c_sym : t.sym;
    if (!sp.valid) GX_CHECK(build_sym_near(h, goff, gdat, (uint32_t)nl, sp));
    if (d_phi) {  // the potentials' slots, sized for this tree (a rebuilt tree may have more pairs)
      GX_CHECK(dbuf_reserve(sp.ownp, std::max<size_t>(sp.n_items, 1) * SYM_LEAF * 8, s));
      GX_CHECK(dbuf_reserve(sp.pairp, std::max<size_t>(sp.n_pairs, 1) * SYM_LEAF * 4, s));
      GX_CHECK(dbuf_reserve(sp.cpartp, std::max<size_t>(sp.n_cpart, 1) * SYM_LEAF * 8, s));
    }
    sym = &sp;
  }
  // the mixed forms' M2L tiles read FP32 multipole rows, written by P2M and M2M
  const bool rows32 = !corrected && h->mx_rows32 && h->cfg.precision != GALAX_PREC_FP64 &&
                      h->cfg.precision != GALAX_PREC_FP64_EXACT;
  if (h->timing) GX_HIP(hipEventRecord(ev[1], s));
  {
    uint32_t* changed = h->flag.as<uint32_t>() + 4 + h->san_cur;
    uint32_t* next = h->flag.as<uint32_t>() + 4 + (h->san_cur ^ 1u);
    k_p2m<P><<<nblk(nl, num_moments(P) <= 32 ? 8 : 4), 256, 0, s>>>(
        d_x, d_y, d_m, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), comx, comy, (uint32_t)nl, mp, changed,
        next, sym ? sym->rel.as<float4>() : nullptr, rows32 ? h->mp32.as<float>() : nullptr,
        rows32 ? h->mx_nexp.as<int8_t>() : nullptr);
    GX_HIP(hipGetLastError());
  }
  if (h->timing) GX_HIP(hipEventRecord(ev[2], s));
  {
    if (corrected) GX_CHECK(chain_check(h));
    for (uint32_t l = 1; l < top; ++l) {
      uint32_t lo = t.level_offset[l], hi = t.level_offset[l + 1];
      if (corrected && t.chain_ok[l]) {  // a chain of one-child levels l .. l2 in one launch
        uint32_t l2 = l;
        while (l2 + 1 < top && t.chain_ok[l2 + 1]) ++l2;
        constexpr int NB = 256 / num_moments(P);
        k_m2m_chain<P><<<nblk(hi - lo, NB), 256, 0, s>>>(comx, comy, lo, hi - lo, l2 - l + 1, mp);
        GX_HIP(hipGetLastError());
        l = l2;
        continue;
      }
      if (corrected && hi - lo == lo - t.level_offset[l - 1]) {  // every parent has one child
        k_m2m_single<P><<<nblk((size_t)(hi - lo) * num_moments(P)), 256, 0, s>>>(
            t.child_first.as<uint32_t>(), comx, comy, lo, hi, mp);
        GX_HIP(hipGetLastError());
        continue;
      }
      if constexpr (P <= 8) {
        if (corrected) {
          k_m2m_nodes<P><<<nblk(hi - lo, 64), 64, 0, s>>>(t.child_first.as<uint32_t>(), t.child_count.as<uint8_t>(),
                                                        comx, comy, lo, hi, mp);
          GX_HIP(hipGetLastError());
          continue;
        }
      }
      k_m2m<P><<<nblk(hi - lo, num_moments(P) <= 32 ? 8 : 4), 256, 0, s>>>(
          t.child_first.as<uint32_t>(), t.child_count.as<uint8_t>(), comx, comy, lo, hi, mp,
          rows32 && l == 1 ? h->mp32.as<float>() : nullptr, rows32 ? h->mx_nexp.as<int8_t>() : nullptr);
      GX_HIP(hipGetLastError());
    }
  }
  if (h->timing) GX_HIP(hipEventRecord(ev[3], s));
  if (h->timing) GX_HIP(hipEventRecord(ev[4], s));
  if (corrected) {
    // lane per pair up to p = 8 (at p = 8 some of the table spills, still faster);
    // above, the wave-per-pair kernel
    if constexpr (P <= 8) {
      const bool mixed = h->cfg.precision == GALAX_PREC_MIXED_FP32 || h->cfg.precision == GALAX_PREC_MIXED_BF16X3;
      if (mixed)
        k_m2l_exact_lp32<P><<<(unsigned)nn, 64, 0, s>>>(h->c_m2l_off.as<uint32_t>(), h->c_m2l_dat.as<uint32_t>(),
                                                        h->c_m2l_order.as<uint32_t>(), comx, comy, eps * eps, mp,
                                                        loc);
      else
        k_m2l_exact_lp<P><<<(unsigned)nn, 64, 0, s>>>(h->c_m2l_off.as<uint32_t>(), h->c_m2l_dat.as<uint32_t>(),
                                                      h->c_m2l_order.as<uint32_t>(), comx, comy, eps * eps, mp, loc);
    } else {
      k_m2l_exact<P><<<nblk(nn, 4), 256, 0, s>>>(h->c_m2l_off.as<uint32_t>(), h->c_m2l_dat.as<uint32_t>(),
                                                 (uint32_t)nn, comx, comy, eps * eps, mp, loc);
    }
    GX_HIP(hipGetLastError());
  } else if (live_hi > 0 && h->cfg.precision != GALAX_PREC_FP64_EXACT) {
    GX_CHECK(m2l_mfma_run(h, mp, loc));
  } else if (live_hi > 0) {
    k_m2l_ref<P><<<nblk(live_hi, 4), 256, 0, s>>>(t.m2l_off.as<uint32_t>(), t.m2l_dat.as<uint32_t>(),
                                                  h->key_of_pair.as<uint32_t>(), h->gtab.as<double>(), live_hi, mp,
                                                  loc);
    GX_HIP(hipGetLastError());
  }
  if (h->timing) GX_HIP(hipEventRecord(ev[5], s));
  // L2L top-down; the last level (parents to leaves) runs inside the L2P+P2P
  // kernel when the forces are computed (the VJP's expansions-only forward
  // needs the leaves' locals in memory)
  const bool fuse_l2l = near && t.n_levels >= 2;
  const uint32_t l2l_stop = fuse_l2l ? 2 : 1;
  if (corrected) {
    for (uint32_t l = t.n_levels - 1; l >= l2l_stop; --l) {
      uint32_t lo = t.level_offset[l - 1], hi = t.level_offset[l];
      if (t.chain_ok[l]) {  // a chain of one-child levels l .. l2 (descending) in one launch
        uint32_t l2 = l;
        while (l2 - 1 >= l2l_stop && t.chain_ok[l2 - 1]) --l2;
        constexpr int NB = 256 / num_moments(P);
        k_l2l_chain<P><<<nblk(hi - lo, NB), 256, 0, s>>>(comx, comy, hi, hi - lo, l - l2 + 1, loc);
        GX_HIP(hipGetLastError());
        l = l2;
        continue;
      }
      if (t.level_offset[l + 1] - hi == hi - lo)  // every parent has one child
        k_l2l_single<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi,
                                                                     loc);
      else if constexpr (P <= 8)
        k_l2l_nodes<P><<<nblk(hi - lo, 64), 64, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi, loc);
      else
        k_l2l<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi, loc);
      GX_HIP(hipGetLastError());
    }
  } else if (t.n_levels >= 2 && !fuse_l2l) {
    k_l2l<P><<<nblk(nl * S), 256, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, 0, (uint32_t)nl, loc);
    GX_HIP(hipGetLastError());
  }
  if (h->timing) GX_HIP(hipEventRecord(ev[6], s));
  const P2PItems& it = corrected ? h->c_items : t.items;
  const uint32_t nmax = h->cfg.n_max;
  if (!sym) GX_CHECK(dbuf_reserve(h->p2p_part, std::max<size_t>(it.n_part, 1) * nmax * 3 * 8, s));
#define GX_LAUNCH_P2P(T, PH)                                                                                 \
  k_l2p_p2p<P, T, PH><<<(unsigned)it.n_items, 64, 0, s>>>(                                                  \
      d_x, d_y, d_m, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), comx, comy, loc, goff, gdat,     \
      it.tgt.as<uint32_t>(), it.g0.as<uint32_t>(), it.g1.as<uint32_t>(), it.slot.as<uint32_t>(), nmax, eps * eps, \
      d_ax, d_ay, d_phi, h->p2p_part.as<double>(), t.parent.as<uint32_t>(), fuse_l2l ? 1 : 0)
  if (h->timing) GX_HIP(hipEventRecord(ev[10], s));
  if (!near) {  // expansions only (the VJP's forward): no L2P / P2P
    if (h->timing) GX_HIP(hipEventRecord(ev[11], s));
    if (h->timing) GX_HIP(hipEventRecord(ev[7], s));
    return 0;
  }
  if (sym) {
    // the leaves, or their halves
    const uint32_t* sbs = sym->halves ? sym->vbs.as<uint32_t>() : t.body_start.as<uint32_t>();
    const uint32_t* sbe = sym->halves ? sym->vbe.as<uint32_t>() : t.body_end.as<uint32_t>();
    const double* scx = sym->halves ? sym->vcx.as<double>() : comx;
    const double* scy = sym->halves ? sym->vcy.as<double>() : comy;
    const uint32_t* svl = sym->halves ? sym->vleaf.as<uint32_t>() : nullptr;
#define GX_LAUNCH_SYM(PH)                                                                                       \
  k_p2p_sym<P, PH><<<(unsigned)sym->n_items, 64, 0, s>>>(                                                     \
      sym->rel.as<float4>(), d_x, d_y, sbs, sbe, scx, scy, svl, comx, comy, loc,                                \
      t.parent.as<uint32_t>(), fuse_l2l ? 1 : 0, sym->own_off.as<uint32_t>(), sym->own_dat.as<uint32_t>(),      \
      sym->itgt.as<uint32_t>(), sym->iu0.as<uint32_t>(), sym->iu1.as<uint32_t>(), sym->iown.as<uint32_t>(),      \
      (float)(eps * eps), sym->own2.as<double2>(), sym->ownp.as<double>(), sym->pair2.as<float2>(),            \
      sym->pairp.as<float>());                                                                                  \
  GX_HIP(hipGetLastError());                                                                                    \
  if (h->timing) GX_HIP(hipEventRecord(ev[11], s));                                                             \
  k_sym_chunk_sum<PH><<<(unsigned)((sym->n_chunks + 7) / 8), 256, 0, s>>>(                                     \
      (uint32_t)sym->n_chunks, sym->ch_leaf.as<uint32_t>(), sym->ch_r0.as<uint32_t>(), sym->ch_r1.as<uint32_t>(), \
      sym->ch_dst.as<uint32_t>(), sym->refs.as<uint32_t>(), sbs, sbe, sym->own2.as<double2>(),                  \
      sym->ownp.as<double>(), sym->pair2.as<float2>(),                                                          \
      sym->pairp.as<float>(), sym->cpart2.as<double2>(), sym->cpartp.as<double>(), d_ax, d_ay, d_phi);         \
  GX_HIP(hipGetLastError());                                                                                    \
  if (sym->n_heavy)                                                                                             \
    k_sym_heavy_sum<PH><<<(unsigned)((sym->n_heavy + 7) / 8), 256, 0, s>>>(                                    \
        (uint32_t)sym->n_heavy, sym->heavy_leaf.as<uint32_t>(), sym->heavy_p0.as<uint32_t>(),                   \
        sym->heavy_n.as<uint32_t>(), sbs, sbe,                                                                  \
        sym->cpart2.as<double2>(), sym->cpartp.as<double>(), d_ax, d_ay, d_phi)
    if (d_phi) {
      GX_LAUNCH_SYM(true);
    } else {
      GX_LAUNCH_SYM(false);
    }
#undef GX_LAUNCH_SYM
    GX_HIP(hipGetLastError());
    if (h->timing) GX_HIP(hipEventRecord(ev[7], s));
    return 0;
  }
  if (fp32_near) {
    if (d_phi)
      GX_LAUNCH_P2P(float, true);
    else
      GX_LAUNCH_P2P(float, false);
  } else {
    if (d_phi)
      GX_LAUNCH_P2P(double, true);
    else
      GX_LAUNCH_P2P(double, false);
  }
  GX_HIP(hipGetLastError());
  if (h->timing) GX_HIP(hipEventRecord(ev[11], s));
  GX_HIP(p2p_combine(it, goff, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), nmax,
                     h->p2p_part.as<double>(), d_ax, d_ay, d_phi, s));
#undef GX_LAUNCH_P2P
  GX_HIP(hipGetLastError());
  if (h->timing) GX_HIP(hipEventRecord(ev[7], s));
  return 0;
}

int evaluate(galax_fmm* h, double* d_x, double* d_y, double* d_m, double* d_ax, double* d_ay, double* d_phi,
             bool near) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const size_t n = t.n_bodies;
  GX_CHECK(upload_tabs(s));
  for (int i = 0; i < 12; ++i)
    if (!h->ev[i]) GX_HIP(hipEventCreate(&h->ev[i]));
  if (n == 0) return 0;
  if (t.n_nodes == 0) return set_error(GALAX_ERR_INVALID, "no tree installed");
  const bool corrected = h->cfg.algo == GALAX_ALGO_CORRECTED;
  if (corrected && !h->corrected_ready) return set_error(GALAX_ERR_INVALID, "corrected lists not built");
  if (!corrected && !h->keys_valid) {
    uint32_t top = std::min<uint32_t>(t.n_levels, 2);
    GX_CHECK(m2l_keys_setup(h, t.level_offset[top]));
    h->mx_valid = h->mx_adj_valid = h->mx_rows32 = false;
  }
  if (!corrected && h->cfg.precision != GALAX_PREC_FP64_EXACT && !h->mx_valid) GX_CHECK(m2l_mfma_setup(h));
  GX_CHECK(dbuf_reserve(h->flag, 64, s));
  if (!h->san_init) {
    GX_HIP(hipMemsetAsync(h->flag.as<uint32_t>() + 4, 0, 8, s));
    h->san_init = true;
    h->san_cur = 1;
  }
  h->san_cur ^= 1u;  // the sanitize runs inside the P2M kernel
  if (h->timing) GX_HIP(hipEventRecord(h->ev[0], s));
  int rc = 0;
  switch (h->cfg.p) {
#define GX_CASE(PP)                                                  \
  case PP:                                                           \
    rc = run_passes<PP>(h, d_x, d_y, d_m, d_ax, d_ay, d_phi, near);  \
    break;
    GX_P_CASES(GX_CASE)
#undef GX_CASE
    default:
      return set_error(GALAX_ERR_INVALID, "p = %u outside 0..16", h->cfg.p);
  }
  GX_CHECK(rc);
  GX_HIP(hipStreamSynchronize(s));
  galax_timings_t& tm = h->timings;
  // pass times when the events were recorded (galax_fmm_set_timing), else 0
  float ms[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (h->timing)
    for (int i = 0; i < 7; ++i) GX_HIP(hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]));
  tm.sanitize = ms[0];
  tm.p2m = ms[1];
  tm.m2m = ms[2];
  tm.m2l_setup = ms[3];
  tm.m2l = ms[4];
  tm.l2l = ms[5];
  tm.l2p_p2p = ms[6];
  float tot = 0.f;
  if (h->timing) GX_HIP(hipEventElapsedTime(&tot, h->ev[0], h->ev[7]));
  tm.total = tot;
  tm.m2l_pairs_executed = corrected ? h->c_m2l_total : h->n_live_pairs;
  tm.m2l_pairs_reference = corrected ? h->c_m2l_total : (t.m2l_full ? t.m2l_total : 0);  // 0: not built yet
  tm.m2l_keys = corrected ? 0 : h->n_keys;
  tm.p2p_interactions = corrected ? h->c_p2p_interactions : t.p2p_interactions;
  const bool tiled = !corrected && h->cfg.precision != GALAX_PREC_FP64_EXACT;
  tm.m2l_groups = tiled ? h->mx_fwd.ngroups : 0;
  tm.m2l_steps = tiled ? h->mx_fwd.nsteps : 0;
  tm.m2l_items = tiled ? h->mx_fwd.nitems : 0;
  float km = 0.f, kp = 0.f;
  if (h->timing && tiled && h->mx_fwd.nitems && h->n_live_targets)
    GX_HIP(hipEventElapsedTime(&km, h->ev[8], h->ev[9]));
  if (h->timing) GX_HIP(hipEventElapsedTime(&kp, h->ev[10], h->ev[11]));
  tm.m2l_kernel = km;
  tm.p2p_kernel = kp;
  return 0;
}

}  // namespace gx
