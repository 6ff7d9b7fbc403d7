// tree.hip — the reference's Morton-chunk tree and its interaction lists, built
// on the device and bit-exact with galax-core:
//   Morton keys            galax-core/src/lib.rs:73-109
//   stable sort + permute  galax-core/src/lib.rs:146-217 (sort_by_key is stable)
//   leaves                 galax-core/src/lib.rs:219-267
//   parent rounds          galax-core/src/lib.rs:269-354
//   M2L lists              galax-core/src/lib.rs:947-1025
//   P2P lists              galax-core/src/lib.rs:1358-1392
// No FMA contraction anywhere in this file: every product is rounded before it
// is added, as in Rust, so the f64 geometry matches bit for bit.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace gx {

// ------------------------------------------------------ Rust semantics ----

__device__ __forceinline__ uint64_t rust_f64_as_u64(double v) {
  if (!(v > 0.0)) return 0;  // NaN, zero and negatives
  if (v >= 18446744073709551616.0) return ~0ull;
  return (uint64_t)v;
}

__device__ __forceinline__ uint64_t normalize(double x, double lo, double hi) {
  double t = (x - lo) / (hi - lo);
  const double top = 1.0 - 1e-15;  // f64::clamp(0.0, 1.0 - 1e-15), NaN passes
  if (t < 0.0) t = 0.0;
  if (t > top) t = top;
  return rust_f64_as_u64(t * 2097152.0);  // (1 << MORTON_BITS) as f64
}

__device__ __forceinline__ uint64_t spread_bits(uint64_t v) {
  uint64_t x = v & ((1ull << MORTON_BITS) - 1);
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

__global__ void k_morton(const double* __restrict__ x, const double* __restrict__ y, size_t n, double x0,
                         double x1, double y0, double y1, uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = spread_bits(normalize(x[i], x0, x1)) | (spread_bits(normalize(y[i], y0, y1)) << 1);
  if (idx) idx[i] = (uint32_t)i;
}

// ------------------------------------------------- stable radix sort ----
// LSD radix sort of (42-bit key, u32 payload), 7-bit digits, 6 passes. Each
// 256-thread block owns a tile of 4096 keys read in 16 rounds of 256
// consecutive keys; within a round, lanes holding the same digit are found
// with 7 wave ballots and ranked by lane order, and per-wave digit counts are
// combined through LDS. Ranks therefore follow input order: the sort is stable.

constexpr int RS_BITS = 7, RS_RADIX = 1 << RS_BITS, RS_THREADS = 256, RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

// Digit counts of each tile: hist[digit][tile] (scanned into global offsets
// next) and a copy by tile, tcnt[tile][digit], for the scatter's local order.
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t* __restrict__ keys, size_t n, int shift,
                                                        uint32_t* __restrict__ hist, uint32_t nb,
                                                        uint32_t* __restrict__ tcnt) {
  __shared__ uint32_t hcount[RS_RADIX];
  for (int t = threadIdx.x; t < RS_RADIX; t += RS_THREADS) hcount[t] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int k = 0; k < RS_ITEMS; ++k) {
    size_t i = base + (size_t)k * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&hcount[(keys[i] >> shift) & (RS_RADIX - 1)], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < RS_RADIX; t += RS_THREADS) {
    hist[(size_t)t * nb + blockIdx.x] = hcount[t];
    tcnt[(size_t)blockIdx.x * RS_RADIX + t] = hcount[t];
  }
}

// One tile of RS_TILE keys per block, stably: the tile's digit histogram and
// its exclusive scan (bucket starts within the tile); each key's rank among
// the tile's keys of its digit (ballot ranking per wave, 16 rounds in key
// order); the keys and values placed in LDS in (digit, rank) order; then
// consecutive threads write consecutive LDS entries, so each bucket's keys
// leave the tile as one contiguous run (~RS_TILE / RS_RADIX keys) instead of
// one scattered 8-byte write per key.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, size_t n, int shift,
                                                           const uint32_t* __restrict__ offs, uint32_t nb,
                                                           const uint32_t* __restrict__ tcnt) {
  static_assert(RS_RADIX == 128, "the tile scan takes two digits per lane of one wave");
  __shared__ uint64_t skey[RS_TILE];
  __shared__ uint32_t sval[RS_TILE];
  __shared__ uint32_t run[RS_RADIX];
  __shared__ uint32_t boff[RS_RADIX];
  __shared__ uint32_t tstart[RS_RADIX];
  __shared__ uint32_t cnt[RS_WAVES][RS_RADIX];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const size_t base = (size_t)blockIdx.x * RS_TILE;
  const uint32_t tn = (uint32_t)min((size_t)RS_TILE, n - base);
  for (int t = tid; t < RS_RADIX; t += RS_THREADS) {
    run[t] = 0;
    boff[t] = offs[(size_t)t * nb + blockIdx.x];
  }
  // bucket starts within the tile: exclusive scan of its digit counts, two
  // digits per lane of wave 0
  if (w == 0) {
    const uint32_t c0 = tcnt[(size_t)blockIdx.x * RS_RADIX + 2 * lane];
    const uint32_t c1 = tcnt[(size_t)blockIdx.x * RS_RADIX + 2 * lane + 1];
    uint32_t incl = c0 + c1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= o) incl += v;
    }
    tstart[2 * lane] = incl - (c0 + c1);
    tstart[2 * lane + 1] = incl - c1;
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int k = 0; k < RS_ITEMS; ++k) {
    for (int t = tid; t < RS_WAVES * RS_RADIX; t += RS_THREADS) (&cnt[0][0])[t] = 0;
    __syncthreads();
    const uint32_t j = (uint32_t)k * RS_THREADS + tid;
    const bool valid = j < tn;
    const uint64_t key = valid ? kin[base + j] : 0ull;
    const uint32_t val = valid ? vin[base + j] : 0u;
    const uint32_t d = (uint32_t)(key >> shift) & (RS_RADIX - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    const bool leader = valid && ((peers & lt_mask) == 0);
    if (leader) cnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    for (int t = tid; t < RS_RADIX; t += RS_THREADS) {
      uint32_t sacc = run[t];
#pragma unroll
      for (int ww = 0; ww < RS_WAVES; ++ww) {
        const uint32_t c = cnt[ww][t];
        cnt[ww][t] = sacc;
        sacc += c;
      }
      run[t] = sacc;
    }
    __syncthreads();
    if (valid) {
      const uint32_t lpos = tstart[d] + cnt[w][d] + rank;
      skey[lpos] = key;
      sval[lpos] = val;
    }
    __syncthreads();  // cnt is cleared for the next round
  }
  for (uint32_t j = tid; j < tn; j += RS_THREADS) {
    const uint64_t key = skey[j];
    const uint32_t d = (uint32_t)(key >> shift) & (RS_RADIX - 1);
    const uint32_t pos = boff[d] + (j - tstart[d]);
    kout[pos] = key;
    vout[pos] = sval[j];
  }
}

// Sorts keys (in ka) with payload vals (in va) on the low `bits` key bits;
// result in place. kb/vb are scratch of n entries each.
int radix_sort_bits(galax_fmm* h, uint64_t* ka, uint32_t* va, uint64_t* kb, uint32_t* vb, size_t n, int bits) {
  if (n <= 1) return 0;
  hipStream_t s = h->stream;
  uint32_t nb = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  GX_CHECK(dbuf_reserve(h->ws.sort_hist, 2 * (size_t)RS_RADIX * nb * sizeof(uint32_t), s));
  uint32_t* hist = h->ws.sort_hist.as<uint32_t>();
  uint32_t* tcnt = hist + (size_t)RS_RADIX * nb;
  int passes = (bits + RS_BITS - 1) / RS_BITS;
  if (passes & 1) ++passes;  // an even number of passes leaves the result in (ka, va)
  uint64_t *ki = ka, *ko = kb;
  uint32_t *vi = va, *vo = vb;
  for (int pass = 0; pass < passes; ++pass) {
    int shift = pass * RS_BITS;
    if (shift >= 64) shift = 63;  // padding pass: every digit equal, order kept
    k_rs_hist<<<nb, RS_THREADS, 0, s>>>(ki, n, shift, hist, nb, tcnt);
    GX_HIP(hipGetLastError());
    GX_CHECK(scan_u32(h, hist, hist, (size_t)RS_RADIX * nb, false, false, nullptr));
    k_rs_scatter<<<nb, RS_THREADS, 0, s>>>(ki, vi, ko, vo, n, shift, hist, nb, tcnt);
    GX_HIP(hipGetLastError());
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  return 0;
}

static int radix_sort(galax_fmm* h, uint64_t* ka, uint32_t* va, uint64_t* kb, uint32_t* vb, size_t n) {
  return radix_sort_bits(h, ka, va, kb, vb, n, 2 * MORTON_BITS);
}

// ---------------------------------------------------------- permute ----

__global__ void k_gather_f64(const double* __restrict__ src, const uint32_t* __restrict__ idx, size_t n,
                             double* __restrict__ dst) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// ------------------------------------------------------------ leaves ----

__device__ __forceinline__ void leaf_geometry(const double* __restrict__ x, const double* __restrict__ y,
                                              const double* __restrict__ m, uint32_t s, uint32_t e, double* cx,
                                              double* cy, double* hw, double* ms, double* comx, double* comy) {
  double mass_sum = 0.0, sx = 0.0, sy = 0.0;
  double bx0 = 1.7976931348623157e308, bx1 = -1.7976931348623157e308;
  double by0 = 1.7976931348623157e308, by1 = -1.7976931348623157e308;
  for (uint32_t i = s; i < e; ++i) {
    double mi = m[i], xi = x[i], yi = y[i];
    mass_sum += mi;
    sx += xi * mi;
    sy += yi * mi;
    bx0 = fmin(bx0, xi);  // f64::min / f64::max
    bx1 = fmax(bx1, xi);
    by0 = fmin(by0, yi);
    by1 = fmax(by1, yi);
  }
  double c_x = (bx0 + bx1) * 0.5, c_y = (by0 + by1) * 0.5;
  *cx = c_x;
  *cy = c_y;
  *hw = fmax(bx1 - bx0, by1 - by0) * 0.5 + 1e-15;
  *ms = mass_sum;
  *comx = mass_sum > 0.0 ? sx / mass_sum : c_x;
  *comy = mass_sum > 0.0 ? sy / mass_sum : c_y;
}

struct NodeArrays {
  double *cx, *cy, *hw, *mass, *comx, *comy;
  uint32_t *bs, *be, *cf, *par;
  uint8_t *cc, *lv;
};

static NodeArrays node_arrays(galax_fmm* h) {
  DevTree& t = h->tree;
  return NodeArrays{t.center_x.as<double>(),   t.center_y.as<double>(),   t.half_width.as<double>(),
                    t.mass.as<double>(),       t.com_x.as<double>(),      t.com_y.as<double>(),
                    t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), t.child_first.as<uint32_t>(),
                    t.parent.as<uint32_t>(),   t.child_count.as<uint8_t>(), t.level.as<uint8_t>()};
}

// leaf ranges: [lstart[i], lstart[i+1]) with lstart[n_leaves] = n
__global__ void k_leaves(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ m,
                         const uint32_t* __restrict__ lstart, uint32_t n_leaves, NodeArrays a) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_leaves) return;
  uint32_t s = lstart[i], e = lstart[i + 1];
  leaf_geometry(x, y, m, s, e, &a.cx[i], &a.cy[i], &a.hw[i], &a.mass[i], &a.comx[i], &a.comy[i]);
  a.bs[i] = s;
  a.be[i] = e;
  a.cf[i] = NONE;
  a.cc[i] = 0;
  a.par[i] = NONE;
  a.lv[i] = 0;
}

__global__ void k_uniform_leaf_starts(uint32_t* lstart, uint32_t n_leaves, uint32_t n_max, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n_leaves) return;
  uint64_t v = (uint64_t)i * n_max;
  lstart[i] = (uint32_t)(v < n ? v : n);
}

// ----------------------------------------------------- parent rounds ----
// One round groups the nodes [cs, ce) into runs of equal keys[body_start] >> shift,
// split into chunks of at most 4 from the start of each run; chunk j becomes
// parent ce + j (galax-core/src/lib.rs:274-354).

__global__ void k_round_heads(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ bs, uint32_t cs,
                              uint32_t c, uint32_t shift, uint32_t* __restrict__ head) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= c) return;
  uint32_t sh = shift & 63;  // release-mode Rust masks the shift amount
  uint64_t g = keys[bs[cs + j]] >> sh;
  bool start = (j == 0) || (g != (keys[bs[cs + j - 1]] >> sh));
  head[j] = start ? j : 0u;
}

__global__ void k_round_chunks(const uint32_t* __restrict__ run_start, uint32_t c, uint32_t* __restrict__ flag) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= c) return;
  flag[j] = ((j - run_start[j]) % 4u == 0u) ? 1u : 0u;
}

__global__ void k_round_mark(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pidx, uint32_t cs,
                             uint32_t ce, uint32_t c, uint32_t* __restrict__ cf) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= c) return;
  if (flag[j]) cf[ce + pidx[j]] = cs + j;
}

__device__ __forceinline__ void round_parent(NodeArrays a, uint32_t ce, uint32_t np, uint32_t q, uint8_t level,
                                             int constrained_style) {
  uint32_t p = ce + q;
  uint32_t first = a.cf[p];
  uint32_t last = (q + 1 < np) ? a.cf[p + 1] : ce;
  double mass_sum = 0.0, sx = 0.0, sy = 0.0;
  double bx0 = 1.7976931348623157e308, bx1 = -1.7976931348623157e308;
  double by0 = 1.7976931348623157e308, by1 = -1.7976931348623157e308;
  double max_hw = 0.0;
  for (uint32_t c = first; c < last; ++c) {
    double cm = a.mass[c], chw = a.hw[c], ccx = a.cx[c], ccy = a.cy[c];
    mass_sum += cm;
    sx += a.comx[c] * cm;
    sy += a.comy[c] * cm;
    if (constrained_style) {  // build_with_constraints, galax-core/src/lib.rs:660-668
      max_hw = fmax(max_hw, chw);
      bx0 = fmin(bx0, ccx - chw);
      bx1 = fmax(bx1, ccx + chw);
      by0 = fmin(by0, ccy - chw);
      by1 = fmax(by1, ccy + chw);
    } else {  // Tree::build, galax-core/src/lib.rs:302-318
      if (chw > max_hw) max_hw = chw;
      double cl = ccx - chw, cr = ccx + chw;
      if (cl < bx0) bx0 = cl;
      if (cr > bx1) bx1 = cr;
      double cb = ccy - chw, ct = ccy + chw;
      if (cb < by0) by0 = cb;
      if (ct > by1) by1 = ct;
    }
    a.par[c] = p;
  }
  double c_x = (bx0 + bx1) * 0.5, c_y = (by0 + by1) * 0.5;
  a.cx[p] = c_x;
  a.cy[p] = c_y;
  a.hw[p] = max_hw * 2.0;
  a.mass[p] = mass_sum;
  a.comx[p] = mass_sum > 0.0 ? sx / mass_sum : c_x;
  a.comy[p] = mass_sum > 0.0 ? sy / mass_sum : c_y;
  a.bs[p] = a.bs[first];
  a.be[p] = a.be[last - 1];
  a.cc[p] = (uint8_t)(last - first);
  a.par[p] = NONE;
  a.lv[p] = level;
}

__global__ void k_round_parents(NodeArrays a, uint32_t ce, uint32_t np, uint8_t level, int constrained_style) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  round_parent(a, ce, np, q, level, constrained_style);
}

// Block-wide exclusive scan (sum or max) of one value per thread of a
// 1024-thread block; *tot gets the reduction. Ends with a barrier.
template <bool MAX>
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* wsum, uint32_t* tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(inc, d, 64);
    if (lane >= d) inc = MAX ? (u > inc ? u : inc) : inc + u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < 16; ++k) {
      const uint32_t x = wsum[k];
      wsum[k] = acc;
      acc = MAX ? (x > acc ? x : acc) : acc + x;
    }
    wsum[16] = acc;
  }
  __syncthreads();
  const uint32_t pre = wsum[w];
  const uint32_t ex_in_wave = MAX ? __shfl_up(inc, 1, 64) : inc - v;  // exclusive within the wave
  uint32_t r;
  if (MAX)
    r = lane == 0 ? pre : (ex_in_wave > pre ? ex_in_wave : pre);
  else
    r = pre + ex_in_wave;
  *tot = wsum[16];
  __syncthreads();
  return r;
}

// Every remaining parent round of at most FUSED_ROUND nodes in one
// 1024-thread workgroup, topology only: which nodes start a parent (runs of
// equal key >> shift, a new parent every 4 nodes of a run; k_round_heads,
// the inclusive max scan, k_round_chunks, the exclusive sum scan and
// k_round_mark of the multi-kernel path) and each parent's first child. The
// parents' geometry (round_parent) runs afterwards per level with a thread
// per parent over the whole GPU (k_round_parents): one workgroup computing it
// round by round waited on memory for ~1.4 ms at 2^20 bodies.
// A node's group key is the Morton key of its first body (a parent shares its
// first child's, galax-core/src/lib.rs:278-282), and nodes are in key order,
// so the run test needs only hb[j], the bit length of key_j ^ key_(j-1):
// (key_j >> sh) != (key_(j-1) >> sh) iff hb[j] > sh. A parent's hb is the
// largest hb over its first child and the previous parent's other children
// (the bit length of an XOR of sorted keys is the largest over the steps
// between them). hb and the chunk-head flags are bytes in LDS (96 KB). Stops
// early (state left in st) when a round is larger than FUSED_ROUND or its
// parents would not fit in `cap` nodes; the host goes on from there.
// st: [cs, ce, shift, level, rounds done]; lvl_end[r] = ce after round r.
constexpr uint32_t FUSED_ROUND = 32768;

__global__ void __launch_bounds__(1024) k_rounds_topo(NodeArrays a, const uint64_t* __restrict__ keys, uint32_t cap,
                                                      uint32_t* __restrict__ st, uint32_t* __restrict__ lvl_end,
                                                      uint32_t max_rounds) {
  __shared__ uint32_t wsum[17];
  __shared__ uint8_t hb[2][FUSED_ROUND];
  __shared__ uint8_t flg[FUSED_ROUND];
  uint32_t cs = st[0], ce = st[1], shift = st[2], lvl = st[3], r = 0;
  const uint32_t tid = threadIdx.x;
  int cur = 0;
  if (ce - cs <= FUSED_ROUND)
    for (uint32_t j = tid; j < ce - cs; j += 1024) {
      uint32_t v = 255;  // the first node always starts a run
      if (j > 0) {
        const uint64_t x = keys[a.bs[cs + j]] ^ keys[a.bs[cs + j - 1]];
        v = x ? 64u - (uint32_t)__clzll((long long)x) : 0u;
      }
      hb[0][j] = (uint8_t)v;
    }
  __syncthreads();
  while (ce - cs > 1 && ce - cs <= FUSED_ROUND && r < max_rounds && lvl < 250) {
    const uint32_t c = ce - cs;
    const uint32_t sh = shift & 63;  // release-mode Rust masks the shift amount
    const uint8_t* h = hb[cur];
    const uint32_t K = (c + 1023) / 1024, lo = min(tid * K, c), hi = min(lo + K, c);
    uint32_t m = 0;  // last run start in this thread's chunk (0 = none: j = 0 always starts one)
    for (uint32_t j = lo; j < hi; ++j)
      if (h[j] > sh) m = j;
    uint32_t tot;
    uint32_t run = block_scan_excl<true>(m, wsum, &tot);
    uint32_t nf = 0;
    for (uint32_t j = lo; j < hi; ++j) {
      if (h[j] > sh) run = j;
      const uint8_t f = ((j - run) % 4u == 0u) ? 1 : 0;  // a new parent every 4 nodes of a run
      flg[j] = f;
      nf += f;
    }
    uint32_t np = 0;
    uint32_t pidx = block_scan_excl<false>(nf, wsum, &np);
    if (ce + np > cap) break;  // uniform: the host grows the arrays and goes on
    uint8_t* hn = hb[cur ^ 1];
    for (uint32_t j = lo; j < hi; ++j)
      if (flg[j]) {
        const uint32_t q = pidx++;
        a.cf[ce + q] = cs + j;
        uint32_t mx = h[j];
        if (j > 0)
          for (uint32_t t = j - 1; !flg[t]; --t) mx = max(mx, (uint32_t)h[t]);  // flg[0] is set
        hn[q] = (uint8_t)(q == 0 ? 255u : mx);
      }
    __syncthreads();
    cur ^= 1;
    ++lvl;
    cs = ce;
    ce += np;
    shift += 2;
    if (tid == 0) lvl_end[r] = ce;
    ++r;
  }
  if (tid == 0) {
    st[0] = cs;
    st[1] = ce;
    st[2] = shift;
    st[3] = lvl;
    st[4] = r;
  }
}

// Grows the node arrays to at least `nodes` entries, keeping their contents:
// every new array is allocated and its copy queued first, then one wait, then
// the old arrays are freed (one round trip instead of one per array).
int ensure_node_capacity(galax_fmm* h, size_t nodes) {
  DevTree& t = h->tree;
  if (nodes <= t.capacity) return 0;
  const size_t nc = std::max(nodes, t.capacity * 3 / 2 + 64);
  const size_t old = t.capacity;
  hipStream_t s = h->stream;
  DBuf* arr[] = {&t.center_x, &t.center_y, &t.half_width, &t.mass,        &t.com_x,       &t.com_y,
                 &t.body_start, &t.body_end, &t.child_first, &t.parent, &t.child_count, &t.level};
  const size_t esz[] = {8, 8, 8, 8, 8, 8, 4, 4, 4, 4, 1, 1};
  constexpr int NA = sizeof(esz) / sizeof(esz[0]);
  static_assert(sizeof(arr) / sizeof(arr[0]) == NA, "one element size per array");
  DBuf nb[NA];
  bool copied = false;
  for (int i = 0; i < NA; ++i) {
    if (int rc = dbuf_reserve(nb[i], nc * esz[i], s)) {
      for (int k = 0; k <= i; ++k) dbuf_free(nb[k]);
      return rc;
    }
    if (arr[i]->p && old) {
      GX_HIP(hipMemcpyAsync(nb[i].p, arr[i]->p, old * esz[i], hipMemcpyDeviceToDevice, s));
      copied = true;
    }
  }
  if (copied) GX_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < NA; ++i) {
    dbuf_free(*arr[i]);
    *arr[i] = nb[i];
  }
  t.capacity = nc;
  return 0;
}

static inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// Builds leaves from the leaf starts in ws.c (n_leaves + 1 entries) and runs the
// parent rounds. Bodies are in h->x, h->y, h->m (tree order); keys in tree.keys.
static int build_from_leaf_starts(galax_fmm* h, size_t n, size_t n_leaves, bool constrained_style) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  // room for the rounds: a tree small enough for the fused rounds gets the
  // capacity they ask for (32 more rounds of n_leaves nodes) at once
  GX_CHECK(ensure_node_capacity(h, n_leaves <= FUSED_ROUND ? n_leaves * 33 + 64 : n_leaves * 4 + 64));
  NodeArrays a = node_arrays(h);
  k_leaves<<<nblk(n_leaves), 256, 0, s>>>(h->x.as<double>(), h->y.as<double>(), h->m.as<double>(),
                                          h->ws.c.as<uint32_t>(), (uint32_t)n_leaves, a);
  GX_HIP(hipGetLastError());
  t.chain_ok.clear();
  t.level_offset.assign(1, 0);
  t.level_offset.push_back((uint32_t)n_leaves);
  size_t cs = 0, ce = n_leaves;
  uint32_t shift = 2;
  uint32_t lvl = 0;
  GX_CHECK(dbuf_reserve(h->ws.e, (n_leaves + 1) * sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(h->ws.f, (n_leaves + 1) * sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(h->ws.g, (n_leaves + 1) * sizeof(uint32_t), s));
  while (ce - cs > 1) {
    uint32_t c = (uint32_t)(ce - cs);
    if (c <= FUSED_ROUND) {
      // the remaining rounds in one launch; the arrays sized for 32 more
      // rounds of c nodes (the kernel stops early if they would not fit)
      const uint32_t max_rounds = 64;
      GX_CHECK(ensure_node_capacity(h, ce + (size_t)c * 32 + 64));
      a = node_arrays(h);
      GX_CHECK(dbuf_reserve(h->ws.h, (8 + max_rounds) * sizeof(uint32_t), s));
      uint32_t* dst = h->ws.h.as<uint32_t>();
      uint32_t hst[8 + 64] = {(uint32_t)cs, (uint32_t)ce, shift, lvl, 0, 0, 0, 0};
      GX_HIP(hipMemcpyAsync(dst, hst, 8 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
      k_rounds_topo<<<1, 1024, 0, s>>>(a, t.keys.as<uint64_t>(), (uint32_t)t.capacity, dst, dst + 8, max_rounds);
      GX_HIP(hipGetLastError());
      GX_HIP(hipMemcpyAsync(hst, dst, (8 + max_rounds) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      // the rounds' geometry, level by level (each reads its children's)
      for (uint32_t k = 0, pe = (uint32_t)ce; k < hst[4]; ++k) {
        const uint32_t ne = hst[8 + k], np = ne - pe;
        k_round_parents<<<nblk(np), 256, 0, s>>>(a, pe, np, (uint8_t)(lvl + 1 + k), constrained_style ? 1 : 0);
        GX_HIP(hipGetLastError());
        t.level_offset.push_back(ne);
        pe = ne;
      }
      cs = hst[0];
      ce = hst[1];
      shift = hst[2];
      lvl = hst[3];
      if (lvl > 250) return set_error(GALAX_ERR_INVALID, "tree deeper than 250 levels");
      if (ce - cs <= 1) break;
      c = (uint32_t)(ce - cs);
      if (c <= FUSED_ROUND && hst[4] == 0) {
        GX_CHECK(ensure_node_capacity(h, ce + (size_t)c + 64));  // room for one more round, then fused again
        a = node_arrays(h);
      }
    }
    ++lvl;
    if (lvl > 250) return set_error(GALAX_ERR_INVALID, "tree deeper than 250 levels");
    uint32_t* head = h->ws.e.as<uint32_t>();
    uint32_t* flag = h->ws.f.as<uint32_t>();
    uint32_t* pidx = h->ws.g.as<uint32_t>();
    k_round_heads<<<nblk(c), 256, 0, s>>>(t.keys.as<uint64_t>(), a.bs, (uint32_t)cs, c, shift, head);
    GX_HIP(hipGetLastError());
    GX_CHECK(scan_u32(h, head, head, c, true, true, nullptr));
    k_round_chunks<<<nblk(c), 256, 0, s>>>(head, c, flag);
    GX_HIP(hipGetLastError());
    uint32_t np = 0;
    GX_CHECK(scan_u32(h, flag, pidx, c, false, false, &np));
    GX_CHECK(ensure_node_capacity(h, ce + np));
    a = node_arrays(h);
    k_round_mark<<<nblk(c), 256, 0, s>>>(flag, pidx, (uint32_t)cs, (uint32_t)ce, c, a.cf);
    GX_HIP(hipGetLastError());
    k_round_parents<<<nblk(np), 256, 0, s>>>(a, (uint32_t)ce, np, (uint8_t)lvl, constrained_style ? 1 : 0);
    GX_HIP(hipGetLastError());
    cs = ce;
    ce += np;
    t.level_offset.push_back((uint32_t)ce);
    shift += 2;
  }
  t.n_nodes = ce;
  t.n_leaves = n_leaves;
  t.n_bodies = n;
  t.n_levels = (uint32_t)t.level_offset.size() - 1;
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

// Tree::build: keys, stable sort, permute the 7 arrays, re-key, leaves, rounds.
int build_tree_plain(galax_fmm* h, size_t n) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const galax_fmm_config_t& c = h->cfg;
  t.n_nodes = t.n_leaves = 0;
  t.n_levels = 0;
  t.n_bodies = n;
  t.chain_ok.clear();
  t.level_offset.assign(1, 0);
  if (n == 0) return 0;
  GX_CHECK(dbuf_reserve(t.keys, n * sizeof(uint64_t), s));
  GX_CHECK(dbuf_reserve(h->ws.a, n * sizeof(uint64_t), s));
  GX_CHECK(dbuf_reserve(h->ws.b, n * sizeof(uint32_t) * 2, s));
  uint64_t* ka = t.keys.as<uint64_t>();
  uint64_t* kb = h->ws.a.as<uint64_t>();
  uint32_t* va = h->ws.b.as<uint32_t>();
  uint32_t* vb = va + n;
  k_morton<<<nblk(n), 256, 0, s>>>(h->x.as<double>(), h->y.as<double>(), n, c.x_min, c.x_max, c.y_min, c.y_max, ka,
                                   va);
  GX_HIP(hipGetLastError());
  GX_CHECK(radix_sort(h, ka, va, kb, vb, n));
  GX_CHECK(build_mark(h, 1));
  // permutation: new[i] = old[va[i]]; kept for galax_fmm_build_tree's perm_out
  GX_CHECK(dbuf_reserve(h->perm, n * sizeof(uint32_t), s));
  GX_HIP(hipMemcpyAsync(h->perm.p, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  GX_CHECK(dbuf_reserve(h->ws.h, n * sizeof(double), s));
  double* tmp = h->ws.h.as<double>();
  DBuf* arrs[7] = {&h->x, &h->y, &h->vx, &h->vy, &h->m, &h->ax, &h->ay};
  for (DBuf* b : arrs) {
    if (!b->p) continue;
    k_gather_f64<<<nblk(n), 256, 0, s>>>(b->as<double>(), va, n, tmp);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(b->p, tmp, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  GX_CHECK(build_mark(h, 2));
  // Re-encoding the permuted bodies reproduces the sorted keys exactly (the
  // encoding is a pure function of the body), so ka already holds them.
  size_t n_leaves = (n + c.n_max - 1) / c.n_max;
  GX_CHECK(dbuf_reserve(h->ws.c, (n_leaves + 1) * sizeof(uint32_t), s));
  k_uniform_leaf_starts<<<nblk(n_leaves + 1), 256, 0, s>>>(h->ws.c.as<uint32_t>(), (uint32_t)n_leaves, c.n_max,
                                                           (uint32_t)n);
  GX_HIP(hipGetLastError());
  return build_from_leaf_starts(h, n, n_leaves, false);
}

// ------------------------------------------- constrained tree (A6/A7) ----

// neighbors_geometric (galax-core/src/lib.rs:382-423) for one leaf and all 8
// test points, as a pruned traversal: nodes whose box contains the point and
// that are cells_adjacent to the leaf; the nearest centre wins, the lowest id
// among equal distances (the reference scans ids in ascending order with `<`).
struct UBox {
  const double *x0, *x1, *y0, *y1;
};

__global__ void k_ubox(NodeArrays a, uint32_t lo, uint32_t hi, double* ux0, double* ux1, double* uy0, double* uy1) {
  uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  double x0 = a.cx[i] - a.hw[i], x1 = a.cx[i] + a.hw[i], y0 = a.cy[i] - a.hw[i], y1 = a.cy[i] + a.hw[i];
  uint32_t cf = a.cf[i];
  if (cf != NONE) {
    for (uint32_t c = cf; c < cf + a.cc[i]; ++c) {
      x0 = fmin(x0, ux0[c]);
      x1 = fmax(x1, ux1[c]);
      y0 = fmin(y0, uy0[c]);
      y1 = fmax(y1, uy1[c]);
    }
  }
  ux0[i] = x0;
  ux1[i] = x1;
  uy0[i] = y0;
  uy1[i] = y1;
}

static int compute_uboxes(galax_fmm* h) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  size_t nn = t.n_nodes;
  GX_CHECK(dbuf_reserve(h->ws.h, nn * 4 * sizeof(double), s));
  double* u = h->ws.h.as<double>();
  NodeArrays a = node_arrays(h);
  for (uint32_t l = 0; l < t.n_levels; ++l) {
    uint32_t lo = t.level_offset[l], hi = t.level_offset[l + 1];
    k_ubox<<<nblk(hi - lo), 256, 0, s>>>(a, lo, hi, u, u + nn, u + 2 * nn, u + 3 * nn);
    GX_HIP(hipGetLastError());
  }
  return 0;
}

constexpr int DFS_STACK = 256;

__device__ __forceinline__ bool cells_adjacent(const NodeArrays& a, uint32_t i, uint32_t j) {
  double hi = a.hw[i], hj = a.hw[j];
  double ratio = fmax(hi, hj) / fmin(hi, hj);
  if (ratio > 8.0) return false;
  double dx = fabs(a.cx[i] - a.cx[j]), dy = fabs(a.cy[i] - a.cy[j]);
  return dx <= hi + hj + 1e-15 && dy <= hi + hj + 1e-15;
}

// For each leaf: does any geometric neighbour have hw*2.1 < this leaf's hw?
// (the 2:1 check of build_constrained, galax-core/src/lib.rs:558-578)
__global__ void k_constrained_check(NodeArrays a, uint32_t n_leaves, uint32_t root, const double* __restrict__ ux0,
                                    const double* __restrict__ ux1, const double* __restrict__ uy0,
                                    const double* __restrict__ uy1, uint32_t* __restrict__ violate,
                                    uint32_t* __restrict__ err) {
  uint32_t lid = blockIdx.x * blockDim.x + threadIdx.x;
  if (lid >= n_leaves) return;
  double cx = a.cx[lid], cy = a.cy[lid], hw = a.hw[lid];
  const double px[8] = {cx + hw * 2.0, cx + hw * 2.0, cx, cx - hw * 2.0, cx - hw * 2.0, cx - hw * 2.0, cx, cx + hw * 2.0};
  const double py[8] = {cy, cy + hw * 2.0, cy + hw * 2.0, cy + hw * 2.0, cy, cy - hw * 2.0, cy - hw * 2.0, cy - hw * 2.0};
  uint32_t result = 0;
  for (int k = 0; k < 8; ++k) {
    uint32_t best = NONE;
    double best_d = 1.7976931348623157e308;
    uint32_t stack[DFS_STACK];
    int sp = 0;
    stack[sp++] = root;
    while (sp > 0) {
      uint32_t nd = stack[--sp];
      // prune with the subtree box (conservative slack on top of the 1e-15)
      double tol = 1e-15 + 1e-12 * (fabs(px[k]) + fabs(py[k]) + 1.0);
      if (ux0[nd] > px[k] + tol || ux1[nd] < px[k] - tol || uy0[nd] > py[k] + tol || uy1[nd] < py[k] - tol) continue;
      if (nd != lid) {
        double dx = fabs(px[k] - a.cx[nd]), dy = fabs(py[k] - a.cy[nd]);
        if (dx <= a.hw[nd] + 1e-15 && dy <= a.hw[nd] + 1e-15 && cells_adjacent(a, lid, nd)) {
          double ex = px[k] - a.cx[nd], ey = py[k] - a.cy[nd];
          double dist = ex * ex + ey * ey;
          if (dist < best_d || (dist == best_d && nd < best)) {
            best_d = dist;
            best = nd;
          }
        }
      }
      uint32_t cf = a.cf[nd];
      if (cf != NONE) {
        for (int c = (int)a.cc[nd] - 1; c >= 0; --c) {
          if (sp < DFS_STACK)
            stack[sp++] = cf + (uint32_t)c;
          else
            atomicOr(err, 1u);
        }
      }
    }
    if (best != NONE && hw > a.hw[best] * 2.1) {
      result = 1;
      break;  // one split per leaf per iteration
    }
  }
  violate[lid] = result;
}

// build_constrained (galax-core/src/lib.rs:541-588): up to 10 rebuilds on the
// current body order; a leaf of more than 2 bodies with a geometric neighbour
// of less than 1/2.1 its half width is split at (start+end)/2.
int build_tree_constrained(galax_fmm* h, size_t n) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const galax_fmm_config_t& c = h->cfg;
  t.n_nodes = t.n_leaves = 0;
  t.n_levels = 0;
  t.n_bodies = n;
  t.chain_ok.clear();
  t.level_offset.assign(1, 0);
  if (n == 0) return 0;
  GX_CHECK(dbuf_reserve(t.keys, n * sizeof(uint64_t), s));
  k_morton<<<nblk(n), 256, 0, s>>>(h->x.as<double>(), h->y.as<double>(), n, c.x_min, c.x_max, c.y_min, c.y_max,
                                   t.keys.as<uint64_t>(), nullptr);
  GX_HIP(hipGetLastError());
  std::vector<uint64_t> splits;
  std::vector<uint32_t> starts;
  auto make_starts = [&]() {
    starts.clear();
    uint64_t st = 0;
    while (st < n) {
      uint64_t bound = std::min<uint64_t>(st + c.n_max, n);
      uint64_t e = bound;
      auto it = std::upper_bound(splits.begin(), splits.end(), st);
      if (it != splits.end() && *it < bound) e = *it;
      starts.push_back((uint32_t)st);
      st = e;
    }
    starts.push_back((uint32_t)n);
  };
  for (int iter = 0; iter <= 10; ++iter) {
    make_starts();
    size_t nl = starts.size() - 1;
    GX_CHECK(dbuf_reserve(h->ws.c, starts.size() * sizeof(uint32_t), s));
    GX_HIP(hipMemcpyAsync(h->ws.c.p, starts.data(), starts.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    GX_CHECK(build_from_leaf_starts(h, n, nl, true));
    if (iter == 10) break;
    GX_CHECK(compute_uboxes(h));
    GX_CHECK(dbuf_reserve(h->ws.d, (nl + 16) * sizeof(uint32_t), s));
    uint32_t* viol = h->ws.d.as<uint32_t>();
    uint32_t* err = viol + nl;
    GX_HIP(hipMemsetAsync(err, 0, sizeof(uint32_t), s));
    double* u = h->ws.h.as<double>();
    size_t nn = t.n_nodes;
    NodeArrays a = node_arrays(h);
    k_constrained_check<<<nblk(nl, 128), 128, 0, s>>>(a, (uint32_t)nl, (uint32_t)(nn - 1), u, u + nn, u + 2 * nn,
                                                  u + 3 * nn, viol, err);
    GX_HIP(hipGetLastError());
    std::vector<uint32_t> hv(nl + 1);
    GX_HIP(hipMemcpyAsync(hv.data(), viol, (nl + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (hv[nl]) return set_error(GALAX_ERR_INVALID, "constrained build: traversal stack overflow");
    std::vector<uint64_t> new_splits;
    for (size_t l = 0; l < nl; ++l) {
      if (!hv[l]) continue;
      uint64_t st = starts[l], e = starts[l + 1];
      if (e - st > 2) {
        uint64_t mid = (st + e) / 2;
        if (!std::binary_search(splits.begin(), splits.end(), mid) &&
            std::find(new_splits.begin(), new_splits.end(), mid) == new_splits.end())
          new_splits.push_back(mid);
      }
    }
    if (new_splits.empty()) break;
    splits.insert(splits.end(), new_splits.begin(), new_splits.end());
    std::sort(splits.begin(), splits.end());
    splits.erase(std::unique(splits.begin(), splits.end()), splits.end());
  }
  return 0;
}

// ------------------------------------------------------------ lists ----

// ------------------------------------------- list candidates, pruned ----
//
// InteractionLists::build (galax-core/src/lib.rs:972-1016) tests, for a node
// c with parent P, every node pn of level(c) + 1 in ascending id, and appends
// each child q of a pn that touches P when q does not touch c, until 64
// entries; build_p2p_lists (:1371-1389) tests every leaf against a leaf.
// Scanned directly that is O(sum of n_l n_(l+1)): at 2^20 Plummer bodies
// ~4.6e9 candidates, most producing nothing (from height 7 up every child
// touches every node: the lists are empty, yet each level is scanned).
//
// Here every level L has a segment tree over its nodes in id order (a heap of
// 2 P2 entries, P2 = 2^ceil(log2 n_L)). A segment stores
//   the union of its nodes' boxes [cx - hw, cx + hw] (x and y), and
//   over all children q of its nodes: max (q.c - q.hw) and min (q.c + q.hw).
// A node touches P only if its box meets P's box widened by 1e-15; a child
// q fails to touch c only if q.c - q.hw > c.c + c.hw + 1e-15 or
// q.c + q.hw < c.c - c.hw - 1e-15 (x or y). A query walks the tree depth
// first in ascending id, skips segments where either condition cannot hold
// (with a slack of 1e-13 of the magnitudes, far above the rounding of the
// reference's own comparisons), applies the reference's exact tests to single
// nodes and stops at 64 entries: the same entries in the same order, visiting
// only segments that can contribute.
struct SegLevel {
  uint32_t base, p2, lo, n;  // heap entries [base, base + 2 p2); node ids [lo, lo + n)
};
// One 32-byte record per segment (a query reads it with two 16-byte loads):
// the box union x0, x1, y0, y1, then the children's extremes
// ax = max(q.c - q.hw), bx = min(q.c + q.hw) in x, ay, by in y. In FP32,
// each bound rounded outwards (minima down, maxima up), so a record holds
// at least what the FP64 one would and the pruning stays conservative.
struct SegRec {
  float x0, x1, y0, y1, ax, bx, ay, by;
};
__device__ __forceinline__ SegRec seg_rec(double x0, double x1, double y0, double y1, double ax, double bx, double ay,
                                          double by) {
  return SegRec{__double2float_rd(x0), __double2float_ru(x1), __double2float_rd(y0), __double2float_ru(y1),
                __double2float_ru(ax), __double2float_rd(bx), __double2float_ru(ay), __double2float_rd(by)};
}
struct SegTree {
  SegRec* r;
};

__global__ void k_seg_leaves(NodeArrays a, const SegLevel* __restrict__ lvls, uint32_t nlv, SegTree st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // one thread per heap leaf slot of every level: levels in order, slots [0, p2)
  uint32_t l = 0, off = 0;
  while (l < nlv && i >= off + lvls[l].p2) off += lvls[l++].p2;
  if (l >= nlv) return;
  const SegLevel L = lvls[l];
  const uint32_t k = i - off;
  const size_t e = (size_t)L.base + L.p2 + k;
  const double inf = __builtin_huge_val();
  if (k >= L.n) {  // padding: meets nothing, no child
    st.r[e] = seg_rec(inf, -inf, inf, -inf, -inf, inf, -inf, inf);
    return;
  }
  const uint32_t nd = L.lo + k;
  const double cx = a.cx[nd], cy = a.cy[nd], hw = a.hw[nd];
  double ax = -inf, bx = inf, ay = -inf, by = inf;
  const uint32_t cf = a.cf[nd];
  if (cf != NONE)
    for (uint32_t q = cf; q < cf + a.cc[nd]; ++q) {
      const double qx = a.cx[q], qy = a.cy[q], qh = a.hw[q];
      ax = fmax(ax, qx - qh);
      bx = fmin(bx, qx + qh);
      ay = fmax(ay, qy - qh);
      by = fmin(by, qy + qh);
    }
  st.r[e] = seg_rec(cx - hw, cx + hw, cy - hw, cy + hw, ax, bx, ay, by);
}

// internal entries at height h (1 = parents of the leaf slots) of every level
// whose tree is that tall
__global__ void k_seg_up(const SegLevel* __restrict__ lvls, uint32_t nlv, uint32_t h, SegTree st) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t l = 0; l < nlv; ++l) {
    const uint32_t cnt = lvls[l].p2 >> h;
    if (i < cnt) {
      const size_t e = (size_t)lvls[l].base + cnt + i;
      const SegRec A = st.r[(size_t)lvls[l].base + 2 * (cnt + i)], B = st.r[(size_t)lvls[l].base + 2 * (cnt + i) + 1];
      st.r[e] = SegRec{fminf(A.x0, B.x0), fmaxf(A.x1, B.x1), fminf(A.y0, B.y0), fmaxf(A.y1, B.y1),
                       fmaxf(A.ax, B.ax), fminf(A.bx, B.bx), fmaxf(A.ay, B.ay), fminf(A.by, B.by)};
      return;
    }
    i -= cnt;
  }
}

constexpr double SEG_SLACK = 1e-13;

// Can the segment hold a node whose box meets [x0, x1] x [y0, y1]?
__device__ __forceinline__ bool seg_meets(const SegTree& st, size_t e, double x0, double x1, double y0, double y1,
                                          double mag) {
  const float4 b = *reinterpret_cast<const float4*>(&st.r[e].x0);
  const double sx0 = b.x, sx1 = b.y, sy0 = b.z, sy1 = b.w;
  if (!(sx0 <= sx1)) return false;  // empty
  const double sl = SEG_SLACK * (mag + fmax(fmax(fabs(sx0), fabs(sx1)), fmax(fabs(sy0), fabs(sy1))));
  return sx1 >= x0 - sl && sx0 <= x1 + sl && sy1 >= y0 - sl && sy0 <= y1 + sl;
}

// Can the segment hold a child that does not touch c (centre cx, cy, half
// width ch)?
__device__ __forceinline__ bool seg_far_child(const SegTree& st, size_t e, double cx, double cy, double ch,
                                              double mag) {
  const float4 c = *reinterpret_cast<const float4*>(&st.r[e].ax);
  const double ax = c.x, bx = c.y, ay = c.z, by = c.w;
  if (ax == -__builtin_huge_val()) return false;  // no child in the segment
  const double sl = SEG_SLACK * (mag + fmax(fmax(fabs(ax), fabs(bx)), fmax(fabs(ay), fabs(by))));
  const double r = ch + 1e-15;
  return ax > cx + r - sl || bx < cx - r + sl || ay > cy + r - sl || by < cy - r + sl;
}

// The walk needs no stack: in a heap (children of s are 2s and 2s + 1) the
// subtree after s in depth-first, ascending order starts at the right
// sibling of s or of its nearest ancestor that is a left child; 0 ends it.
__device__ __forceinline__ uint32_t seg_next(uint32_t s) {
  while (s & 1u) s >>= 1;
  return s ? s + 1 : 0u;
}

// M2L row of node c (below the root); rows of nodes c >= nwalk are left empty.
__device__ __forceinline__ void m2l_row_seg(const NodeArrays& a, uint32_t nn, uint32_t nwalk,
                                            const SegLevel* __restrict__ lvls, uint32_t nlv, const SegTree& st,
                                            uint32_t* __restrict__ rows, uint32_t* __restrict__ lens, uint32_t c) {
  if (c >= nn) return;
  if (c >= nwalk) {
    lens[c] = 0;
    return;
  }
  const uint32_t P = a.par[c];
  const uint32_t lc = a.lv[c];
  if (P == NONE || lc + 1 >= nlv) {
    lens[c] = 0;
    return;
  }
  const SegLevel L = lvls[lc + 1];
  const double ccx = a.cx[c], ccy = a.cy[c], chw = a.hw[c];
  const double pcx = a.cx[P], pcy = a.cy[P], phw = a.hw[P];
  const double magp = 1.0 + fabs(pcx) + fabs(pcy) + phw;
  const double magc = 1.0 + fabs(ccx) + fabs(ccy) + chw;
  const double px0 = pcx - phw - 1e-15, px1 = pcx + phw + 1e-15, py0 = pcy - phw - 1e-15, py1 = pcy + phw + 1e-15;
  uint32_t* row = rows + (size_t)c * LIST_CAP;
  uint32_t len = 0;
  // entries go out four at a time (16-byte stores; rows are 256-byte aligned
  // and LIST_CAP is a multiple of 4): a thread's stores are far apart in
  // time, and single 4-byte ones each cost a partial-line write to HBM
  uint32_t b0 = 0, b1 = 0, b2 = 0;
  auto put = [&](uint32_t q) {
    const uint32_t r = len & 3u;
    if (r == 3u) *reinterpret_cast<uint4*>(row + (len - 3u)) = make_uint4(b0, b1, b2, q);
    b0 = r == 0u ? q : b0;
    b1 = r == 1u ? q : b1;
    b2 = r == 2u ? q : b2;
    ++len;
  };
  uint32_t s = 1;
  while (s && len < LIST_CAP) {
    if (s >= L.p2) {  // a single node
      const uint32_t k = s - L.p2;
      const uint32_t pn = L.lo + k;
      s = seg_next(s);
      if (k >= L.n || pn == P) continue;
      // galax-core/src/lib.rs:991-997
      const double dx = fabs(a.cx[pn] - pcx), dy = fabs(a.cy[pn] - pcy);
      const double h_sum = phw + a.hw[pn];
      if (dx > h_sum + 1e-15 || dy > h_sum + 1e-15) continue;
      const uint32_t qf = a.cf[pn], qc = a.cc[pn];
      for (uint32_t q = qf; qf != NONE && q < qf + qc && len < LIST_CAP; ++q) {
        // :999-1006
        const double cdx = fabs(a.cx[q] - ccx), cdy = fabs(a.cy[q] - ccy);
        const double hs = chw + a.hw[q];
        if (cdx <= hs + 1e-15 && cdy <= hs + 1e-15) continue;
        put(q);
      }
      continue;
    }
    const size_t e = (size_t)L.base + s;
    s = (seg_meets(st, e, px0, px1, py0, py1, magp) && seg_far_child(st, e, ccx, ccy, chw, magc)) ? 2 * s
                                                                                                   : seg_next(s);
  }
  const uint32_t r = len & 3u, l0 = len - r;  // the last, partial group
  if (r > 0u) row[l0] = b0;
  if (r > 1u) row[l0 + 1] = b1;
  if (r > 2u) row[l0 + 2] = b2;
  lens[c] = len;
}

// P2P row of leaf c: the leaf itself first, then every other leaf whose box
// touches it, ascending (galax-core/src/lib.rs:1371-1389).
__device__ __forceinline__ void p2p_row_seg(const NodeArrays& a, uint32_t n_leaves, const SegLevel* __restrict__ lvls,
                                            const SegTree& st, uint32_t* __restrict__ rows,
                                            uint32_t* __restrict__ lens, uint32_t c) {
  if (c >= n_leaves) return;
  const SegLevel L = lvls[0];
  const double cx = a.cx[c], cy = a.cy[c], hw = a.hw[c];
  const double mag = 1.0 + fabs(cx) + fabs(cy) + hw;
  const double x0 = cx - hw - 1e-15, x1 = cx + hw + 1e-15, y0 = cy - hw - 1e-15, y1 = cy + hw + 1e-15;
  uint32_t* row = rows + (size_t)c * LIST_CAP;
  row[0] = c;
  uint32_t len = 1;
  uint32_t s = 1;
  while (s && len < LIST_CAP) {
    if (s >= L.p2) {
      const uint32_t k = s - L.p2, b = L.lo + k;
      s = seg_next(s);
      if (k >= L.n || b == c) continue;
      const double dx = fabs(cx - a.cx[b]), dy = fabs(cy - a.cy[b]);
      const double hs = hw + a.hw[b];
      if (dx <= hs + 1e-15 && dy <= hs + 1e-15) row[len++] = b;
      continue;
    }
    s = seg_meets(st, (size_t)L.base + s, x0, x1, y0, y1, mag) ? 2 * s : seg_next(s);
  }
  lens[c] = len;
}

// Both list builds in one launch, a thread per row: the first nbp blocks walk
// the leaves' P2P rows (few rows, long walks: they start first), the rest the
// nodes' M2L rows (walked for nodes below nwalk). Each build has its own rows
// and lengths.
__global__ void __launch_bounds__(256) k_rows_seg(NodeArrays a, uint32_t nn, uint32_t nwalk, uint32_t n_leaves,
                                                  uint32_t nbp, const SegLevel* __restrict__ lvls, uint32_t nlv,
                                                  SegTree st, uint32_t* __restrict__ mrows, uint32_t* __restrict__ mlens,
                                                  uint32_t* __restrict__ prows, uint32_t* __restrict__ plens) {
  if (blockIdx.x < nbp)
    p2p_row_seg(a, n_leaves, lvls, st, prows, plens, blockIdx.x * blockDim.x + threadIdx.x);
  else
    m2l_row_seg(a, nn, nwalk, lvls, nlv, st, mrows, mlens, (blockIdx.x - nbp) * blockDim.x + threadIdx.x);
}

__global__ void k_rows_compact(const uint32_t* __restrict__ rows, const uint32_t* __restrict__ off, uint32_t nn,
                               uint32_t* __restrict__ dat) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t i = (uint32_t)(t / LIST_CAP), k = (uint32_t)(t % LIST_CAP);
  if (i >= nn) return;
  if (k < off[i + 1] - off[i]) dat[off[i] + k] = rows[t];
}

// Symmetric P2P gather lists from the reference lists: leaf t gathers from
// itself, from b in list(t) with b > t, and from every a < t with t in list(a)
// (galax-core/src/lib.rs:1415-1437 computes pair (a,b) iff b in list(a), a < b).
__global__ void k_gather_count(uint32_t n_leaves, const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                               uint32_t* __restrict__ gcnt) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_leaves) return;
  uint32_t own = 0;
  for (uint32_t k = off[a]; k < off[a + 1]; ++k) {
    uint32_t b = dat[k];
    if (b > a) {
      ++own;
      atomicAdd(&gcnt[b], 1u);
    }
  }
  atomicAdd(&gcnt[a], own + 1u);
}

__global__ void k_gather_fill(uint32_t n_leaves, const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                              const uint32_t* __restrict__ goff, uint32_t* __restrict__ cursor,
                              uint32_t* __restrict__ gdat) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_leaves) return;
  gdat[goff[a]] = a;
  for (uint32_t k = off[a]; k < off[a + 1]; ++k) {
    uint32_t b = dat[k];
    if (b > a) {
      uint32_t pa = atomicAdd(&cursor[a], 1u);
      gdat[goff[a] + 1 + pa] = b;
      uint32_t pb = atomicAdd(&cursor[b], 1u);
      gdat[goff[b] + 1 + pb] = a;
    }
  }
}


// Symmetric P2P gather lists from the reference lists: leaf t gathers from
// itself, from b in list(t) with b > t, and from every a < t with t in list(a)
// (galax-core/src/lib.rs:1415-1437 computes pair (a,b) iff b in list(a), a < b).
// Every directed pair (t, s) becomes a key t * 2^w + s; one radix sort gives
// the CSR in ascending (t, s) order. A leaf can sit in the lists of thousands
// of smaller leaves (a large outer box touching a dense core), so the lists
// are unbounded and are not sorted per segment.
__global__ void k_pair_count(uint32_t n_leaves, const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                             uint32_t* __restrict__ cnt) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_leaves) return;
  uint32_t c = 1;  // self
  for (uint32_t k = off[a]; k < off[a + 1]; ++k) c += dat[k] > a ? 2u : 0u;
  cnt[a] = c;
}

__global__ void k_pair_emit(uint32_t n_leaves, const uint32_t* __restrict__ off, const uint32_t* __restrict__ dat,
                            const uint32_t* __restrict__ pos, int w, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_leaves) return;
  uint32_t p = pos[a];
  keys[p] = ((uint64_t)a << w) | a;
  vals[p++] = 0;
  for (uint32_t k = off[a]; k < off[a + 1]; ++k) {
    uint32_t b = dat[k];
    if (b > a) {
      keys[p] = ((uint64_t)a << w) | b;
      vals[p++] = 0;
      keys[p] = ((uint64_t)b << w) | a;
      vals[p++] = 0;
    }
  }
}

// The sorted keys as CSR: dat[i] = source of key i; off[u] = the first i
// whose target is u (key i starts every leaf u after the previous key's
// target, up to its own), off[n_leaves] = n. (Counting with atomics instead
// serialises: neighbouring keys share their target.)
__global__ void k_pair_split(const uint64_t* __restrict__ keys, uint32_t n, int w, uint32_t n_leaves,
                             uint32_t* __restrict__ dat, uint32_t* __restrict__ off) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  dat[i] = (uint32_t)(k & ((1ull << w) - 1));
  const uint32_t t = (uint32_t)(k >> w);
  const uint32_t u0 = i ? (uint32_t)(keys[i - 1] >> w) + 1u : 0u;
  for (uint32_t u = u0; u <= t; ++u) off[u] = i;
  if (i == n - 1)
    for (uint32_t u = t + 1; u <= n_leaves; ++u) off[u] = n;
}

// Directed body-body interactions of the gather lists (self pairs excluded).
// One thread per gather entry (t, s): n_t n_s directed pairs, less the
// body itself when s = t; its leaf t found by binary search of the offsets.
// (A thread per leaf left the longest lists, thousands of entries, to one
// thread.) Block sums, then one 64-bit atomic add per block.
__global__ void __launch_bounds__(256) k_count_interactions(uint32_t n_leaves, const uint32_t* __restrict__ goff,
                                                            const uint32_t* __restrict__ gdat,
                                                            const uint32_t* __restrict__ bs,
                                                            const uint32_t* __restrict__ be,
                                                            unsigned long long* __restrict__ total) {
  __shared__ unsigned long long part[4];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (g < goff[n_leaves]) {
    uint32_t lo = 0, hi = n_leaves;  // the leaf t with goff[t] <= g < goff[t + 1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (goff[mid] <= g)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t t = lo, src = gdat[g];
    const unsigned long long nt = be[t] - bs[t];
    v = nt * (unsigned long long)(be[src] - bs[src]) - (src == t ? nt : 0ull);
  }
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

int count_interactions(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint64_t* out) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  GX_CHECK(dbuf_reserve(h->flag, 64, s));
  unsigned long long* d = reinterpret_cast<unsigned long long*>(h->flag.as<uint32_t>() + 12);
  GX_HIP(hipMemsetAsync(d, 0, 8, s));
  uint32_t ng = 0;  // gather entries
  GX_HIP(hipMemcpyAsync(&ng, goff + t.n_leaves, 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  if (ng)
    k_count_interactions<<<nblk(ng), 256, 0, s>>>((uint32_t)t.n_leaves, goff, gdat, t.body_start.as<uint32_t>(),
                                                  t.body_end.as<uint32_t>(), d);
  GX_HIP(hipGetLastError());
  unsigned long long v = 0;
  GX_HIP(hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  *out = v;
  return 0;
}

static int bits_for(uint64_t n) {
  int w = 1;
  while (w < 32 && (1ull << w) < n) ++w;
  return w;
}

__global__ void k_items_count(const uint32_t* __restrict__ goff, uint32_t nl, uint32_t* __restrict__ nit,
                              uint32_t* __restrict__ npart, uint32_t* __restrict__ ismulti) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  uint32_t len = goff[t + 1] - goff[t];
  uint32_t n = (len + P2P_ITEM_LEAVES - 1) / P2P_ITEM_LEAVES;
  if (n == 0) n = 1;
  nit[t] = n;
  npart[t] = n > 1 ? n : 0u;
  ismulti[t] = n > 1 ? 1u : 0u;
}

// One thread per item i: its leaf t (ioff[t] <= i < ioff[t + 1], every leaf
// has an item) by binary search, so a leaf with very many items is not left
// to one thread.
__global__ void k_items_fill(const uint32_t* __restrict__ goff, uint32_t nl, uint32_t ni,
                             const uint32_t* __restrict__ ioff, const uint32_t* __restrict__ poff,
                             const uint32_t* __restrict__ moff, uint32_t* __restrict__ tgt, uint32_t* __restrict__ g0,
                             uint32_t* __restrict__ g1, uint32_t* __restrict__ slot, uint32_t* __restrict__ multi) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  uint32_t lo = 0, hi = nl;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (ioff[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t t = lo, k = i - ioff[t], n = ioff[t + 1] - ioff[t];
  const uint32_t a = goff[t] + k * P2P_ITEM_LEAVES;
  tgt[i] = t;
  g0[i] = a;
  g1[i] = min(a + P2P_ITEM_LEAVES, goff[t + 1]);
  slot[i] = n > 1 ? poff[t] + k : NONE;
  if (k == 0 && n > 1) multi[moff[t]] = t;
}

// Work class of an item: target bodies in 32-body passes times source bodies,
// as a power of two.
__global__ void k_items_cost(const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ g0,
                             const uint32_t* __restrict__ g1, const uint32_t* __restrict__ gdat,
                             const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be, uint32_t ni,
                             uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const uint32_t t = tgt[i];
  uint64_t ns = 0;
  for (uint32_t g = g0[i]; g < g1[i]; ++g) ns += be[gdat[g]] - bs[gdat[g]];
  const uint64_t cost = (uint64_t)((be[t] - bs[t] + 31u) / 32u) * ns;
  // power-of-two cost classes, costliest first; the stable sort keeps the
  // spatial order within a class (neighbouring items share sources in L2)
  const uint32_t cls = cost ? 64u - (uint32_t)__clzll((unsigned long long)cost) : 0u;
  key[i] = 64u - cls;
  val[i] = i;
}

__global__ void k_items_permute(const uint32_t* __restrict__ ord, uint32_t ni, const uint32_t* __restrict__ a0,
                                const uint32_t* __restrict__ a1, const uint32_t* __restrict__ a2,
                                const uint32_t* __restrict__ a3, uint32_t* __restrict__ b) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ni) return;
  const uint32_t i = ord[j];
  b[j] = a0[i];
  b[ni + j] = a1[i];
  b[2 * ni + j] = a2[i];
  b[3 * ni + j] = a3[i];
}

int build_p2p_items(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t nl, P2PItems& it) {
  hipStream_t s = h->stream;
  GX_CHECK(dbuf_reserve(it.part_off, (size_t)(nl + 1) * 4, s));
  GX_CHECK(dbuf_reserve(h->ws.f, (size_t)(nl + 1) * 8, s));
  uint32_t* ioff = h->ws.f.as<uint32_t>();
  uint32_t* moff = ioff + nl + 1;
  uint32_t* poff = it.part_off.as<uint32_t>();
  GX_HIP(hipMemsetAsync(ioff, 0, (size_t)(nl + 1) * 8, s));
  GX_HIP(hipMemsetAsync(poff, 0, (size_t)(nl + 1) * 4, s));
  k_items_count<<<nblk(nl), 256, 0, s>>>(goff, nl, ioff, poff, moff);
  GX_HIP(hipGetLastError());
  uint32_t ni = 0, np = 0, nm = 0;
  GX_CHECK(scan_u32(h, ioff, ioff, nl + 1, false, false, &ni));
  GX_CHECK(scan_u32(h, poff, poff, nl + 1, false, false, &np));
  GX_CHECK(scan_u32(h, moff, moff, nl + 1, false, false, &nm));
  it.n_items = ni;
  it.n_part = np;
  it.n_multi = nm;
  GX_CHECK(dbuf_reserve(it.tgt, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(it.g0, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(it.g1, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(it.slot, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(it.multi, (size_t)nm * 4 + 4, s));
  if (ni) {
    k_items_fill<<<nblk(ni), 256, 0, s>>>(goff, nl, ni, ioff, poff, moff, it.tgt.as<uint32_t>(), it.g0.as<uint32_t>(),
                                          it.g1.as<uint32_t>(), it.slot.as<uint32_t>(), it.multi.as<uint32_t>());
    GX_HIP(hipGetLastError());
  }
  // items by decreasing power-of-two work class (stable): the costly items of
  // dense regions start first instead of forming the last round of blocks.
  // Slots, and so the combine's summation order, do not depend on the order
  // items run in.
  if (ni > 1) {
    Workspace& ws = h->ws;
    GX_CHECK(dbuf_reserve(ws.sort_ka, (size_t)ni * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_kb, (size_t)ni * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_va, (size_t)ni * 4, s));
    GX_CHECK(dbuf_reserve(ws.sort_vb, (size_t)ni * 4, s));
    k_items_cost<<<nblk(ni), 256, 0, s>>>(it.tgt.as<uint32_t>(), it.g0.as<uint32_t>(), it.g1.as<uint32_t>(), gdat,
                                          h->tree.body_start.as<uint32_t>(), h->tree.body_end.as<uint32_t>(), ni,
                                          ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>());
    GX_HIP(hipGetLastError());
    GX_CHECK(radix_sort_bits(h, ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>(), ws.sort_kb.as<uint64_t>(),
                             ws.sort_vb.as<uint32_t>(), ni, 7));
    GX_CHECK(dbuf_reserve(ws.e, (size_t)ni * 16, s));  // free here in both list builds
    uint32_t* b = ws.e.as<uint32_t>();
    k_items_permute<<<nblk(ni), 256, 0, s>>>(ws.sort_va.as<uint32_t>(), ni, it.tgt.as<uint32_t>(),
                                             it.g0.as<uint32_t>(), it.g1.as<uint32_t>(), it.slot.as<uint32_t>(), b);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(it.tgt.p, b, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(it.g0.p, b + ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(it.g1.p, b + 2 * (size_t)ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(it.slot.p, b + 3 * (size_t)ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
  }
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

// ------------------------------------------------ symmetric near field ----

// Does leaf t own the pair {t, b}? The one with the longer gather list; on a
// tie the lower id.
__device__ __forceinline__ bool sym_owns(const uint32_t* __restrict__ goff, uint32_t t, uint32_t b) {
  const uint32_t gt = goff[t + 1] - goff[t], gb = goff[b + 1] - goff[b];
  return gt > gb || (gt == gb && t < b);
}

// Row of CSR entry g: the last r with off[r] <= g (off ascending, off[0] = 0)
__device__ __forceinline__ uint32_t csr_row(const uint32_t* __restrict__ off, uint32_t nrows, uint32_t g) {
  uint32_t lo = 0, hi = nrows;  // off[lo] <= g < off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// One thread per gather entry (lists of hub leaves run to thousands of
// entries): does the entry's leaf own the pair? Entry ng (the last thread)
// writes 0, so the exclusive scan ends with the number of owned pairs.
__global__ void k_sym_flag(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gdat, uint32_t nl,
                           uint32_t ng, uint32_t* __restrict__ flag) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g > ng) return;
  uint32_t f = 0;
  if (g < ng) {
    const uint32_t t = csr_row(goff, nl, g), b = gdat[g];
    f = (b != t && sym_owns(goff, t, b)) ? 1u : 0u;
  }
  flag[g] = f;
}

// Per leaf: first owned pair (the scan at its first gather entry) and the
// number of items (at least one: the self block).
__global__ void k_sym_offsets(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ pos, uint32_t nl,
                              uint32_t* __restrict__ own_off, uint32_t* __restrict__ nit) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nl) return;
  own_off[t] = pos[goff[t]];
  if (t < nl) {
    const uint32_t c = pos[goff[t + 1]] - pos[goff[t]];
    nit[t] = max(1u, (c + SYM_ITEM_PARTNERS - 1) / SYM_ITEM_PARTNERS);
  } else {
    nit[t] = 0;
  }
}

// Owned partners in gather-list order; each pair's key (partner, owner)
// sorts the non-owned pairs of every leaf by owner.
__global__ void k_sym_own_fill(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gdat, uint32_t nl,
                               uint32_t ng, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                               uint32_t* __restrict__ own_dat, int w, uint64_t* __restrict__ key,
                               uint32_t* __restrict__ val) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng || !flag[g]) return;
  const uint32_t t = csr_row(goff, nl, g), b = gdat[g], u = pos[g];
  own_dat[u] = b;
  key[u] = ((uint64_t)b << w) | t;
  val[u] = u;
}

__global__ void k_sym_low_split(const uint64_t* __restrict__ key, uint32_t np, int w, uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  atomicAdd(&cnt[(uint32_t)(key[i] >> w)], 1u);
}

__global__ void k_sym_items(const uint32_t* __restrict__ own_off, const uint32_t* __restrict__ item_off, uint32_t nl,
                            uint32_t* __restrict__ itgt, uint32_t* __restrict__ iu0, uint32_t* __restrict__ iu1,
                            uint32_t* __restrict__ iown, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  const uint32_t e = own_off[t + 1];
  for (uint32_t i = item_off[t], u = own_off[t]; i < item_off[t + 1]; ++i, u += SYM_ITEM_PARTNERS) {
    itgt[i] = t;
    iu0[i] = u;
    iu1[i] = min(u + SYM_ITEM_PARTNERS, e);
    iown[i] = i;
    // costliest first (partners, plus the self block of the first item); the
    // stable sort keeps the spatial order within a class
    const uint32_t cost = (iu1[i] - u) + (u == own_off[t] ? 1u : 0u);
    key[i] = 64u - (cost ? 32u - (uint32_t)__clz(cost) : 0u);
    val[i] = i;
  }
}

// Per leaf: the number of refs (its items, then its non-owned pairs) and of
// chunks; heavy leaves (several chunks) need chunk sums.
__global__ void k_sym_ref_count(const uint32_t* __restrict__ item_off, const uint32_t* __restrict__ low_off,
                                uint32_t nl, uint32_t* __restrict__ nref, uint32_t* __restrict__ nch,
                                uint32_t* __restrict__ npart, uint32_t* __restrict__ heavy) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nl) return;
  if (t == nl) {
    nref[t] = nch[t] = npart[t] = heavy[t] = 0;
    return;
  }
  const uint32_t r = (item_off[t + 1] - item_off[t]) + (low_off[t + 1] - low_off[t]);
  const uint32_t c = (r + SYM_CHUNK - 1) / SYM_CHUNK;
  nref[t] = r;
  nch[t] = c;
  npart[t] = c > 1 ? c : 0u;
  heavy[t] = c > 1 ? 1u : 0u;
}

__global__ void k_sym_ref_fill(const uint32_t* __restrict__ item_off, const uint32_t* __restrict__ low_off,
                               const uint32_t* __restrict__ low_slot, uint32_t nl, const uint32_t* __restrict__ ref_off,
                               const uint32_t* __restrict__ ch_off, const uint32_t* __restrict__ part_off,
                               const uint32_t* __restrict__ heavy_off, uint32_t* __restrict__ refs,
                               uint32_t* __restrict__ ch_leaf, uint32_t* __restrict__ ch_r0,
                               uint32_t* __restrict__ ch_r1, uint32_t* __restrict__ ch_dst,
                               uint32_t* __restrict__ heavy_leaf, uint32_t* __restrict__ heavy_p0,
                               uint32_t* __restrict__ heavy_n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  uint32_t r = ref_off[t];
  for (uint32_t i = item_off[t]; i < item_off[t + 1]; ++i) refs[r++] = SYM_OWN_TAG | i;
  for (uint32_t l = low_off[t]; l < low_off[t + 1]; ++l) refs[r++] = low_slot[l];
  const uint32_t r0 = ref_off[t], r1 = ref_off[t + 1];
  const uint32_t nc = ch_off[t + 1] - ch_off[t];
  for (uint32_t c = 0; c < nc; ++c) {
    const uint32_t ci = ch_off[t] + c;
    ch_leaf[ci] = t;
    ch_r0[ci] = r0 + c * SYM_CHUNK;
    ch_r1[ci] = min(r0 + (c + 1) * SYM_CHUNK, r1);
    ch_dst[ci] = nc > 1 ? part_off[t] + c : NONE;
  }
  if (nc > 1) {
    const uint32_t hi = heavy_off[t];
    heavy_leaf[hi] = t;
    heavy_p0[hi] = part_off[t];
    heavy_n[hi] = nc;
  }
}

void sym_p2p_free(SymP2P& sp) {
  DBuf* b[] = {&sp.own_off, &sp.own_dat, &sp.itgt,       &sp.iu0,      &sp.iu1,     &sp.iown,  &sp.refs,
               &sp.ch_leaf, &sp.ch_r0,   &sp.ch_r1,      &sp.ch_dst,   &sp.heavy_leaf, &sp.heavy_p0, &sp.heavy_n,
               &sp.rel,     &sp.own2,    &sp.ownp,       &sp.pair2,    &sp.pairp,   &sp.cpart2, &sp.cpartp,
               &sp.vbs,     &sp.vbe,     &sp.vcx,        &sp.vcy,      &sp.vleaf,   &sp.ab2};
  for (DBuf* d : b) dbuf_free(*d);
  sp.valid = sp.halves = false;
  sp.n_items = sp.n_pairs = sp.n_leaves = sp.n_chunks = sp.n_heavy = sp.n_cpart = 0;
}

int build_sym_p2p(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t nl, SymP2P& sp) {
  hipStream_t s = h->stream;
  Workspace& ws = h->ws;
  sp.valid = false;
  const size_t L1 = (size_t)nl + 1;
  // per-leaf counters in workspace a: nown/own_off, nit/item_off, low_off,
  // nref/ref_off, nch/ch_off, npart/part_off, heavy/heavy_off
  GX_CHECK(dbuf_reserve(ws.a, L1 * 7 * 4, s));
  uint32_t* own_off = ws.a.as<uint32_t>();
  uint32_t* item_off = own_off + L1;
  uint32_t* low_off = item_off + L1;
  uint32_t* ref_off = low_off + L1;
  uint32_t* ch_off = ref_off + L1;
  uint32_t* part_off = ch_off + L1;
  uint32_t* heavy_off = part_off + L1;
  // gather entries: owned flags, their exclusive scan, per-leaf offsets
  uint32_t ng = 0;
  GX_HIP(hipMemcpyAsync(&ng, goff + nl, 4, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  ScopedDBuf gflag;
  GX_CHECK(dbuf_reserve(gflag, ((size_t)ng + 1) * 8, s));
  uint32_t* flag = gflag.as<uint32_t>();
  uint32_t* pos = flag + ng + 1;
  k_sym_flag<<<nblk((size_t)ng + 1), 256, 0, s>>>(goff, gdat, nl, ng, flag);
  GX_HIP(hipGetLastError());
  GX_CHECK(scan_u32(h, flag, pos, (size_t)ng + 1, false, false, nullptr));
  k_sym_offsets<<<nblk(L1), 256, 0, s>>>(goff, pos, nl, own_off, item_off);
  GX_HIP(hipGetLastError());
  uint64_t np64 = 0, ni64 = 0;
  GX_CHECK(sum_u32(h, flag, (size_t)ng + 1, &np64));
  GX_CHECK(sum_u32(h, item_off, L1, &ni64));
  if ((np64 + 1) * SYM_LEAF >= 0xFFFFFFF0ull || (ni64 + 1) * SYM_LEAF >= 0xFFFFFFF0ull)
    return set_error(GALAX_ERR_UNSUPPORTED, "symmetric near field: too many leaf pairs");
  const uint32_t np = (uint32_t)np64;
  uint32_t ni = 0;
  GX_CHECK(scan_u32(h, item_off, item_off, L1, false, false, &ni));
  GX_CHECK(dbuf_reserve(sp.own_off, L1 * 4, s));
  GX_HIP(hipMemcpyAsync(sp.own_off.p, own_off, L1 * 4, hipMemcpyDeviceToDevice, s));
  const size_t nsort = std::max<size_t>(std::max<size_t>(ni, np), 1);
  GX_CHECK(dbuf_reserve(ws.sort_ka, nsort * 8, s));
  GX_CHECK(dbuf_reserve(ws.sort_kb, nsort * 8, s));
  GX_CHECK(dbuf_reserve(ws.sort_va, nsort * 4, s));
  GX_CHECK(dbuf_reserve(ws.sort_vb, nsort * 4, s));
  // owned partners; the non-owned pairs of each leaf in ascending owner
  GX_CHECK(dbuf_reserve(sp.own_dat, (size_t)np * 4 + 4, s));
  GX_HIP(hipMemsetAsync(low_off, 0, L1 * 4, s));
  const int w = bits_for(nl);
  k_sym_own_fill<<<nblk(ng), 256, 0, s>>>(goff, gdat, nl, ng, flag, pos, sp.own_dat.as<uint32_t>(), w,
                                          ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>());
  GX_HIP(hipGetLastError());
  DBuf low_slot;
  GX_CHECK(dbuf_reserve(low_slot, (size_t)np * 4 + 4, s));
  if (np) {
    GX_CHECK(radix_sort_bits(h, ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>(), ws.sort_kb.as<uint64_t>(),
                             ws.sort_vb.as<uint32_t>(), np, 2 * w));
    k_sym_low_split<<<nblk(np), 256, 0, s>>>(ws.sort_ka.as<uint64_t>(), np, w, low_off);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(low_slot.p, ws.sort_va.p, (size_t)np * 4, hipMemcpyDeviceToDevice, s));
  }
  GX_CHECK(scan_u32(h, low_off, low_off, L1, false, false, nullptr));
  // items, ordered by decreasing work class
  GX_CHECK(dbuf_reserve(sp.itgt, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.iu0, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.iu1, (size_t)ni * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.iown, (size_t)ni * 4 + 4, s));
  k_sym_items<<<nblk(nl), 256, 0, s>>>(own_off, item_off, nl, sp.itgt.as<uint32_t>(), sp.iu0.as<uint32_t>(),
                                       sp.iu1.as<uint32_t>(), sp.iown.as<uint32_t>(), ws.sort_ka.as<uint64_t>(),
                                       ws.sort_va.as<uint32_t>());
  GX_HIP(hipGetLastError());
  if (ni > 1) {
    GX_CHECK(radix_sort_bits(h, ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>(), ws.sort_kb.as<uint64_t>(),
                             ws.sort_vb.as<uint32_t>(), ni, 7));
    DBuf tmp;
    GX_CHECK(dbuf_reserve(tmp, (size_t)ni * 16, s));
    uint32_t* b = tmp.as<uint32_t>();
    k_items_permute<<<nblk(ni), 256, 0, s>>>(ws.sort_va.as<uint32_t>(), ni, sp.itgt.as<uint32_t>(),
                                             sp.iu0.as<uint32_t>(), sp.iu1.as<uint32_t>(), sp.iown.as<uint32_t>(), b);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(sp.itgt.p, b, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(sp.iu0.p, b + ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(sp.iu1.p, b + 2 * (size_t)ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipMemcpyAsync(sp.iown.p, b + 3 * (size_t)ni, (size_t)ni * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipStreamSynchronize(s));
    dbuf_free(tmp);
  }
  // refs and chunks
  k_sym_ref_count<<<nblk(L1), 256, 0, s>>>(item_off, low_off, nl, ref_off, ch_off, part_off, heavy_off);
  GX_HIP(hipGetLastError());
  uint32_t nref = 0, nch = 0, ncp = 0, nheavy = 0;
  GX_CHECK(scan_u32(h, ref_off, ref_off, L1, false, false, &nref));
  GX_CHECK(scan_u32(h, ch_off, ch_off, L1, false, false, &nch));
  GX_CHECK(scan_u32(h, part_off, part_off, L1, false, false, &ncp));
  GX_CHECK(scan_u32(h, heavy_off, heavy_off, L1, false, false, &nheavy));
  GX_CHECK(dbuf_reserve(sp.refs, (size_t)nref * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.ch_leaf, (size_t)nch * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.ch_r0, (size_t)nch * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.ch_r1, (size_t)nch * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.ch_dst, (size_t)nch * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.heavy_leaf, (size_t)nheavy * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.heavy_p0, (size_t)nheavy * 4 + 4, s));
  GX_CHECK(dbuf_reserve(sp.heavy_n, (size_t)nheavy * 4 + 4, s));
  k_sym_ref_fill<<<nblk(nl), 256, 0, s>>>(item_off, low_off, low_slot.as<uint32_t>(), nl, ref_off, ch_off, part_off,
                                          heavy_off, sp.refs.as<uint32_t>(), sp.ch_leaf.as<uint32_t>(),
                                          sp.ch_r0.as<uint32_t>(), sp.ch_r1.as<uint32_t>(), sp.ch_dst.as<uint32_t>(),
                                          sp.heavy_leaf.as<uint32_t>(), sp.heavy_p0.as<uint32_t>(),
                                          sp.heavy_n.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(dbuf_reserve(sp.rel, std::max<size_t>(h->tree.n_bodies, 1) * 16, s));
  GX_CHECK(dbuf_reserve(sp.own2, std::max<size_t>(ni, 1) * SYM_LEAF * 16, s));
  GX_CHECK(dbuf_reserve(sp.pair2, std::max<size_t>(np, 1) * SYM_LEAF * 8, s));
  GX_CHECK(dbuf_reserve(sp.cpart2, std::max<size_t>(ncp, 1) * SYM_LEAF * 16, s));
  GX_HIP(hipStreamSynchronize(s));
  dbuf_free(low_slot);
  sp.n_items = ni;
  sp.n_pairs = np;
  sp.n_leaves = nl;
  sp.n_chunks = nch;
  sp.n_heavy = nheavy;
  sp.n_cpart = ncp;
  sp.valid = true;
  return 0;
}

// ---- halves of leaves with more than SYM_LEAF bodies ----

// halves per leaf (1 or 2); entry nl is 0, so the exclusive scan ends with the count
__global__ void k_vh_count(const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be, uint32_t nl,
                           uint32_t* __restrict__ nh) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nl) return;
  nh[t] = t < nl ? (be[t] - bs[t] > SYM_LEAF ? 2u : 1u) : 0u;
}

// each leaf's halves: the first SYM_LEAF bodies, then the rest
__global__ void k_vh_fill(const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                          const double* __restrict__ comx, const double* __restrict__ comy, uint32_t nl,
                          const uint32_t* __restrict__ voff, uint32_t* __restrict__ vbs, uint32_t* __restrict__ vbe,
                          double* __restrict__ vcx, double* __restrict__ vcy, uint32_t* __restrict__ vleaf) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  const uint32_t v = voff[t], b0 = bs[t], b1 = be[t], mid = min(b1, b0 + SYM_LEAF);
  vbs[v] = b0;
  vbe[v] = mid;
  vcx[v] = comx[t];
  vcy[v] = comy[t];
  vleaf[v] = t;
  if (b1 > mid) {
    vbs[v + 1] = mid;
    vbe[v + 1] = b1;
    vcx[v + 1] = comx[t];
    vcy[v + 1] = comy[t];
    vleaf[v + 1] = t;
  }
}

// gather-list length of each half: its leaf's partners' halves; entry nv is 0
__global__ void k_vg_count(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gdat, uint32_t nl,
                           const uint32_t* __restrict__ voff, uint32_t* __restrict__ vcnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nl) return;
  if (t == nl) {
    vcnt[voff[nl]] = 0u;
    return;
  }
  uint32_t c = 0;
  for (uint32_t g = goff[t]; g < goff[t + 1]; ++g) c += voff[gdat[g] + 1] - voff[gdat[g]];
  for (uint32_t v = voff[t]; v < voff[t + 1]; ++v) vcnt[v] = c;
}

// gather list of each half: its leaf's list, every partner replaced by its halves in order
__global__ void k_vg_fill(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gdat, uint32_t nl,
                          const uint32_t* __restrict__ voff, const uint32_t* __restrict__ vgoff,
                          uint32_t* __restrict__ vgdat) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  for (uint32_t v = voff[t]; v < voff[t + 1]; ++v) {
    uint32_t q = vgoff[v];
    for (uint32_t g = goff[t]; g < goff[t + 1]; ++g)
      for (uint32_t w = voff[gdat[g]]; w < voff[gdat[g] + 1]; ++w) vgdat[q++] = w;
  }
}

int build_sym_near(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t nl, SymP2P& sp) {
  if (h->cfg.n_max > 2 * SYM_LEAF) return set_error(GALAX_ERR_INVALID, "symmetric near field: leaves too large");
  sp.halves = h->cfg.n_max > SYM_LEAF;
  if (!sp.halves) return build_sym_p2p(h, goff, gdat, nl, sp);
  hipStream_t s = h->stream;
  const DevTree& t = h->tree;
  const size_t L1 = (size_t)nl + 1;
  ScopedDBuf voffb, vgoffb, vgdatb;
  GX_CHECK(dbuf_reserve(voffb, L1 * 4, s));
  uint32_t* voff = voffb.as<uint32_t>();
  k_vh_count<<<nblk(L1), 256, 0, s>>>(t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), nl, voff);
  GX_HIP(hipGetLastError());
  uint32_t nv = 0;
  GX_CHECK(scan_u32(h, voff, voff, L1, false, false, &nv));
  const size_t V1 = (size_t)nv + 1;
  GX_CHECK(dbuf_reserve(sp.vbs, V1 * 4, s));
  GX_CHECK(dbuf_reserve(sp.vbe, V1 * 4, s));
  GX_CHECK(dbuf_reserve(sp.vcx, V1 * 8, s));
  GX_CHECK(dbuf_reserve(sp.vcy, V1 * 8, s));
  GX_CHECK(dbuf_reserve(sp.vleaf, V1 * 4, s));
  k_vh_fill<<<nblk(nl), 256, 0, s>>>(t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), t.com_x.as<double>(),
                                     t.com_y.as<double>(), nl, voff, sp.vbs.as<uint32_t>(), sp.vbe.as<uint32_t>(),
                                     sp.vcx.as<double>(), sp.vcy.as<double>(), sp.vleaf.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(dbuf_reserve(vgoffb, V1 * 4, s));
  uint32_t* vgoff = vgoffb.as<uint32_t>();
  k_vg_count<<<nblk(L1), 256, 0, s>>>(goff, gdat, nl, voff, vgoff);
  GX_HIP(hipGetLastError());
  uint64_t nvg64 = 0;
  GX_CHECK(sum_u32(h, vgoff, V1, &nvg64));
  if (nvg64 >= 0xFFFFFFF0ull) return set_error(GALAX_ERR_UNSUPPORTED, "symmetric near field: too many half pairs");
  uint32_t nvg = 0;
  GX_CHECK(scan_u32(h, vgoff, vgoff, V1, false, false, &nvg));
  GX_CHECK(dbuf_reserve(vgdatb, (size_t)nvg * 4 + 4, s));
  k_vg_fill<<<nblk(nl), 256, 0, s>>>(goff, gdat, nl, voff, vgoff, vgdatb.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(build_sym_p2p(h, vgoff, vgdatb.as<uint32_t>(), nv, sp));
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

static int build_gather(galax_fmm* h) {
  DevTree& t = h->tree;
  t.sym.valid = false;  // built from the gather lists at its first use
  hipStream_t s = h->stream;
  uint32_t nl = (uint32_t)t.n_leaves;
  GX_CHECK(dbuf_reserve(h->ws.e, (nl + 1) * sizeof(uint32_t), s));
  uint32_t* cnt = h->ws.e.as<uint32_t>();
  GX_HIP(hipMemsetAsync(cnt, 0, (nl + 1) * sizeof(uint32_t), s));
  k_pair_count<<<nblk(nl), 256, 0, s>>>(nl, t.p2p_off.as<uint32_t>(), t.p2p_dat.as<uint32_t>(), cnt);
  GX_HIP(hipGetLastError());
  uint32_t total = 0;
  GX_CHECK(scan_u32(h, cnt, cnt, nl + 1, false, false, &total));
  t.gat_total = total;
  const int w = bits_for(nl);
  GX_CHECK(dbuf_reserve(h->ws.sort_ka, (size_t)total * 8 + 8, s));
  GX_CHECK(dbuf_reserve(h->ws.sort_kb, (size_t)total * 8 + 8, s));
  GX_CHECK(dbuf_reserve(h->ws.sort_va, (size_t)total * 4 + 4, s));
  GX_CHECK(dbuf_reserve(h->ws.sort_vb, (size_t)total * 4 + 4, s));
  uint64_t* keys = h->ws.sort_ka.as<uint64_t>();
  k_pair_emit<<<nblk(nl), 256, 0, s>>>(nl, t.p2p_off.as<uint32_t>(), t.p2p_dat.as<uint32_t>(), cnt, w, keys,
                                       h->ws.sort_va.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(radix_sort_bits(h, keys, h->ws.sort_va.as<uint32_t>(), h->ws.sort_kb.as<uint64_t>(),
                           h->ws.sort_vb.as<uint32_t>(), total, 2 * w));
  GX_CHECK(dbuf_reserve(t.gat_off, (nl + 1) * sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(t.gat_dat, (size_t)total * sizeof(uint32_t) + 4, s));
  uint32_t* goff = t.gat_off.as<uint32_t>();
  GX_HIP(hipMemsetAsync(goff, 0, (nl + 1) * sizeof(uint32_t), s));
  if (total) {
    k_pair_split<<<nblk(total), 256, 0, s>>>(keys, total, w, nl, t.gat_dat.as<uint32_t>(), goff);
    GX_HIP(hipGetLastError());
  }
  GX_CHECK(count_interactions(h, goff, t.gat_dat.as<uint32_t>(), &t.p2p_interactions));
  GX_CHECK(build_p2p_items(h, goff, t.gat_dat.as<uint32_t>(), nl, t.items));
  return 0;
}

// Segment trees over every level's nodes (the candidates of both list
// builds), in h->ws.g. lv (host, the levels' layout) is copied from
// asynchronously: the caller keeps it until its next wait on the stream.
static int build_segtrees(galax_fmm* h, const NodeArrays& a, std::vector<SegLevel>& lv, SegTree& st,
                          SegLevel*& dlv) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  lv.assign(t.n_levels, SegLevel{});
  size_t tot = 0, slots = 0;
  uint32_t maxh = 0;
  for (uint32_t l = 0; l < t.n_levels; ++l) {
    const uint32_t lo = t.level_offset[l], n = t.level_offset[l + 1] - lo;
    uint32_t p2 = 1, hgt = 0;
    while (p2 < n) {
      p2 <<= 1;
      ++hgt;
    }
    lv[l] = SegLevel{(uint32_t)tot, p2, lo, n};
    tot += 2 * (size_t)p2;
    slots += p2;
    maxh = std::max(maxh, hgt);
  }
  GX_CHECK(dbuf_reserve(h->ws.g, tot * sizeof(SegRec) + t.n_levels * sizeof(SegLevel), s));
  st = SegTree{h->ws.g.as<SegRec>()};
  dlv = reinterpret_cast<SegLevel*>(st.r + tot);
  GX_HIP(hipMemcpyAsync(dlv, lv.data(), t.n_levels * sizeof(SegLevel), hipMemcpyHostToDevice, s));
  k_seg_leaves<<<nblk(slots), 256, 0, s>>>(a, dlv, t.n_levels, st);
  GX_HIP(hipGetLastError());
  for (uint32_t hh = 1; hh <= maxh; ++hh) {
    size_t cnt = 0;
    for (const SegLevel& L : lv) cnt += L.p2 >> hh;
    k_seg_up<<<nblk(cnt), 256, 0, s>>>(dlv, t.n_levels, hh, st);
    GX_HIP(hipGetLastError());
  }
  return 0;
}

// The M2L CSR (m2l_off, m2l_dat, m2l_total) from per-node rows (of the first
// nrows nodes) and lengths (lens: nn + 1 entries, 0 for rows from nrows on;
// the last is cleared here).
static int compact_m2l(galax_fmm* h, const uint32_t* rows, uint32_t* lens, size_t nrows) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const size_t nn = t.n_nodes;
  GX_HIP(hipMemsetAsync(lens + nn, 0, sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(t.m2l_off, (nn + 1) * sizeof(uint32_t), s));
  uint32_t mt = 0;
  GX_CHECK(scan_u32(h, lens, t.m2l_off.as<uint32_t>(), nn + 1, false, false, &mt));
  t.m2l_total = mt;
  GX_CHECK(dbuf_reserve(t.m2l_dat, (size_t)mt * sizeof(uint32_t) + 4, s));
  if (nrows) {
    k_rows_compact<<<nblk(nrows * LIST_CAP), 256, 0, s>>>(rows, t.m2l_off.as<uint32_t>(), (uint32_t)nrows,
                                                          t.m2l_dat.as<uint32_t>());
    GX_HIP(hipGetLastError());
  }
  return 0;
}

// The reference's lists. The M2L rows are walked for the nodes of levels 0
// and 1 only, the rows any evaluation reads (a leaf receives its own M2L and
// its parent's: the ascending-id L2L, galax-core/src/lib.rs:1266); the rows of
// levels >= 2 stay empty until ensure_full_m2l_lists builds them (for an
// export of the lists or their total). Node ids run level by level, so the
// live rows, their offsets and their entries are the same either way.
int build_lists_reference(galax_fmm* h) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  size_t nn = t.n_nodes;
  t.m2l_full = true;
  if (nn == 0) {
    t.m2l_total = t.p2p_total = t.gat_total = 0;
    return 0;
  }
  NodeArrays a = node_arrays(h);
  const uint32_t nwalk = t.level_offset[std::min<uint32_t>(t.n_levels, 2)];
  GX_CHECK(dbuf_reserve(h->ws.d, (size_t)nwalk * LIST_CAP * sizeof(uint32_t), s));  // rows of the walked nodes
  GX_CHECK(dbuf_reserve(h->ws.c, (nn + 1) * sizeof(uint32_t), s));
  uint32_t* rows = h->ws.d.as<uint32_t>();
  uint32_t* lens = h->ws.c.as<uint32_t>();
  SegTree st{};
  SegLevel* dlv = nullptr;
  std::vector<SegLevel> lv;  // kept until the scans below wait on the stream
  GX_CHECK(build_segtrees(h, a, lv, st, dlv));
  // M2L and P2P rows in one launch (P2P: leaves only; other rows empty)
  GX_CHECK(dbuf_reserve(h->ws.e, std::max<size_t>(t.n_leaves, 1) * LIST_CAP * sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(h->ws.f, (nn + 1) * sizeof(uint32_t), s));
  uint32_t* prows = h->ws.e.as<uint32_t>();
  uint32_t* plens = h->ws.f.as<uint32_t>();
  GX_HIP(hipMemsetAsync(plens, 0, (nn + 1) * sizeof(uint32_t), s));
  {
    const uint32_t nbp = nblk(t.n_leaves);
    k_rows_seg<<<nbp + nblk(nn), 256, 0, s>>>(a, (uint32_t)nn, nwalk, (uint32_t)t.n_leaves, nbp, dlv, t.n_levels, st,
                                              rows, lens, prows, plens);
    GX_HIP(hipGetLastError());
  }
  GX_CHECK(compact_m2l(h, rows, lens, nwalk));
  t.m2l_full = nwalk >= nn;
  // P2P lists
  GX_CHECK(dbuf_reserve(t.p2p_off, (nn + 1) * sizeof(uint32_t), s));
  uint32_t pt = 0;
  GX_CHECK(scan_u32(h, plens, t.p2p_off.as<uint32_t>(), nn + 1, false, false, &pt));
  t.p2p_total = pt;
  GX_CHECK(dbuf_reserve(t.p2p_dat, (size_t)pt * sizeof(uint32_t) + 4, s));
  k_rows_compact<<<nblk(t.n_leaves * LIST_CAP), 256, 0, s>>>(prows, t.p2p_off.as<uint32_t>(), (uint32_t)t.n_leaves,
                                                             t.p2p_dat.as<uint32_t>());
  GX_HIP(hipGetLastError());
  GX_CHECK(build_mark(h, 4));
  GX_CHECK(build_gather(h));
  h->keys_valid = false;
  return 0;
}

static int build_full_m2l_lists(galax_fmm* h);

int ensure_full_m2l_lists(galax_fmm* h) {
  if (h->tree.m2l_full || h->tree.n_nodes == 0) return 0;
  const int rc = build_full_m2l_lists(h);
  // a failure can leave the live rows overwritten: the tree is then unusable
  if (rc) h->has_tree = false;
  return rc;
}

static int build_full_m2l_lists(galax_fmm* h) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const size_t nn = t.n_nodes;
  NodeArrays a = node_arrays(h);
  GX_CHECK(dbuf_reserve(h->ws.d, nn * LIST_CAP * sizeof(uint32_t), s));
  GX_CHECK(dbuf_reserve(h->ws.c, (nn + 1) * sizeof(uint32_t), s));
  uint32_t* rows = h->ws.d.as<uint32_t>();
  uint32_t* lens = h->ws.c.as<uint32_t>();
  SegTree st{};
  SegLevel* dlv = nullptr;
  std::vector<SegLevel> lv;  // kept until the scans below wait on the stream
  GX_CHECK(build_segtrees(h, a, lv, st, dlv));
  k_rows_seg<<<nblk(nn), 256, 0, s>>>(a, (uint32_t)nn, (uint32_t)nn, (uint32_t)t.n_leaves, 0u, dlv, t.n_levels, st,
                                      rows, lens, nullptr, nullptr);
  GX_HIP(hipGetLastError());
  GX_CHECK(compact_m2l(h, rows, lens, nn));
  GX_HIP(hipStreamSynchronize(s));
  t.m2l_full = true;
  return 0;
}

// --------------------------------------------------- host tree install ----

int install_tree_host(galax_fmm* h, const galax_tree_view_t* tv, const galax_lists_view_t* m2l,
                      const galax_lists_view_t* p2p) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  size_t nn = tv->n_nodes;
  if (m2l->n_nodes != nn || p2p->n_nodes != nn) return set_error(GALAX_ERR_INVALID, "list sizes != tree size");
  if (nn >= 0xFFFFFFF0ull) return set_error(GALAX_ERR_INVALID, "too many nodes");
  if (nn && (!tv->center_x || !tv->center_y || !tv->half_width || !tv->mass || !tv->com_x || !tv->com_y ||
             !tv->body_start || !tv->body_end || !tv->children || !tv->parent || !tv->is_leaf))
    return set_error(GALAX_ERR_INVALID, "NULL tree array");
  if (!m2l->offsets || !p2p->offsets) return set_error(GALAX_ERR_INVALID, "NULL list offsets");
  std::vector<uint32_t> level(nn, 0), cf(nn, NONE), par(nn, NONE), bs(nn), be(nn);
  std::vector<uint8_t> cc(nn, 0), lv8(nn);
  size_t nl = 0, nb = 0;
  for (size_t i = 0; i < nn; ++i) {
    uint32_t mx = 0;
    int k = 0;
    for (; k < 4; ++k) {
      int32_t c = tv->children[i * 4 + k];
      if (c < 0) break;
      if ((size_t)c >= i) return set_error(GALAX_ERR_INVALID, "child id %d >= parent id %zu", c, i);
      if (k > 0 && c != tv->children[i * 4 + k - 1] + 1)
        return set_error(GALAX_ERR_INVALID, "children of node %zu are not consecutive", i);
      mx = std::max(mx, level[c] + 1);
    }
    for (int k2 = k; k2 < 4; ++k2)
      if (tv->children[i * 4 + k2] >= 0) return set_error(GALAX_ERR_INVALID, "gap in children of node %zu", i);
    if (k > 0) {
      cf[i] = (uint32_t)tv->children[i * 4];
      cc[i] = (uint8_t)k;
    }
    if ((k == 0) != (tv->is_leaf[i] != 0)) return set_error(GALAX_ERR_INVALID, "is_leaf inconsistent at %zu", i);
    level[i] = mx;
    if (mx > 255) return set_error(GALAX_ERR_INVALID, "tree deeper than 255 levels");
    lv8[i] = (uint8_t)mx;
    const int32_t pi = tv->parent[i];
    if (pi >= 0 && ((size_t)pi <= i || (size_t)pi >= nn))
      return set_error(GALAX_ERR_INVALID, "parent %d of node %zu is out of range", pi, i);
    par[i] = pi < 0 ? NONE : (uint32_t)pi;
    if (tv->body_start[i] > tv->body_end[i] || tv->body_end[i] > h->max_n)
      return set_error(GALAX_ERR_INVALID, "node %zu: body range [%llu, %llu) not within [0, max_n = %zu]", i,
                       (unsigned long long)tv->body_start[i], (unsigned long long)tv->body_end[i], h->max_n);
    bs[i] = (uint32_t)tv->body_start[i];
    be[i] = (uint32_t)tv->body_end[i];
    if (k > 0 && (bs[i] != bs[cf[i]] || be[i] != be[cf[i] + k - 1]))
      return set_error(GALAX_ERR_INVALID, "node %zu: body range differs from its children's", i);
    if (k == 0) {
      ++nl;
      nb = std::max<size_t>(nb, be[i]);
      if (be[i] - bs[i] > h->cfg.n_max)
        return set_error(GALAX_ERR_INVALID, "leaf %zu holds %u bodies, more than n_max = %u", i, be[i] - bs[i],
                         h->cfg.n_max);
    }
  }
  for (size_t i = 1; i < nn; ++i)
    if (level[i] < level[i - 1]) return set_error(GALAX_ERR_INVALID, "node ids are not level-ordered");
  for (size_t i = 0; i < nl; ++i)
    if (level[i] != 0) return set_error(GALAX_ERR_INVALID, "leaves must come first");
  for (size_t i = 0; i < nn; ++i)
    for (uint32_t c = cf[i]; cc[i] && c < cf[i] + cc[i]; ++c)
      if (par[c] != i) return set_error(GALAX_ERR_INVALID, "child %u of node %zu has parent %d", c, i, tv->parent[c]);
  // lists: offsets from 0, non-decreasing, u32-sized; entries name nodes (M2L)
  // or leaves (P2P), which the device kernels index without further checks
  auto check_lists = [&](const galax_lists_view_t* lv, size_t bound, const char* what) -> int {
    if (lv->offsets[0] != 0) return set_error(GALAX_ERR_INVALID, "%s offsets[0] != 0", what);
    for (size_t i = 0; i < nn; ++i)
      if (lv->offsets[i + 1] < lv->offsets[i])
        return set_error(GALAX_ERR_INVALID, "%s offsets decrease at node %zu", what, i);
    const uint64_t total = lv->offsets[nn];
    if (total >= 0xFFFFFFF0ull) return set_error(GALAX_ERR_INVALID, "%s lists too long", what);
    if (total && !lv->data) return set_error(GALAX_ERR_INVALID, "NULL %s list data", what);
    for (uint64_t k = 0; k < total; ++k)
      if (lv->data[k] >= bound)
        return set_error(GALAX_ERR_INVALID, "%s list entry %llu = %u is not below %zu", what, (unsigned long long)k,
                         lv->data[k], bound);
    return 0;
  };
  GX_CHECK(check_lists(m2l, nn, "M2L"));
  GX_CHECK(check_lists(p2p, nl, "P2P"));
  h->has_tree = false;  // from here on the previous tree is overwritten
  GX_CHECK(ensure_node_capacity(h, nn));
  auto up = [&](DBuf& d, const void* src, size_t bytes) -> int {
    GX_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s));
    return 0;
  };
  GX_CHECK(up(t.center_x, tv->center_x, nn * 8));
  GX_CHECK(up(t.center_y, tv->center_y, nn * 8));
  GX_CHECK(up(t.half_width, tv->half_width, nn * 8));
  GX_CHECK(up(t.mass, tv->mass, nn * 8));
  GX_CHECK(up(t.com_x, tv->com_x, nn * 8));
  GX_CHECK(up(t.com_y, tv->com_y, nn * 8));
  GX_CHECK(up(t.body_start, bs.data(), nn * 4));
  GX_CHECK(up(t.body_end, be.data(), nn * 4));
  GX_CHECK(up(t.child_first, cf.data(), nn * 4));
  GX_CHECK(up(t.parent, par.data(), nn * 4));
  GX_CHECK(up(t.child_count, cc.data(), nn));
  GX_CHECK(up(t.level, lv8.data(), nn));
  t.n_nodes = nn;
  t.n_leaves = nl;
  t.n_bodies = nb;
  t.n_levels = nn ? level[nn - 1] + 1 : 0;
  t.chain_ok.clear();
  t.level_offset.assign(t.n_levels + 1, (uint32_t)nn);
  t.level_offset[0] = 0;
  for (uint32_t l = 1; l <= t.n_levels; ++l) {
    size_t i = t.level_offset[l - 1];
    while (i < nn && level[i] < l) ++i;
    t.level_offset[l] = (uint32_t)i;
  }
  auto up_lists = [&](const galax_lists_view_t* lv, DBuf& off, DBuf& dat, size_t& total) -> int {
    std::vector<uint32_t> o(nn + 1);
    for (size_t i = 0; i <= nn; ++i) o[i] = (uint32_t)lv->offsets[i];
    total = lv->offsets[nn];
    GX_CHECK(dbuf_reserve(off, (nn + 1) * 4, s));
    GX_CHECK(dbuf_reserve(dat, total * 4 + 4, s));
    GX_HIP(hipMemcpyAsync(off.p, o.data(), (nn + 1) * 4, hipMemcpyHostToDevice, s));
    if (total) GX_HIP(hipMemcpyAsync(dat.p, lv->data, total * 4, hipMemcpyHostToDevice, s));
    GX_HIP(hipStreamSynchronize(s));
    return 0;
  };
  t.m2l_full = true;  // the caller's lists, every row
  GX_CHECK(up_lists(m2l, t.m2l_off, t.m2l_dat, t.m2l_total));
  GX_CHECK(up_lists(p2p, t.p2p_off, t.p2p_dat, t.p2p_total));
  GX_CHECK(build_gather(h));
  h->keys_valid = false;
  return 0;
}

}  // namespace gx
