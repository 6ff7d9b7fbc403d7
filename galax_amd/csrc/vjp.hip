// vjp.hip — vector-Jacobian product of the FMM accelerations (no reference
// counterpart; for differentiable use and checking through torch.autograd).
//
// For cotangents (ax_bar, ay_bar) the adjoint passes run in reverse order of
// the forward ones, on the same static tree, lists and node centres (the
// reference freezes node geometry at build time, galax-cli/src/main.rs:159,
// galax-integrate/src/lib.rs:217), so the map (x, y, m) -> (ax, ay) is the
// composition of polynomial and linear stages:
//   near  a_i += sum_j m_j F(x_j - x_i),  F(d) = d / (|d|^2 + eps^2)^{3/2}
//         x_bar_j += sum_i K(x_i - x_j) (m_j a_bar_i - m_i a_bar_j),   K = dF/dd
//         m_bar_j -= sum_i a_bar_i . F(x_i - x_j)
//   L2P   a_b = -grad Phi_leaf(u_b):  x_bar_b -= H(u_b) a_bar_b (Hessian),
//         Lbar[IJ] -= sum_b a_bar_b . grad(u^I v^J)(u_b)
//   L2L^T, M2L^T (transposed operators), M2M^T, then P2M^T:
//         x_bar_b += m_b sum Mbar[ij] d/dx (u^i v^j / i!j!), m_bar_b += sum Mbar[ij] u^i v^j / i!j!
// Every stage is a gather (one writer per output), so results are deterministic.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.h"

namespace gx {

static inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <int P>
__device__ __forceinline__ void cij(int k, int& i, int& j) {
  int n = 0;
#pragma unroll
  for (int q = 0; q <= P; ++q)
    if (k >= (q + 1) * (q + 2) / 2) n = q + 1;
  j = k - n * (n + 1) / 2;
  i = n - j;
}

__device__ __forceinline__ double dfact(int n) {
  double f = 1.0;
  for (int q = 2; q <= n; ++q) f *= q;
  return f;
}

__device__ __forceinline__ double dbinom(int n, int k) {
  double r = 1.0;
  for (int q = 1; q <= k; ++q) r = r * (n - k + q) / q;
  return r;
}

// ------------------------------------------------------------- near ----

// Adjoint of the near field for the targets j of one P2P work item (a target
// leaf and at most P2P_ITEM_LEAVES of its gather list), as the forward P2P
// kernel is organised: the item's sources are staged through LDS in chunks of
// VJP_CHUNK as (x, y, m, abar_x) and abar_y, in T precision with x, y relative
// to the target leaf's COM; lanes 0-31 and 32-63 hold the same 32 targets and
// take the even / odd sources (the chunk is padded to an even length with a
// zero-mass, zero-cotangent source far away); only the chunks holding the
// target leaf itself run with the self mask. With d = x_i - x_j and
// v = m_j abar_i - m_i abar_j (i the source):
//   xbar_j += v / r^3 - 3 d (d . v) / r^5,   mbar_j -= abar_i . d / r^3.
// T = float for the mixed forms (the precision of their forward near field).
constexpr int VJP_CHUNK = 256;
template <typename T, bool MASK>
__device__ __forceinline__ void vjp_near_sum(const typename Near<T>::V4* __restrict__ sa, const T* __restrict__ sb,
                                             int half, int n, int self, T xj, T yj, T mj, T abx, T aby, T eps2,
                                             T& gx, T& gy, T& gm) {
  const typename Near<T>::V4* ah = sa + half;
  const T* bh = sb + half;
#pragma unroll 2
  for (int jj = 0; jj < n; jj += 2) {
    const typename Near<T>::V4 s = ah[jj];
    const T sby = bh[jj];
    const T dx = s.x - xj, dy = s.y - yj;
    const T r2 = fma(dx, dx, fma(dy, dy, eps2));
    T inv = Near<T>::rsq(r2);
    if (MASK) inv = jj + half == self ? (T)0 : inv;
    const T inv2 = inv * inv, inv3 = inv2 * inv, inv5 = inv3 * inv2;
    const T vx = fma(mj, s.w, -s.z * abx), vy = fma(mj, sby, -s.z * aby);
    const T t = (T)3 * fma(dx, vx, dy * vy) * inv5;
    gx += fma(vx, inv3, -dx * t);
    gy += fma(vy, inv3, -dy * t);
    gm -= fma(s.w, dx, sby * dy) * inv3;
  }
}

template <typename T>
__global__ void __launch_bounds__(64) k_vjp_near(const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ m, const double* __restrict__ axb,
                                                 const double* __restrict__ ayb, const uint32_t* __restrict__ bs,
                                                 const uint32_t* __restrict__ be, const double* __restrict__ comx,
                                                 const double* __restrict__ comy, const uint32_t* __restrict__ gdat,
                                                 const uint32_t* __restrict__ itgt, const uint32_t* __restrict__ ig0,
                                                 const uint32_t* __restrict__ ig1, const uint32_t* __restrict__ islot,
                                                 uint32_t n_max, double eps2d, double* __restrict__ xb,
                                                 double* __restrict__ yb, double* __restrict__ mb,
                                                 double* __restrict__ part) {
  typedef typename Near<T>::V4 V4;
  static_assert(P2P_ITEM_LEAVES <= 64, "one lane per source leaf of an item");
  __shared__ V4 sa[VJP_CHUNK + 1];
  __shared__ T sb[VJP_CHUNK + 1];
  const uint32_t item = xcd_runs(blockIdx.x, gridDim.x);
  const uint32_t leaf = itgt[item], g0 = ig0[item], g1 = ig1[item], slot = islot[item];
  const int lane = threadIdx.x, tl = lane & 31, half = lane >> 5;
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
  const unsigned long long sm_ = __ballot(is_self);
  uint32_t self_lo = 0, self_hi = 0;
  if (sm_) {
    self_lo = __shfl(inc - cnt, __ffsll((long long)sm_) - 1, 64);
    self_hi = self_lo + (te - ts);
  }
  for (uint32_t tb = ts; tb < te; tb += 32) {
    const uint32_t j = tb + tl;
    const bool valid = j < te;
    // lanes past the leaf's end sit at its COM with zero mass and cotangent
    const T xj = valid ? (T)(x[j] - ox) : (T)0, yj = valid ? (T)(y[j] - oy) : (T)0;
    const T mj = valid ? (T)m[j] : (T)0, abx = valid ? (T)axb[j] : (T)0, aby = valid ? (T)ayb[j] : (T)0;
    T gx_ = 0, gy_ = 0, gm_ = 0;
    uint32_t jcur = 0;
    for (uint32_t c0 = 0; c0 < total; c0 += VJP_CHUNK) {
      const int n = (int)min((uint32_t)VJP_CHUNK, total - c0);
      __syncthreads();
      for (int w0 = 0; w0 < n; w0 += 64) {  // uniform: 64-element windows
        const uint32_t g = item_source(c0 + (uint32_t)w0, lane, nlv, offs, ss, jcur);
        const int pp = w0 + lane;
        if (pp >= n) continue;
        V4 v;
        v.x = (T)(x[g] - ox);
        v.y = (T)(y[g] - oy);
        v.z = (T)m[g];
        v.w = (T)axb[g];
        sa[pp] = v;
        sb[pp] = (T)ayb[g];
      }
      if (lane == 0 && (n & 1)) {  // pad to even: zero mass and cotangent, far away
        V4 v;
        v.x = (T)3e19;
        v.y = (T)3e19;
        v.z = (T)0;
        v.w = (T)0;
        sa[n] = v;
        sb[n] = (T)0;
      }
      __syncthreads();
      const int ne = (n + 1) & ~1;
      if (c0 < self_hi && c0 + (uint32_t)n > self_lo)
        vjp_near_sum<T, true>(sa, sb, half, ne, (int)(self_lo + (j - ts)) - (int)c0, xj, yj, mj, abx, aby, eps2,
                              gx_, gy_, gm_);
      else
        vjp_near_sum<T, false>(sa, sb, half, ne, 0, xj, yj, mj, abx, aby, eps2, gx_, gy_, gm_);
    }
    gx_ += __shfl_xor(gx_, 32, 64);
    gy_ += __shfl_xor(gy_, 32, 64);
    gm_ += __shfl_xor(gm_, 32, 64);
    if (valid && half == 0) {
      if (slot == NONE) {
        xb[j] = (double)gx_;
        yb[j] = (double)gy_;
        mb[j] = (double)gm_;
      } else {
        double* pp = part + ((size_t)slot * n_max + (j - ts)) * 3;
        pp[0] = (double)gx_;
        pp[1] = (double)gy_;
        pp[2] = (double)gm_;
      }
    }
  }
}

// Adjoint of the FP32 forms' symmetric near field (k_p2p_sym's work items,
// leaves or halves, and pair layout): each unordered pair once, in FP32 on
// leaf-relative coordinates. With d = x_i - x_j, w = m_j abar_i - m_i abar_j
// and v = w / r^3 - 3 d (d . w) / r^5 (the gather form's terms for both bodies):
//   xbar_j += v, xbar_i -= v, mbar_j -= abar_i . d / r^3, mbar_i += abar_j . d / r^3.
// The i side (the item's leaf, sums per rotation K rotated back at the end)
// and, on its first item, the leaf's self block (every ordered pair, j side)
// go to the item's own slot; the j side to the pair's slot. sym_combine then
// adds each body's slots (xbar, ybar as the force pair, mbar as the third
// value). Padding bodies sit at 3e19 with zero mass and cotangent: r^2
// overflows to inf, 1/r = 0, and every term of a real body with them is 0.
__global__ void k_vjp_sym_prep(uint32_t nlv, const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                               const double* __restrict__ cx, const double* __restrict__ cy,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ m, const double* __restrict__ axb,
                               const double* __restrict__ ayb, float4* __restrict__ rel, float2* __restrict__ ab) {
  const uint32_t v = blockIdx.x * 8u + (threadIdx.x >> 5), e = threadIdx.x & 31u;
  if (v >= nlv) return;
  const uint32_t b = bs[v] + e;
  if (b >= be[v]) return;
  rel[b] = make_float4((float)(x[b] - cx[v]), (float)(y[b] - cy[v]), (float)m[b], 0.f);
  ab[b] = make_float2((float)axb[b], (float)ayb[b]);
}

__global__ void __launch_bounds__(64) k_vjp_sym(const float4* __restrict__ rel, const float2* __restrict__ ab,
                                                const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                                const double* __restrict__ cx, const double* __restrict__ cy,
                                                const uint32_t* __restrict__ own_off,
                                                const uint32_t* __restrict__ own_dat,
                                                const uint32_t* __restrict__ itgt, const uint32_t* __restrict__ iu0,
                                                const uint32_t* __restrict__ iu1, const uint32_t* __restrict__ iown,
                                                float eps2, double2* __restrict__ own2, double* __restrict__ ownp,
                                                float2* __restrict__ pair2, float* __restrict__ pairp) {
  constexpr float FAR = 3e19f;
  const uint32_t item = xcd_runs(blockIdx.x, gridDim.x);
  const uint32_t t = itgt[item], u0 = iu0[item], u1 = iu1[item], own = iown[item];
  const bool first = u0 == own_off[t];
  const int lane = threadIdx.x, r = lane >> 4, q = lane & 15;
  const uint32_t ts = bs[t], na = be[t] - ts;
  const double ox = cx[t], oy = cy[t];
  const uint32_t ia = 16u * (uint32_t)(r & 1) + (uint32_t)q;   // i-layout
  const uint32_t jb = 16u * (uint32_t)(r >> 1) + (uint32_t)q;  // j-layout
  float xi = FAR, yi = FAR, mi = 0.f, axi = 0.f, ayi = 0.f;
  if (ia < na) {
    const float4 v = rel[ts + ia];
    const float2 a = ab[ts + ia];
    xi = v.x;
    yi = v.y;
    mi = v.z;
    axi = a.x;
    ayi = a.y;
  }
  float gix[16], giy[16], gim[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) gix[k] = giy[k] = gim[k] = 0.f;
  // one pair (i rotated by K, j in the lane): the j side's terms; the i side's
  // into gix[K] etc. when BOTH
  auto pair = [&](auto kc, float xj, float yj, float mj, float axj, float ayj, bool zero, float& sx, float& sy,
                  float& sm, auto both) {
    constexpr int K = decltype(kc)::value;
    const float dx = row_rot<K>(xi) - xj, dy = row_rot<K>(yi) - yj;  // x_i - x_j
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2));
    float rs = Near<float>::rsq(r2);
    if (zero) rs = 0.f;
    const float rs2 = rs * rs, r3 = rs2 * rs, r5x3 = r3 * (3.f * rs2);
    const float mri = row_rot<K>(mi);
    const float wx = fmaf(mj, row_rot<K>(axi), -mri * axj), wy = fmaf(mj, row_rot<K>(ayi), -mri * ayj);
    const float tt = fmaf(dx, wx, dy * wy) * r5x3;
    const float vx = fmaf(wx, r3, -dx * tt), vy = fmaf(wy, r3, -dy * tt);
    sx += vx;
    sy += vy;
    sm -= fmaf(row_rot<K>(axi), dx, row_rot<K>(ayi) * dy) * r3;
    if constexpr (decltype(both)::value) {
      gix[K] -= vx;
      giy[K] -= vy;
      gim[K] += fmaf(axj, dx, ayj * dy) * r3;
    }
  };
  // self block: every ordered pair (i, j) of t, sums on the j side
  float sx = 0.f, sy = 0.f, sm = 0.f;
  if (first) {
    float xj = FAR, yj = FAR, mj = 0.f, axj = 0.f, ayj = 0.f;
    if (jb < na) {
      const float4 v = rel[ts + jb];
      const float2 a = ab[ts + jb];
      xj = v.x;
      yj = v.y;
      mj = v.z;
      axj = a.x;
      ayj = a.y;
    }
    const bool diag = (r & 1) == (r >> 1);  // rows 0 and 3: i == j at K = 0
    static_for<0, 16>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      pair(kc, xj, yj, mj, axj, ayj, K == 0 && diag, sx, sy, sm, std::false_type{});
    });
  }
  for (uint32_t u = u0; u < u1; ++u) {
    const uint32_t b = own_dat[u];
    const uint32_t b0 = bs[b], nb = be[b] - b0;
    float xj = FAR, yj = FAR, mj = 0.f, axj = 0.f, ayj = 0.f;
    if (jb < nb) {
      const float4 v = rel[b0 + jb];
      const float2 a = ab[b0 + jb];
      xj = v.x + (float)(cx[b] - ox);
      yj = v.y + (float)(cy[b] - oy);
      mj = v.z;
      axj = a.x;
      ayj = a.y;
    }
    float ajx = 0.f, ajy = 0.f, ajm = 0.f;
    static_for<0, 16>([&](auto kc) { pair(kc, xj, yj, mj, axj, ayj, false, ajx, ajy, ajm, std::true_type{}); });
    // rows (0, 1) and (2, 3) hold the same j-bodies
    ajx += __shfl_xor(ajx, 16, 64);
    ajy += __shfl_xor(ajy, 16, 64);
    ajm += __shfl_xor(ajm, 16, 64);
    if ((r & 1) == 0 && jb < nb) {
      const size_t o = (size_t)u * SYM_LEAF + jb;
      pair2[o] = make_float2(ajx, ajy);
      pairp[o] = ajm;
    }
  }
  // the i side: acc_i[K] of lane q belongs to the i of its rotated lane
  float tix = 0.f, tiy = 0.f, tim = 0.f;
  static_for<0, 16>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    tix += row_rot<16 - K>(gix[K]);
    tiy += row_rot<16 - K>(giy[K]);
    tim += row_rot<16 - K>(gim[K]);
  });
  tix += __shfl_xor(tix, 32, 64);  // rows r and r ^ 2 hold the same i-bodies
  tiy += __shfl_xor(tiy, 32, 64);
  tim += __shfl_xor(tim, 32, 64);
  if (first) {  // uniform over the block: the self block's j-side sums to the i layout
    sx += __shfl_xor(sx, 16, 64);
    sy += __shfl_xor(sy, 16, 64);
    sm += __shfl_xor(sm, 16, 64);
    const int src = lane < 16 ? lane : lane + 16;  // body 16 + q sits in row 2
    sx = __shfl(sx, src, 64);
    sy = __shfl(sy, src, 64);
    sm = __shfl(sm, src, 64);
  } else {
    sx = sy = sm = 0.f;
  }
  if (lane < 32 && (uint32_t)lane < na) {
    const size_t o = (size_t)own * SYM_LEAF + (size_t)lane;
    own2[o] = make_double2((double)(tix + sx), (double)(tiy + sy));
    ownp[o] = (double)(tim + sm);
  }
}

// ------------------------------------------------------ L2P adjoint ----

// Lbar of each leaf: Lbar[IJ] = -sum_b a_bar_b . grad(u^I v^J)(u_b). One wave
// per leaf; chunks of 32 bodies stage their power tables u^0..u^P, v^0..v^P and
// cotangents in LDS, then each lane sums its coefficient over the chunk.
template <int P>
__global__ void __launch_bounds__(256) k_vjp_l2p_coef(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ axb, const double* __restrict__ ayb,
                                                      const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be,
                                                      const double* __restrict__ comx, const double* __restrict__ comy,
                                                      uint32_t n_leaves, double* __restrict__ lbar) {
  constexpr int S = num_moments(P), CH = 32;
  __shared__ double pu[4][CH][P + 1], pv[4][CH][P + 1], gx_[4][CH], gy_[4][CH];
  __shared__ uint32_t blen[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t leaf = blockIdx.x * 4 + w;
  const bool live = leaf < n_leaves;
  const uint32_t s0 = live ? bs[leaf] : 0u, s1 = live ? be[leaf] : 0u;
  const double cx = live ? comx[leaf] : 0.0, cy = live ? comy[leaf] : 0.0;
  if (lane == 0) blen[w] = s1 - s0;
  __syncthreads();
  const uint32_t len = max(max(blen[0], blen[1]), max(blen[2], blen[3]));
  for (int k0 = 0; k0 < S; k0 += 64) {
    const int k = k0 + lane;
    int I = 0, J = 0;
    if (k < S) cij<P>(k, I, J);
    double acc = 0.0;
    for (uint32_t c = 0; c < len; c += CH) {
      __syncthreads();
      if (lane < CH) {
        const uint32_t b = s0 + c + lane;
        double u = 0.0, v = 0.0, gxv = 0.0, gyv = 0.0;
        if (b < s1) {
          u = x[b] - cx;
          v = y[b] - cy;
          gxv = axb[b];
          gyv = ayb[b];
        }
        double a = 1.0, bb = 1.0;
#pragma unroll
        for (int q = 0; q <= P; ++q) {
          pu[w][lane][q] = a;
          pv[w][lane][q] = bb;
          a *= u;
          bb *= v;
        }
        gx_[w][lane] = gxv;
        gy_[w][lane] = gyv;
      }
      __syncthreads();
      const uint32_t n = s1 > s0 + c ? min((uint32_t)CH, s1 - s0 - c) : 0u;
      if (k < S) {
        for (uint32_t q = 0; q < n; ++q) {
          const double gu = (double)I * pu[w][q][I > 0 ? I - 1 : 0] * pv[w][q][J];
          const double gv = (double)J * pu[w][q][I] * pv[w][q][J > 0 ? J - 1 : 0];
          acc -= gx_[w][q] * gu + gy_[w][q] * gv;
        }
      }
    }
    if (live && k < S) lbar[(size_t)leaf * S + k] = acc;
  }
}

// per body: far-field Hessian term and the P2M adjoint (leaf COM frozen)
template <int P>
__global__ void k_vjp_body(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ m,
                           const double* __restrict__ axb, const double* __restrict__ ayb,
                           const uint32_t* __restrict__ leaf_of, const double* __restrict__ comx,
                           const double* __restrict__ comy, const double* __restrict__ loc,
                           const double* __restrict__ mbar, size_t n, double* __restrict__ xb, double* __restrict__ yb,
                           double* __restrict__ mb) {
  constexpr int S = num_moments(P);
  size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const uint32_t lf = leaf_of[b];
  const double u = x[b] - comx[lf], v = y[b] - comy[lf];
  double up[P + 1], vp[P + 1];
  up[0] = 1.0;
  vp[0] = 1.0;
#pragma unroll
  for (int q = 1; q <= P; ++q) {
    up[q] = up[q - 1] * u;
    vp[q] = vp[q - 1] * v;
  }
  const double* L = loc + (size_t)lf * S;
  const double* Mb = mbar + (size_t)lf * S;
  double hxx = 0.0, hxy = 0.0, hyy = 0.0, px = 0.0, py = 0.0, pm = 0.0;
  double fi = 1.0;
#pragma unroll
  for (int i = 0; i <= P; ++i) {
    if (i > 0) fi *= i;
    double fj = 1.0;
#pragma unroll
    for (int j = 0; j <= P - i; ++j) {
      if (j > 0) fj *= j;
      const double c = L[moment_index(i, j)];
      if (i >= 2) hxx += c * i * (i - 1) * up[i - 2] * vp[j];
      if (i >= 1 && j >= 1) hxy += c * i * j * up[i - 1] * vp[j - 1];
      if (j >= 2) hyy += c * j * (j - 1) * up[i] * vp[j - 2];
      const double mm = Mb[moment_index(i, j)] / (fi * fj);
      if (i >= 1) px += mm * i * up[i - 1] * vp[j];
      if (j >= 1) py += mm * j * up[i] * vp[j - 1];
      pm += mm * up[i] * vp[j];
    }
  }
  const double ax_ = axb[b], ay_ = ayb[b], mass = m[b];
  xb[b] += -(ax_ * hxx + ay_ * hxy) + mass * px;
  yb[b] += -(ax_ * hxy + ay_ * hyy) + mass * py;
  mb[b] += pm;
}

__global__ void k_leaf_of(const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be, uint32_t n_leaves,
                          uint32_t* __restrict__ leaf_of) {
  uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_leaves) return;
  for (uint32_t b = bs[l]; b < be[l]; ++b) leaf_of[b] = l;
}

// ------------------------------------------------------ L2L adjoint ----

// factorials and binomials up to MAX_P, filled at compile time
struct VTabs {
  double fact[MAX_P + 1];
  double binom[MAX_P + 1][MAX_P + 1];
};
constexpr VTabs make_vtabs() {
  VTabs t{};
  t.fact[0] = 1.0;
  for (int n = 1; n <= MAX_P; ++n) t.fact[n] = t.fact[n - 1] * n;
  for (int n = 0; n <= MAX_P; ++n) {
    t.binom[n][0] = 1.0;
    for (int k = 1; k <= n; ++k) t.binom[n][k] = t.binom[n][k - 1] * (n - k + 1) / k;
  }
  return t;
}
static __constant__ VTabs v_tabs = make_vtabs();

// parents in [lo, hi): Lbar_P[i,j] += sum_c sum_{oi<=i, oj<=j} Lbar_c[oi,oj] C(i,oi) C(j,oj) dx^(i-oi) dy^(j-oj)
template <int P>
__global__ void k_vjp_l2l(const uint32_t* __restrict__ cf, const uint8_t* __restrict__ cc,
                          const double* __restrict__ comx, const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                          double* __restrict__ lbar) {
  constexpr int S = num_moments(P);
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)(hi - lo) * S) return;
  uint32_t p = lo + (uint32_t)(t / S);
  int k = (int)(t % S);
  int i, j;
  cij<P>(k, i, j);
  double acc = 0.0;
  for (uint32_t c = cf[p]; c < cf[p] + cc[p]; ++c) {
    const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
    const double* lc = lbar + (size_t)c * S;
    double xw = 1.0;  // dx^(i - oi)
    for (int oi = i; oi >= 0; --oi) {
      double row = 0.0, yw = 1.0;  // yw = dy^(j - oj)
      for (int oj = j; oj >= 0; --oj) {
        row += lc[moment_index(oi, oj)] * v_tabs.binom[j][oj] * yw;
        yw *= dy;
      }
      acc += row * v_tabs.binom[i][oi] * xw;
      xw *= dx;
    }
  }
  lbar[(size_t)p * S + k] += acc;
}

// ------------------------------------------------------ M2M adjoint ----
// children in [lo, hi): Mbar_c[i,j] += sum_{I>=i, J>=j, I+J<=p} Mbar_P[I,J] dx^(I-i) dy^(J-j) / ((I-i)!(J-j)!)
template <int P>
__global__ void k_vjp_m2m(const uint32_t* __restrict__ par, const double* __restrict__ comx,
                          const double* __restrict__ comy, uint32_t lo, uint32_t hi, double* __restrict__ mbar) {
  constexpr int S = num_moments(P);
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)(hi - lo) * S) return;
  uint32_t c = lo + (uint32_t)(t / S);
  int k = (int)(t % S);
  uint32_t p = par[c];
  if (p == NONE) return;
  int i, j;
  cij<P>(k, i, j);
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  const double* mp_ = mbar + (size_t)p * S;
  double acc = 0.0;
  double xw = 1.0;  // dx^(I - i) / (I - i)!
  for (int I = i; I <= P; ++I) {
    double yw = 1.0;  // dy^(J - j) / (J - j)!
    for (int J = j; J <= P - I; ++J) {
      acc += mp_[moment_index(I, J)] * xw * yw;
      yw *= dy / (J - j + 1);
    }
    xw *= dx / (I - i + 1);
  }
  mbar[(size_t)c * S + k] += acc;
}

// CORRECTED variant, levels whose parents each have one child: the two
// adjoints with a thread per (node, coefficient). A child at its parent's COM
// (the usual case) shifts by zero, so each adjoint adds the other node's
// cotangent coefficient for coefficient (what the full shifts give bit for
// bit); an off-COM child takes its coefficient's full adjoint shift.
__device__ __forceinline__ void vj_coef_ij(int k, int& i, int& j) {
  int n = 0;
  while ((n + 1) * (n + 2) / 2 <= k) ++n;
  j = k - n * (n + 1) / 2;
  i = n - j;
}

__device__ __forceinline__ double vj_binom(int n, int k) {
  double r = 1.0;
  for (int q = 1; q <= k; ++q) r = r * (n - k + q) / q;
  return r;
}

// parents in [lo, hi): Lbar_p(i, j) += sum_{oi <= i, oj <= j} Lbar_c(oi, oj) C(i, oi) C(j, oj) dx^(i-oi) dy^(j-oj)
template <int P>
__global__ void __launch_bounds__(256) k_vjp_l2l_single(const uint32_t* __restrict__ cf,
                                                        const double* __restrict__ comx,
                                                        const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                        double* __restrict__ lbar) {
  constexpr int S = num_moments(P);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t p = lo + (uint32_t)(t / S);
  const int k = (int)(t % S);
  if (p >= hi) return;
  const uint32_t c = cf[p];
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  const double* lc = lbar + (size_t)c * S;
  double v;
  if (dx == 0.0 && dy == 0.0) {
    v = lc[k];
  } else {
    int i, j;
    vj_coef_ij(k, i, j);
    v = 0.0;
    for (int oi = 0; oi <= i; ++oi)
      for (int oj = 0; oj <= j; ++oj) {
        double e = vj_binom(i, oi) * vj_binom(j, oj);
        for (int q = 0; q < i - oi; ++q) e *= dx;
        for (int q = 0; q < j - oj; ++q) e *= dy;
        v = fma(lc[moment_index(oi, oj)], e, v);
      }
  }
  lbar[(size_t)p * S + k] += v;
}

// children in [lo, hi): Mbar_c(i, j) += sum_{I >= i, J >= j} Mbar_p(I, J) dx^(I-i) / (I-i)! dy^(J-j) / (J-j)!
template <int P>
__global__ void __launch_bounds__(256) k_vjp_m2m_single(const uint32_t* __restrict__ par,
                                                        const double* __restrict__ comx,
                                                        const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                        double* __restrict__ mbar) {
  constexpr int S = num_moments(P);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = lo + (uint32_t)(t / S);
  const int k = (int)(t % S);
  if (c >= hi) return;
  const uint32_t p = par[c];
  if (p == NONE) return;
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
  const double* mp_ = mbar + (size_t)p * S;
  double v;
  if (dx == 0.0 && dy == 0.0) {
    v = mp_[k];
  } else {
    int i, j;
    vj_coef_ij(k, i, j);
    v = 0.0;
    double ta = 1.0;
    for (int I = i; I <= P - j; ++I) {
      double tb = 1.0;
      for (int J = j; I + J <= P; ++J) {
        v = fma(mp_[moment_index(I, J)], ta * tb, v);
        tb *= dy / (J - j + 1);
      }
      ta *= dx / (I - i + 1);
    }
  }
  mbar[(size_t)c * S + k] += v;
}

// CORRECTED variant, p <= 8: the same two adjoints with one thread per node,
// the cotangents and the offset's powers in registers and the shift
// constants known at compile time (as k_m2m_nodes / k_l2l_nodes).
template <int P>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_vjp_l2l_nodes(const uint32_t* __restrict__ cf,
                                                       const uint8_t* __restrict__ cc,
                                                       const double* __restrict__ comx,
                                                       const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                       double* __restrict__ lbar) {
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
    double ex[P + 1], ey[P + 1];
    ex[0] = ey[0] = 1.0;
#pragma unroll
    for (int a = 1; a <= P; ++a) {
      ex[a] = ex[a - 1] * dx;
      ey[a] = ey[a - 1] * dy;
    }
    const double* lc = lbar + (size_t)c * S;
    double l[S];
#pragma unroll
    for (int k = 0; k < S; ++k) l[k] = lc[k];
    static_for<0, P + 1>([&](auto nc_) {
      constexpr int n = decltype(nc_)::value;
      static_for<0, n + 1>([&](auto jc) {
        constexpr int j = decltype(jc)::value, i = n - j;
        double sum = acc[moment_index(i, j)];
        static_for<0, i + 1>([&](auto oic) {
          constexpr int oi = decltype(oic)::value;
          static_for<0, j + 1>([&](auto ojc) {
            constexpr int oj = decltype(ojc)::value;
            constexpr double bc = ct_binom(i, oi) * ct_binom(j, oj);
            sum = fma(l[moment_index(oi, oj)] * bc, ex[i - oi] * ey[j - oj], sum);
          });
        });
        acc[moment_index(i, j)] = sum;
      });
    });
  }
  double* out = lbar + (size_t)p * S;
#pragma unroll
  for (int k = 0; k < S; ++k) out[k] += acc[k];
}

template <int P>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_vjp_m2m_nodes(const uint32_t* __restrict__ par,
                                                       const double* __restrict__ comx,
                                                       const double* __restrict__ comy, uint32_t lo, uint32_t hi,
                                                       double* __restrict__ mbar) {
  constexpr int S = num_moments(P);
  const uint32_t c = lo + blockIdx.x * 64u + threadIdx.x;
  if (c >= hi) return;
  const uint32_t p = par[c];
  if (p == NONE) return;
  const double dx = comx[c] - comx[p], dy = comy[c] - comy[p];
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
  const double* mp_ = mbar + (size_t)p * S;
  double mv[S];
#pragma unroll
  for (int k = 0; k < S; ++k) mv[k] = mp_[k];
  double* mc = mbar + (size_t)c * S;
  static_for<0, P + 1>([&](auto nc_) {
    constexpr int n = decltype(nc_)::value;
    static_for<0, n + 1>([&](auto jc) {
      constexpr int j = decltype(jc)::value, i = n - j;
      double sum = 0.0;
      static_for<i, P + 1>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        static_for<j, P - I + 1>([&](auto Jc) {
          constexpr int J = decltype(Jc)::value;
          sum = fma(mv[moment_index(I, J)], tx[I - i] * ty[J - j], sum);
        });
      });
      mc[moment_index(i, j)] += sum;
    });
  });
}

// ------------------------------------------------------ M2L adjoint ----
// Mbar_s[i,j] += sum_pairs sum_{I,J} (-1)^{i+j}/(I!J!) G[i+I, j+J] Lbar_t[I,J]

// corrected variant: lists are symmetric; exact derivatives per pair, built
// in LDS by the same recurrence as the forward (D = derivatives of 1/R, G = -D)
template <int P>
__global__ void __launch_bounds__(256) k_vjp_m2l_exact(const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ dat, uint32_t nn,
                                                       const double* __restrict__ comx,
                                                       const double* __restrict__ comy, double eps2,
                                                       const double* __restrict__ lbar, double* __restrict__ mbar) {
  constexpr int S = num_moments(P), ST = 2 * P + 1;
  __shared__ double sD[4][ST * ST];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t s = blockIdx.x * 4 + w;
  if (s >= nn) return;
  double* D = sD[w];
  double acc[(S + 63) / 64];
#pragma unroll
  for (int q = 0; q < (S + 63) / 64; ++q) acc[q] = 0.0;
  for (uint32_t r = off[s]; r < off[s + 1]; ++r) {
    const uint32_t t = dat[r];
    const double x = comx[t] - comx[s], y = comy[t] - comy[s];
    const double r2 = x * x + y * y + eps2, ir2 = 1.0 / r2;
    if (lane == 0) D[0] = 1.0 / sqrt(r2);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int n = 1; n <= 2 * P; ++n) {
      for (int k = lane; k <= n; k += 64) {
        int l = n - k;
        double v;
        if (k >= 1) {
          v = -(2.0 * k - 1.0) * x * D[(n - 1) * ST + l];
          if (k >= 2) v -= (double)(k - 1) * (k - 1) * D[(n - 2) * ST + l];
          if (l >= 1) v -= 2.0 * l * y * D[(n - 1) * ST + (l - 1)];
          if (l >= 2) v -= (double)l * (l - 1) * D[(n - 2) * ST + (l - 2)];
        } else {
          v = -(2.0 * l - 1.0) * y * D[(n - 1) * ST + (l - 1)];
          if (l >= 2) v -= (double)(l - 1) * (l - 1) * D[(n - 2) * ST + (l - 2)];
        }
        D[n * ST + l] = v * ir2;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    const double* Lb = lbar + (size_t)t * S;
#pragma unroll
    for (int q = 0; q < (S + 63) / 64; ++q) {
      int k = lane + 64 * q;
      if (k < S) {
        int i, j;
        cij<P>(k, i, j);
        const double sg = ((i + j) & 1) ? 1.0 : -1.0;  // (-1)^{i+j} * (G = -D)
        double a = 0.0;
#pragma unroll
        for (int I = 0; I <= P; ++I)
#pragma unroll
          for (int J = 0; J <= P - I; ++J)
            a += D[(i + I + j + J) * ST + (j + J)] * Lb[moment_index(I, J)] / (dfact(I) * dfact(J));
        acc[q] += sg * a;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
#pragma unroll
  for (int q = 0; q < (S + 63) / 64; ++q) {
    int k = lane + 64 * q;
    if (k < S) mbar[(size_t)s * S + k] = acc[q];
  }
}

// Adjoint of the corrected M2L, p <= 8: one wave per source s, one lane per
// partner t (the lists are symmetric), as k_m2l_exact_lp: the lane runs the
// pair's derivative recurrence and scatters it against lbar_t scaled by
// 1/(I!J!) (m2l_pair_scatter); the wave sums its lanes in a fixed order.
template <int P>
__global__ void __launch_bounds__(64) k_vjp_m2l_exact_lp(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ dat, const uint32_t* __restrict__ order,
                                                         const double* __restrict__ comx,
                                                         const double* __restrict__ comy, double eps2,
                                                         const double* __restrict__ lbar, double* __restrict__ mbar) {
  constexpr int S = num_moments(P);
  const uint32_t s = order[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t r0 = off[s], r1 = off[s + 1];
  const double sx = comx[s], sy = comy[s];
  double acc = 0.0;  // coefficient `lane`
  for (uint32_t rb = r0; rb < r1; rb += 64) {  // uniform
    const uint32_t r = rb + (uint32_t)lane;
    double A[S];
#pragma unroll
    for (int k = 0; k < S; ++k) A[k] = 0.0;
    if (r < r1) {
      const uint32_t t = dat[r];
      const double x = comx[t] - sx, y = comy[t] - sy;
      const double r2 = x * x + y * y + eps2;
      const double* Lb = lbar + (size_t)t * S;
      double lv[S];
#pragma unroll
      for (int q = 0; q <= P; ++q)
#pragma unroll
        for (int J = 0; J <= q; ++J) lv[moment_index(q - J, J)] = Lb[moment_index(q - J, J)] / (dfact(q - J) * dfact(J));
      // A(i, j) += D(i + I, j + J) lv(I, J)
      m2l_pair_scatter<P, double>(x, y, 1.0 / r2, 1.0 / sqrt(r2), lv, A);
    }
    acc += wave_sum_scatter<S>(A, lane);
  }
  if (lane < S) {
    int i, j;
    cij<P>(lane, i, j);
    const double sg = ((i + j) & 1) ? 1.0 : -1.0;  // (-1)^{i+j} * (G = -D)
    mbar[(size_t)s * S + lane] = sg * acc;
  }
}

template <int P>
static int vjp_impl(galax_fmm* h, const double* x, const double* y, const double* m, const double* axb,
                    const double* ayb, double* xb, double* yb, double* mb) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  constexpr int S = num_moments(P);
  const bool corrected = h->cfg.algo == GALAX_ALGO_CORRECTED;
  const size_t nn = t.n_nodes, nl = t.n_leaves, n = t.n_bodies;
  const double eps2 = h->cfg.softening * h->cfg.softening;
  const double* comx = t.com_x.as<double>();
  const double* comy = t.com_y.as<double>();
  GX_CHECK(dbuf_reserve(h->vj_lbar, nn * S * 8, s));
  GX_CHECK(dbuf_reserve(h->vj_mbar, nn * S * 8, s));
  GX_CHECK(dbuf_reserve(h->vj_leaf_of, n * 4, s));
  double* lbar = h->vj_lbar.as<double>();
  double* mbar = h->vj_mbar.as<double>();
  GX_HIP(hipMemsetAsync(lbar, 0, nn * S * 8, s));
  GX_HIP(hipMemsetAsync(mbar, 0, nn * S * 8, s));
  const uint32_t* goff = corrected ? h->c_gat_off.as<uint32_t>() : t.gat_off.as<uint32_t>();
  const uint32_t* gdat = corrected ? h->c_gat_dat.as<uint32_t>() : t.gat_dat.as<uint32_t>();
  // near field
  const P2PItems& it = corrected ? h->c_items : t.items;
  const uint32_t nmax = h->cfg.n_max;
  GX_CHECK(dbuf_reserve(h->p2p_part, std::max<size_t>(it.n_part, 1) * nmax * 3 * 8, s));
  const bool fp32_near =
      h->cfg.precision == GALAX_PREC_MIXED_FP32 || h->cfg.precision == GALAX_PREC_MIXED_BF16X3;
  if (fp32_near && h->cfg.n_max <= 2 * SYM_LEAF && !h->near_gather && nl > 0) {
    // the FP32 forms: each unordered leaf pair once (the forward's structure)
    SymP2P& sp = corrected ? h->c_sym : t.sym;
    if (!sp.valid) GX_CHECK(build_sym_near(h, goff, gdat, (uint32_t)nl, sp));
    GX_CHECK(dbuf_reserve(sp.ownp, std::max<size_t>(sp.n_items, 1) * SYM_LEAF * 8, s));
    GX_CHECK(dbuf_reserve(sp.pairp, std::max<size_t>(sp.n_pairs, 1) * SYM_LEAF * 4, s));
    GX_CHECK(dbuf_reserve(sp.cpartp, std::max<size_t>(sp.n_cpart, 1) * SYM_LEAF * 8, s));
    GX_CHECK(dbuf_reserve(sp.ab2, std::max<size_t>(n, 1) * 8, s));
    const uint32_t* sbs = sp.halves ? sp.vbs.as<uint32_t>() : t.body_start.as<uint32_t>();
    const uint32_t* sbe = sp.halves ? sp.vbe.as<uint32_t>() : t.body_end.as<uint32_t>();
    const double* scx = sp.halves ? sp.vcx.as<double>() : comx;
    const double* scy = sp.halves ? sp.vcy.as<double>() : comy;
    const uint32_t nlv = (uint32_t)sp.n_leaves;
    k_vjp_sym_prep<<<(nlv + 7) / 8, 256, 0, s>>>(nlv, sbs, sbe, scx, scy, x, y, m, axb, ayb, sp.rel.as<float4>(),
                                                 sp.ab2.as<float2>());
    GX_HIP(hipGetLastError());
    k_vjp_sym<<<(unsigned)sp.n_items, 64, 0, s>>>(
        sp.rel.as<float4>(), sp.ab2.as<float2>(), sbs, sbe, scx, scy, sp.own_off.as<uint32_t>(),
        sp.own_dat.as<uint32_t>(), sp.itgt.as<uint32_t>(), sp.iu0.as<uint32_t>(), sp.iu1.as<uint32_t>(),
        sp.iown.as<uint32_t>(), (float)eps2, sp.own2.as<double2>(), sp.ownp.as<double>(), sp.pair2.as<float2>(),
        sp.pairp.as<float>());
    GX_HIP(hipGetLastError());
    GX_CHECK(sym_combine(sp, sbs, sbe, xb, yb, mb, s));
  } else {
#define GX_LAUNCH_VJP_NEAR(T)                                                                                  \
  k_vjp_near<T><<<(unsigned)it.n_items, 64, 0, s>>>(                                                          \
      x, y, m, axb, ayb, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), t.com_x.as<double>(),          \
      t.com_y.as<double>(), gdat, it.tgt.as<uint32_t>(), it.g0.as<uint32_t>(), it.g1.as<uint32_t>(),            \
      it.slot.as<uint32_t>(), nmax, eps2, xb, yb, mb, h->p2p_part.as<double>())
  if (fp32_near)
    GX_LAUNCH_VJP_NEAR(float);
  else
    GX_LAUNCH_VJP_NEAR(double);
#undef GX_LAUNCH_VJP_NEAR
    GX_HIP(hipGetLastError());
    GX_HIP(p2p_combine(it, goff, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), nmax,
                       h->p2p_part.as<double>(), xb, yb, mb, s));
  }
  // L2P adjoint -> Lbar of leaves
  k_vjp_l2p_coef<P><<<nblk(nl, 4), 256, 0, s>>>(x, y, axb, ayb, t.body_start.as<uint32_t>(),
                                                t.body_end.as<uint32_t>(), comx, comy, (uint32_t)nl, lbar);
  GX_HIP(hipGetLastError());
  // L2L adjoint, bottom-up (reference: level 1 only)
  const uint32_t top = corrected ? t.n_levels : std::min<uint32_t>(t.n_levels, 2);
  for (uint32_t l = 1; l < top; ++l) {
    uint32_t lo = t.level_offset[l], hi = t.level_offset[l + 1];
    if (corrected && hi - lo == lo - t.level_offset[l - 1]) {  // every parent has one child
      k_vjp_l2l_single<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.child_first.as<uint32_t>(), comx, comy, lo,
                                                                       hi, lbar);
      GX_HIP(hipGetLastError());
      continue;
    }
    if constexpr (P <= 8) {
      if (corrected) {
        k_vjp_l2l_nodes<P><<<nblk(hi - lo, 64), 64, 0, s>>>(t.child_first.as<uint32_t>(), t.child_count.as<uint8_t>(),
                                                          comx, comy, lo, hi, lbar);
        GX_HIP(hipGetLastError());
        continue;
      }
    }
    k_vjp_l2l<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.child_first.as<uint32_t>(),
                                                             t.child_count.as<uint8_t>(), comx, comy, lo, hi, lbar);
    GX_HIP(hipGetLastError());
  }
  // M2L adjoint
  if (corrected) {
    if constexpr (P <= 8) {  // above, the lane's window and sums no longer fit in registers
      k_vjp_m2l_exact_lp<P><<<(unsigned)nn, 64, 0, s>>>(h->c_m2l_off.as<uint32_t>(), h->c_m2l_dat.as<uint32_t>(),
                                                        h->c_m2l_order.as<uint32_t>(), comx, comy, eps2, lbar, mbar);
    } else {
      k_vjp_m2l_exact<P><<<nblk(nn, 4), 256, 0, s>>>(h->c_m2l_off.as<uint32_t>(), h->c_m2l_dat.as<uint32_t>(),
                                                     (uint32_t)nn, comx, comy, eps2, lbar, mbar);
    }
  } else if (h->n_live_pairs > 0) {
    // mbar_s = sum over live pairs (t, s) of T_K^T lbar_t, on the matrix cores
    if (!h->mx_adj_valid) GX_CHECK(m2l_adjoint_setup(h));
    GX_CHECK(m2l_adjoint_run(h, lbar, mbar));
  }
  GX_HIP(hipGetLastError());
  // M2M adjoint, top-down
  for (uint32_t l = top; l-- > 1;) {
    uint32_t lo = t.level_offset[l - 1], hi = t.level_offset[l];
    if (corrected && t.level_offset[l + 1] - hi == hi - lo) {  // every parent has one child
      k_vjp_m2m_single<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi,
                                                                       mbar);
      GX_HIP(hipGetLastError());
      continue;
    }
    if constexpr (P <= 8) {
      if (corrected) {
        k_vjp_m2m_nodes<P><<<nblk(hi - lo, 64), 64, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi, mbar);
        GX_HIP(hipGetLastError());
        continue;
      }
    }
    k_vjp_m2m<P><<<nblk((size_t)(hi - lo) * S), 256, 0, s>>>(t.parent.as<uint32_t>(), comx, comy, lo, hi, mbar);
    GX_HIP(hipGetLastError());
  }
  // far-field Hessian + P2M adjoint per body
  k_leaf_of<<<nblk(nl), 256, 0, s>>>(t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), (uint32_t)nl,
                                     h->vj_leaf_of.as<uint32_t>());
  k_vjp_body<P><<<nblk(n), 256, 0, s>>>(x, y, m, axb, ayb, h->vj_leaf_of.as<uint32_t>(), comx, comy,
                                        h->loc.as<double>(), mbar, n, xb, yb, mb);
  GX_HIP(hipGetLastError());
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

// forward (recomputing M and L from the given positions, in the handle's
// precision form) then the adjoint passes, all in FP64
int vjp(galax_fmm* h, const double* d_x, const double* d_y, const double* d_m, const double* d_axb,
        const double* d_ayb, double* d_xb, double* d_yb, double* d_mb) {
  size_t n = h->tree.n_bodies;
  GX_CHECK(dbuf_reserve(h->vj_fwd, std::max<size_t>(n, 1) * 16, h->stream));
  double* fax = h->vj_fwd.as<double>();
  // the adjoint needs the expansions (loc for the L2P Hessian), not the forces
  GX_CHECK(evaluate(h, const_cast<double*>(d_x), const_cast<double*>(d_y), const_cast<double*>(d_m), fax, fax + n,
                    nullptr, false));
  switch (h->cfg.p) {
#define GX_CASE(PP) \
  case PP:          \
    return vjp_impl<PP>(h, d_x, d_y, d_m, d_axb, d_ayb, d_xb, d_yb, d_mb);
    GX_CASE(0) GX_CASE(1) GX_CASE(2) GX_CASE(3) GX_CASE(4) GX_CASE(5) GX_CASE(6) GX_CASE(7) GX_CASE(8)
    GX_CASE(9) GX_CASE(10) GX_CASE(11) GX_CASE(12) GX_CASE(13) GX_CASE(14) GX_CASE(15) GX_CASE(16)
#undef GX_CASE
  }
  return set_error(GALAX_ERR_INVALID, "p out of range");
}

}  // namespace gx
