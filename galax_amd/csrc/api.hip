// api.hip — the extern "C" entry points of include/galax_fmm.h and the
// host-side orchestration around the device passes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"

using namespace gx;

namespace gx {

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

// --------------------------------------------- corrected-variant lists ----

// R_n = max |body - com_n| over the node's bodies. The bodies of a node are
// cut into chunks of RAD_CHUNK, one wave per (node, chunk) item (a node near
// the root has most of the bodies: one wave per node left them to one wave);
// each item's max of the squared distances goes into the node's slot by an
// integer atomic max on its bit pattern (non-negative doubles order as their
// bits; fmax drops NaN, so none arrives), then one pass takes the roots.
// Max is exact in any order: the radii are those of a serial loop.
constexpr uint32_t RAD_CHUNK = 4096;

__global__ void k_radii_items(const uint32_t* __restrict__ bs, const uint32_t* __restrict__ be, uint32_t nn,
                              uint32_t* __restrict__ nit) {
  const uint32_t nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  nit[nd] = max(1u, (be[nd] - bs[nd] + RAD_CHUNK - 1) / RAD_CHUNK);
}

__global__ void k_radii(const double* __restrict__ x, const double* __restrict__ y, const uint32_t* __restrict__ bs,
                        const uint32_t* __restrict__ be, const double* __restrict__ comx,
                        const double* __restrict__ comy, uint32_t nn, const uint32_t* __restrict__ ioff,
                        uint32_t nitems, unsigned long long* __restrict__ r2bits) {
#pragma clang fp contract(off)
  const uint32_t item = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (item >= nitems) return;
  uint32_t lo = 0, hi = nn;  // the node nd with ioff[nd] <= item < ioff[nd + 1]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (ioff[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t nd = lo;
  const uint32_t b0 = bs[nd] + (item - ioff[nd]) * RAD_CHUNK, b1 = min(be[nd], b0 + RAD_CHUNK);
  const double cx = comx[nd], cy = comy[nd];
  double r2 = 0.0;
  for (uint32_t b = b0 + lane; b < b1; b += 64) {
    double dx = x[b] - cx, dy = y[b] - cy;
    r2 = fmax(r2, dx * dx + dy * dy);
  }
  for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
  if (lane == 0) atomicMax(r2bits + nd, (unsigned long long)__double_as_longlong(r2));
}

__global__ void k_radii_sqrt(uint32_t nn, double* __restrict__ R) {
  const uint32_t nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd < nn) R[nd] = sqrt(R[nd]);  // R holds the squared radius's bits, read as a double
}

// The dual tree traversal, breadth-first on the device. Every visited pair
// (a, b) is kept in one buffer (a in the high 32 bits). A round turns each
// pair of the previous round that is neither accepted nor a leaf pair into
// its child pairs: both children lists of a self pair (i <= j), else the
// children of the larger-radius non-leaf node. The pairs visited, hence the
// lists, are those of the recursive traversal whatever the order of visits.
// One pass then emits the accepted pairs (both directions) and the leaf
// pairs as sort keys; sorted, they give each node's ascending M2L list and
// each leaf's P2P list, the leaf itself first (the key 0 of its row).
struct DttTree {
  const double *comx, *comy, *R;
  const uint32_t* cf;
  const uint8_t* cc;
  double theta;
};

enum : uint32_t { DTT_SPLIT_A, DTT_SPLIT_B, DTT_SELF, DTT_LEAF_SELF, DTT_M2L, DTT_P2P };

__device__ __forceinline__ uint32_t dtt_kind(const DttTree& t, uint32_t a, uint32_t b) {
#pragma clang fp contract(off)
  const bool la = t.cc[a] == 0, lb = t.cc[b] == 0;
  if (a == b) return la ? DTT_LEAF_SELF : DTT_SELF;
  const double dx = t.comx[a] - t.comx[b], dy = t.comy[a] - t.comy[b];
  const double d = sqrt(dx * dx + dy * dy);
  if (t.R[a] + t.R[b] < t.theta * d) return DTT_M2L;
  if (la && lb) return DTT_P2P;
  return (lb || (!la && t.R[a] >= t.R[b])) ? DTT_SPLIT_A : DTT_SPLIT_B;
}

__device__ __forceinline__ uint32_t dtt_children(const DttTree& t, uint32_t kind, uint32_t a, uint32_t b) {
  if (kind == DTT_SELF) return (uint32_t)t.cc[a] * (t.cc[a] + 1u) / 2u;
  if (kind == DTT_SPLIT_A) return t.cc[a];
  if (kind == DTT_SPLIT_B) return t.cc[b];
  return 0;
}

__global__ void k_dtt_count(DttTree t, const uint64_t* __restrict__ pairs, uint32_t lo, uint32_t hi,
                            uint32_t* __restrict__ cnt) {
  const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const uint64_t pr = pairs[i];
  const uint32_t a = (uint32_t)(pr >> 32), b = (uint32_t)pr;
  cnt[i - lo] = dtt_children(t, dtt_kind(t, a, b), a, b);
}

// child pairs of [lo, hi) at hi + pos[i - lo]
__global__ void k_dtt_expand(DttTree t, uint64_t* pairs, uint32_t lo, uint32_t hi, const uint32_t* __restrict__ pos) {
  const uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const uint64_t pr = pairs[i];
  const uint32_t a = (uint32_t)(pr >> 32), b = (uint32_t)pr;
  const uint32_t kind = dtt_kind(t, a, b);
  uint64_t* out = pairs + hi + pos[i - lo];
  if (kind == DTT_SELF) {
    const uint32_t c0 = t.cf[a], nc = t.cc[a];
    for (uint32_t ci = 0; ci < nc; ++ci)
      for (uint32_t cj = ci; cj < nc; ++cj) *out++ = ((uint64_t)(c0 + ci) << 32) | (c0 + cj);
  } else if (kind == DTT_SPLIT_A) {
    for (uint32_t ci = 0; ci < t.cc[a]; ++ci) *out++ = ((uint64_t)(t.cf[a] + ci) << 32) | b;
  } else if (kind == DTT_SPLIT_B) {
    for (uint32_t cj = 0; cj < t.cc[b]; ++cj) *out++ = ((uint64_t)a << 32) | (t.cf[b] + cj);
  }
}

// per visited pair: 2 M2L keys if accepted, 2 P2P keys if a pair of distinct leaves
__global__ void k_dtt_emit_count(DttTree t, const uint64_t* __restrict__ pairs, uint32_t n, uint32_t* __restrict__ mcnt,
                                 uint32_t* __restrict__ pcnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t pr = pairs[i];
  const uint32_t kind = dtt_kind(t, (uint32_t)(pr >> 32), (uint32_t)pr);
  mcnt[i] = kind == DTT_M2L ? 2u : 0u;
  pcnt[i] = kind == DTT_P2P ? 2u : 0u;
}

// M2L key (a << wm) | b; P2P key (a << wp) | (b + 1), the leaf itself being 0
__global__ void k_dtt_emit(DttTree t, const uint64_t* __restrict__ pairs, uint32_t n, const uint32_t* __restrict__ mpos,
                           const uint32_t* __restrict__ ppos, int wm, int wp, uint64_t* __restrict__ mkeys,
                           uint64_t* __restrict__ pkeys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t pr = pairs[i];
  const uint32_t a = (uint32_t)(pr >> 32), b = (uint32_t)pr;
  const uint32_t kind = dtt_kind(t, a, b);
  if (kind == DTT_M2L) {
    mkeys[mpos[i]] = ((uint64_t)a << wm) | b;
    mkeys[mpos[i] + 1] = ((uint64_t)b << wm) | a;
  } else if (kind == DTT_P2P) {
    pkeys[ppos[i]] = ((uint64_t)a << wp) | (b + 1u);
    pkeys[ppos[i] + 1] = ((uint64_t)b << wp) | (a + 1u);
  }
}

__global__ void k_dtt_self_keys(uint32_t nl, int wp, uint64_t* __restrict__ pkeys) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < nl) pkeys[a] = (uint64_t)a << wp;
}

// sorted keys -> CSR entries and row offsets (rows + 1 entries): off[u] is
// the first i whose row is u (key i starts every row after the previous
// key's, up to its own), off[rows] = n. (Counting with atomics instead
// serialises: neighbouring keys share their row.)
__global__ void k_dtt_split(const uint64_t* __restrict__ keys, uint32_t n, int w, bool self_first, uint32_t rows,
                            uint32_t* __restrict__ dat, uint32_t* __restrict__ off) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const uint32_t row = (uint32_t)(k >> w), col = (uint32_t)(k & ((1ull << w) - 1));
  dat[i] = self_first ? (col == 0 ? row : col - 1u) : col;
  const uint32_t u0 = i ? (uint32_t)(keys[i - 1] >> w) + 1u : 0u;
  for (uint32_t u = u0; u <= row; ++u) off[u] = i;
  if (i == n - 1)
    for (uint32_t u = row + 1; u <= rows; ++u) off[u] = n;
}

// Targets in decreasing M2L list length (stable), for the lane-per-pair M2L
// kernels: the long lists start first instead of forming the last blocks.
__global__ void k_order_keys(const uint32_t* __restrict__ off, uint32_t n, uint64_t* __restrict__ key,
                             uint32_t* __restrict__ val) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = 0xFFFFFFFFull - (off[i + 1] - off[i]);
  val[i] = i;
}

namespace {

int key_bits(uint64_t n) {  // bits for the values 0 .. n
  int w = 1;
  while (w < 32 && (1ull << w) <= n) ++w;
  return w;
}

int grow_pairs(DBuf& b, size_t keep, size_t need, hipStream_t s) {
  if (need * 8 <= b.bytes) return 0;
  DBuf nb;
  GX_CHECK(dbuf_reserve(nb, std::max(need, 2 * b.bytes / 8) * 8, s));
  if (keep) GX_HIP(hipMemcpyAsync(nb.p, b.p, keep * 8, hipMemcpyDeviceToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  dbuf_free(b);
  b = nb;
  return 0;
}

// sorts n keys of `bits` bits in ws.sort_ka and splits them into dat / off (rows + 1)
int keys_to_csr(galax_fmm* h, size_t n, int w, int bits, bool self_first, uint32_t rows, DBuf& off, DBuf& dat) {
  hipStream_t s = h->stream;
  Workspace& ws = h->ws;
  GX_CHECK(dbuf_reserve(ws.sort_kb, n * 8 + 8, s));
  GX_CHECK(dbuf_reserve(ws.sort_va, n * 4 + 4, s));
  GX_CHECK(dbuf_reserve(ws.sort_vb, n * 4 + 4, s));
  GX_CHECK(radix_sort_bits(h, ws.sort_ka.as<uint64_t>(), ws.sort_va.as<uint32_t>(), ws.sort_kb.as<uint64_t>(),
                           ws.sort_vb.as<uint32_t>(), n, bits));
  GX_CHECK(dbuf_reserve(off, ((size_t)rows + 1) * 4, s));
  GX_CHECK(dbuf_reserve(dat, n * 4 + 4, s));
  GX_HIP(hipMemsetAsync(off.p, 0, ((size_t)rows + 1) * 4, s));
  if (n) {
    k_dtt_split<<<nb256(n), 256, 0, s>>>(ws.sort_ka.as<uint64_t>(), (uint32_t)n, w, self_first, rows,
                                         dat.as<uint32_t>(), off.as<uint32_t>());
    GX_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace

// Dual tree traversal over the installed tree with the acceptance test
// R_A + R_B < theta |com_A - com_B| (radii from the given positions).
int build_lists_corrected(galax_fmm* h, const double* d_x, const double* d_y) {
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const size_t nn = t.n_nodes, nl = t.n_leaves;
  h->corrected_ready = false;
  h->c_sym.valid = false;
  if (nn == 0) {
    h->corrected_ready = true;
    return 0;
  }
  Workspace& ws = h->ws;
  GX_CHECK(dbuf_reserve(ws.a, nn * 8, s));
  {
    GX_CHECK(dbuf_reserve(ws.b, (nn + 1) * 4, s));
    uint32_t* ioff = ws.b.as<uint32_t>();
    k_radii_items<<<nb256(nn), 256, 0, s>>>(t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(), (uint32_t)nn,
                                            ioff);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemsetAsync(ioff + nn, 0, 4, s));
    uint32_t nitems = 0;
    GX_CHECK(scan_u32(h, ioff, ioff, nn + 1, false, false, &nitems));
    GX_HIP(hipMemsetAsync(ws.a.p, 0, nn * 8, s));
    k_radii<<<(nitems + 3) / 4, 256, 0, s>>>(d_x, d_y, t.body_start.as<uint32_t>(), t.body_end.as<uint32_t>(),
                                             t.com_x.as<double>(), t.com_y.as<double>(), (uint32_t)nn, ioff, nitems,
                                             ws.a.as<unsigned long long>());
    GX_HIP(hipGetLastError());
    k_radii_sqrt<<<nb256(nn), 256, 0, s>>>((uint32_t)nn, ws.a.as<double>());
    GX_HIP(hipGetLastError());
  }
  const DttTree dt{t.com_x.as<double>(), t.com_y.as<double>(), ws.a.as<double>(), t.child_first.as<uint32_t>(),
                   t.child_count.as<uint8_t>(), h->cfg.theta};
  // rounds: pairs [lo, hi) expand into [hi, hi + total)
  DBuf& P = h->c_dtt_pairs;
  GX_CHECK(grow_pairs(P, 0, std::max<size_t>(4 * nn, 1024), s));
  const uint64_t root_pair = ((uint64_t)(nn - 1) << 32) | (uint64_t)(nn - 1);
  GX_HIP(hipMemcpyAsync(P.p, &root_pair, 8, hipMemcpyHostToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  size_t lo = 0, hi = 1;
  while (lo < hi) {
    const size_t n = hi - lo;
    GX_CHECK(dbuf_reserve(ws.b, (n + 1) * 4, s));
    uint32_t* cnt = ws.b.as<uint32_t>();
    k_dtt_count<<<nb256(n), 256, 0, s>>>(dt, P.as<uint64_t>(), (uint32_t)lo, (uint32_t)hi, cnt);
    GX_HIP(hipGetLastError());
    // pair ids and the scan are 32-bit: a pair has at most 10 child pairs
    // (a self pair of a node with 4 children), so the exact 64-bit total is
    // needed only when hi + 10 n could reach 2^32 (it costs a wait per round)
    if (hi + 10ull * n >= (1ull << 32)) {
      uint64_t total64 = 0;
      GX_CHECK(sum_u32(h, cnt, n, &total64));
      if (hi + total64 >= (1ull << 32)) return set_error(GALAX_ERR_UNSUPPORTED, "dual tree traversal above 2^32 pairs");
    }
    uint32_t total = 0;
    GX_CHECK(scan_u32(h, cnt, cnt, n, false, false, &total));
    GX_CHECK(grow_pairs(P, hi, hi + total, s));
    if (total) {
      k_dtt_expand<<<nb256(n), 256, 0, s>>>(dt, P.as<uint64_t>(), (uint32_t)lo, (uint32_t)hi, cnt);
      GX_HIP(hipGetLastError());
    }
    lo = hi;
    hi += total;
  }
  const size_t nv = hi;
  // M2L and P2P keys of the visited pairs
  GX_CHECK(dbuf_reserve(ws.c, (nv + 1) * 4, s));
  GX_CHECK(dbuf_reserve(ws.d, (nv + 1) * 4, s));
  uint32_t *mpos = ws.c.as<uint32_t>(), *ppos = ws.d.as<uint32_t>();
  k_dtt_emit_count<<<nb256(nv), 256, 0, s>>>(dt, P.as<uint64_t>(), (uint32_t)nv, mpos, ppos);
  GX_HIP(hipGetLastError());
  uint64_t m64 = 0, p64 = 0;
  GX_CHECK(sum_u32(h, mpos, nv, &m64));
  GX_CHECK(sum_u32(h, ppos, nv, &p64));
  if (m64 >= 0xFFFFFFF0ull || p64 + nl >= 0xFFFFFFF0ull)
    return set_error(GALAX_ERR_UNSUPPORTED, "corrected lists above 2^32 entries");
  uint32_t mtot = 0, ptot = 0;
  GX_CHECK(scan_u32(h, mpos, mpos, nv, false, false, &mtot));
  GX_CHECK(scan_u32(h, ppos, ppos, nv, false, false, &ptot));
  const int wm = key_bits(nn), wp = key_bits(nl + 1);
  // M2L: keys in ws.sort_ka; P2P: keys in ws.e (self keys first), sorted after the M2L lists
  GX_CHECK(dbuf_reserve(ws.sort_ka, ((size_t)mtot + 1) * 8, s));
  GX_CHECK(dbuf_reserve(ws.e, ((size_t)ptot + nl + 1) * 8, s));
  uint64_t* pkeys = ws.e.as<uint64_t>();
  k_dtt_self_keys<<<nb256(nl), 256, 0, s>>>((uint32_t)nl, wp, pkeys);
  GX_HIP(hipGetLastError());
  k_dtt_emit<<<nb256(nv), 256, 0, s>>>(dt, P.as<uint64_t>(), (uint32_t)nv, mpos, ppos, wm, wp,
                                       ws.sort_ka.as<uint64_t>(), pkeys + nl);
  GX_HIP(hipGetLastError());
  GX_CHECK(keys_to_csr(h, mtot, wm, 2 * wm, false, (uint32_t)nn, h->c_m2l_off, h->c_m2l_dat));
  const size_t pn = (size_t)ptot + nl;
  GX_CHECK(dbuf_reserve(ws.sort_ka, pn * 8 + 8, s));
  GX_HIP(hipMemcpyAsync(ws.sort_ka.p, pkeys, pn * 8, hipMemcpyDeviceToDevice, s));
  GX_CHECK(keys_to_csr(h, pn, wp, 2 * wp, true, (uint32_t)nl, h->c_gat_off, h->c_gat_dat));
  GX_HIP(hipStreamSynchronize(s));
  {  // the M2L targets by decreasing list length
    GX_CHECK(dbuf_reserve(ws.sort_ka, nn * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_kb, nn * 8, s));
    GX_CHECK(dbuf_reserve(ws.sort_vb, nn * 4, s));
    GX_CHECK(dbuf_reserve(h->c_m2l_order, nn * 4, s));
    k_order_keys<<<nb256(nn), 256, 0, s>>>(h->c_m2l_off.as<uint32_t>(), (uint32_t)nn, ws.sort_ka.as<uint64_t>(),
                                           h->c_m2l_order.as<uint32_t>());
    GX_HIP(hipGetLastError());
    GX_CHECK(radix_sort_bits(h, ws.sort_ka.as<uint64_t>(), h->c_m2l_order.as<uint32_t>(), ws.sort_kb.as<uint64_t>(),
                             ws.sort_vb.as<uint32_t>(), nn, 32));
  }
  h->c_m2l_total = mtot;
  h->c_gat_total = pn;
  h->c_dtt_visited = nv;
  GX_HIP(hipStreamSynchronize(s));
  dbuf_free(P);  // the visited pairs (8 bytes each) are needed only while the lists are built
  GX_CHECK(count_interactions(h, h->c_gat_off.as<uint32_t>(), h->c_gat_dat.as<uint32_t>(), &h->c_p2p_interactions));
  GX_CHECK(build_p2p_items(h, h->c_gat_off.as<uint32_t>(), h->c_gat_dat.as<uint32_t>(), (uint32_t)t.n_leaves,
                           h->c_items));
  h->corrected_ready = true;
  return 0;
}

// ---------------------------------------------------------------- KDK ----

__global__ void k_kick(double* __restrict__ vx, double* __restrict__ vy, const double* __restrict__ ax,
                       const double* __restrict__ ay, size_t n, double half) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] += ax[i] * half;
  vy[i] += ay[i] * half;
}

__global__ void k_drift(double* __restrict__ x, double* __restrict__ y, const double* __restrict__ vx,
                        const double* __restrict__ vy, size_t n, double dt) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += vx[i] * dt;
  y[i] += vy[i] * dt;
}

// per-block partial sums of {KE, 1/2 m phi, heuristic PE, px, py, Lz}
__global__ void k_diag(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ vx,
                       const double* __restrict__ vy, const double* __restrict__ m, const double* __restrict__ ax,
                       const double* __restrict__ ay, const double* __restrict__ phi, size_t n,
                       double* __restrict__ part) {
  __shared__ double sh[6][256];
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    double mi = m[i];
    v[0] = 0.5 * mi * (vx[i] * vx[i] + vy[i] * vy[i]);
    v[1] = 0.5 * mi * (phi ? phi[i] : 0.0);
    double a = sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
    double r = fmax(sqrt(x[i] * x[i] + y[i] * y[i]), 1.0);
    v[2] = mi * a * r;
    v[3] = mi * vx[i];
    v[4] = mi * vy[i];
    v[5] = mi * (x[i] * vy[i] - y[i] * vx[i]);
  }
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) part[(size_t)blockIdx.x * 6 + q] = sh[q][0];
}

static int diagnostics(galax_fmm* h, uint64_t step, galax_diag_t* out) {
  size_t n = h->tree.n_bodies;
  hipStream_t s = h->stream;
  unsigned nb = nb256(n);
  GX_CHECK(dbuf_reserve(h->ws.b, (size_t)nb * 6 * 8, s));
  k_diag<<<nb, 256, 0, s>>>(h->x.as<double>(), h->y.as<double>(), h->vx.as<double>(), h->vy.as<double>(),
                            h->m.as<double>(), h->ax.as<double>(), h->ay.as<double>(), h->phi.as<double>(), n,
                            h->ws.b.as<double>());
  GX_HIP(hipGetLastError());
  std::vector<double> part((size_t)nb * 6);
  GX_HIP(hipMemcpyAsync(part.data(), h->ws.b.p, part.size() * 8, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (unsigned b = 0; b < nb; ++b)
    for (int q = 0; q < 6; ++q) acc[q] += part[(size_t)b * 6 + q];
  out->step = step;
  out->kinetic = acc[0];
  out->potential = acc[1];
  out->potential_heuristic = -0.5 * acc[2];
  out->momentum_x = acc[3];
  out->momentum_y = acc[4];
  out->angular_momentum = acc[5];
  return 0;
}

int kdk_device(galax_fmm* h, double dt, uint64_t n_steps, uint64_t every, galax_diag_t* diag) {
  size_t n = h->tree.n_bodies;
  hipStream_t s = h->stream;
  double half = 0.5 * dt;
  uint64_t nd = 0;
  for (uint64_t st = 0; st < n_steps; ++st) {
    k_kick<<<nb256(n), 256, 0, s>>>(h->vx.as<double>(), h->vy.as<double>(), h->ax.as<double>(), h->ay.as<double>(),
                                    n, half);
    k_drift<<<nb256(n), 256, 0, s>>>(h->x.as<double>(), h->y.as<double>(), h->vx.as<double>(), h->vy.as<double>(),
                                     n, dt);
    GX_HIP(hipGetLastError());
    GX_CHECK(evaluate(h, h->x.as<double>(), h->y.as<double>(), h->m.as<double>(), h->ax.as<double>(),
                      h->ay.as<double>(), h->phi.as<double>()));
    k_kick<<<nb256(n), 256, 0, s>>>(h->vx.as<double>(), h->vy.as<double>(), h->ax.as<double>(), h->ay.as<double>(),
                                    n, half);
    GX_HIP(hipGetLastError());
    if (diag && every > 0 && (st + 1) % every == 0) GX_CHECK(diagnostics(h, st + 1, &diag[nd++]));
  }
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace gx

// ====================================================================== ABI

namespace {

int check_cfg(const galax_fmm_config_t* c) {
  if (!c) return set_error(GALAX_ERR_INVALID, "config is NULL");
  if (c->p > (uint32_t)MAX_P) return set_error(GALAX_ERR_INVALID, "p = %u outside 0..%d", c->p, MAX_P);
  if (c->n_max == 0) return set_error(GALAX_ERR_INVALID, "n_max must be > 0");
  if (!(c->softening >= 0.0) || !std::isfinite(c->softening))
    return set_error(GALAX_ERR_INVALID, "softening must be finite and >= 0");
  if (c->tree_mode > GALAX_TREE_CONSTRAINED) return set_error(GALAX_ERR_INVALID, "unknown tree_mode");
  if (c->algo > GALAX_ALGO_CORRECTED) return set_error(GALAX_ERR_INVALID, "unknown algo");
  if (c->precision > GALAX_PREC_FP64_EXACT) return set_error(GALAX_ERR_INVALID, "unknown precision");
  if (c->algo == GALAX_ALGO_CORRECTED && !(c->theta > 0.0 && c->theta < 1.0))
    return set_error(GALAX_ERR_INVALID, "theta must be in (0, 1) for the corrected variant");
  return 0;
}

int reserve_bodies(galax_fmm* h, size_t n) {
  hipStream_t s = h->stream;
  DBuf* b[8] = {&h->x, &h->y, &h->m, &h->vx, &h->vy, &h->ax, &h->ay, &h->phi};
  for (DBuf* d : b) GX_CHECK(dbuf_reserve(*d, std::max<size_t>(n, 1) * 8, s));
  return 0;
}

int after_tree_installed(galax_fmm* h) {
  h->has_tree = true;
  h->keys_valid = false;
  h->mx_valid = h->mx_adj_valid = h->mx_rows32 = false;
  h->corrected_ready = false;
  return 0;
}

int up(hipStream_t s, DBuf& d, const double* src, size_t n) {
  if (!src || n == 0) return 0;
  GX_HIP(hipMemcpyAsync(d.p, src, n * 8, hipMemcpyHostToDevice, s));
  return 0;
}
int down(hipStream_t s, double* dst, const DBuf& d, size_t n) {
  if (!dst || n == 0) return 0;
  GX_HIP(hipMemcpyAsync(dst, d.p, n * 8, hipMemcpyDeviceToHost, s));
  return 0;
}

}  // namespace

extern "C" {

const char* galax_fmm_last_error(void) { return gx::last_error(); }

const char* galax_fmm_version(void) { return "galax_fmm 0.1 (gfx950)"; }

int galax_fmm_create(const galax_fmm_config_t* cfg, size_t max_n, galax_fmm_t** out) {
  if (!out) return set_error(GALAX_ERR_INVALID, "out is NULL");
  *out = nullptr;
  GX_CHECK(check_cfg(cfg));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(GALAX_ERR_INIT, "no HIP device");
  galax_fmm* h = new (std::nothrow) galax_fmm();
  if (!h) return set_error(GALAX_ERR_ALLOC, "out of host memory");
  h->cfg = *cfg;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return set_error(GALAX_ERR_INIT, "hipGetDevice failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    delete h;
    return set_error(GALAX_ERR_COMPILE, "device is %s; this library carries gfx950 code only", prop.gcnArchName);
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return set_error(GALAX_ERR_INIT, "hipStreamCreate failed");
  }
  h->max_n = max_n;
  {  // measurement switch: the FP32 forms' near field on the gather kernel
    const char* g = std::getenv("GALAX_NEAR_GATHER");
    h->near_gather = g && g[0] == '1';
  }
  int rc = reserve_bodies(h, max_n);
  if (rc) {
    galax_fmm_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

void galax_fmm_destroy(galax_fmm_t* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  DBuf* bufs[] = {&h->x, &h->y, &h->m, &h->vx, &h->vy, &h->ax, &h->ay, &h->phi, &h->mp, &h->loc, &h->mp32,
                  &h->loc32, &h->gtab, &h->key_of_pair, &h->perm, &h->hash, &h->tmp_u32, &h->flag,
                  &h->c_m2l_off, &h->c_m2l_dat, &h->c_gat_off, &h->c_gat_dat, &h->c_dtt_pairs, &h->c_m2l_order,
                  &h->key_dxy, &h->mx_t64, &h->mx_t64T, &h->mx_that, &h->mx_sig, &h->mx_zero, &h->mx_nexp,
                  &h->vj_lbar, &h->vj_mbar, &h->vj_leaf_of, &h->vj_fwd, &h->vj_out,
                  &h->mx_fwd.src, &h->mx_fwd.tile_off, &h->mx_fwd.step_key, &h->mx_fwd.step_first,
                  &h->mx_fwd.step_cnt, &h->mx_fwd.step_ids, &h->mx_fwd.item_tile, &h->mx_fwd.item_s0, &h->mx_fwd.item_s1,
                  &h->mx_fwd.item_slot, &h->mx_fwd.multi_tile, &h->mx_fwd.multi_slot, &h->mx_fwd.multi_cnt,
                  &h->mx_fwd.part, &h->mx_fwd.lr_first, &h->mx_fwd.lr_len, &h->mx_fwd.lr_order, &h->mx_fwd.hub_first, &h->mx_fwd.hub_len, &h->mx_fwd.hub_coff, &h->mx_fwd.hub_part, &h->mx_fwd.virt, &h->mx_fwd.hubs, &h->mx_fwd.lr_key, &h->mx_fwd.step_sh, &h->mx_fwd.lr_tab, &h->mx_fwd.lr_psh, &h->mx_adj.src, &h->mx_adj.tile_off, &h->mx_adj.step_key,
                  &h->mx_adj.step_first, &h->mx_adj.step_cnt, &h->mx_adj.step_ids, &h->mx_adj.item_tile, &h->mx_adj.item_s0,
                  &h->mx_adj.item_s1, &h->mx_adj.item_slot, &h->mx_adj.multi_tile, &h->mx_adj.multi_slot,
                  &h->mx_adj.multi_cnt, &h->mx_adj.part, &h->mx_adj.lr_first, &h->mx_adj.lr_len, &h->mx_adj.lr_order, &h->mx_adj.hub_first, &h->mx_adj.hub_len, &h->mx_adj.hub_coff, &h->mx_adj.hub_part, &h->mx_adj.virt, &h->mx_adj.hubs, &h->mx_adj.lr_key, &h->mx_adj.step_sh, &h->mx_adj.lr_tab, &h->mx_adj.lr_psh};
  for (DBuf* b : bufs) dbuf_free(*b);
  DevTree& t = h->tree;
  DBuf* tb[] = {&t.center_x, &t.center_y, &t.half_width, &t.mass, &t.com_x, &t.com_y, &t.body_start, &t.body_end,
                &t.child_first, &t.child_count, &t.parent, &t.level, &t.keys, &t.m2l_off, &t.m2l_dat, &t.p2p_off,
                &t.p2p_dat, &t.gat_off, &t.gat_dat};
  for (DBuf* b : tb) dbuf_free(*b);
  for (P2PItems* it : {&t.items, &h->c_items}) {
    DBuf* ib[] = {&it->tgt, &it->g0, &it->g1, &it->slot, &it->part_off, &it->multi};
    for (DBuf* b : ib) dbuf_free(*b);
  }
  sym_p2p_free(t.sym);
  sym_p2p_free(h->c_sym);
  dbuf_free(h->p2p_part);
  Workspace& w = h->ws;
  DBuf* wb[] = {&w.a, &w.b, &w.c, &w.d, &w.e, &w.f, &w.g, &w.h, &w.sort_ka, &w.sort_kb, &w.sort_va, &w.sort_vb,
                &w.sort_hist};
  for (DBuf* b : wb) dbuf_free(*b);
  for (DBuf& b : w.scan) dbuf_free(b);
  for (hipEvent_t& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t& e : h->bev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static int build_common(galax_fmm_t* h, size_t n) {
  h->has_tree = false;  // a build that fails part-way leaves no usable tree
  for (hipEvent_t& e : h->bev)
    if (!e) GX_HIP(hipEventCreate(&e));
  h->build_timed = true;
  const bool constrained = h->cfg.tree_mode == GALAX_TREE_CONSTRAINED;
  GX_CHECK(build_mark(h, 0));
  if (constrained || n == 0) {  // no sort or permutation stage
    GX_CHECK(build_mark(h, 1));
    GX_CHECK(build_mark(h, 2));
  }
  int rc = constrained ? build_tree_constrained(h, n) : build_tree_plain(h, n);
  if (rc) return rc;
  GX_CHECK(build_mark(h, 3));
  GX_CHECK(build_lists_reference(h));
  if (h->tree.n_nodes == 0) GX_CHECK(build_mark(h, 4));
  GX_CHECK(build_mark(h, 5));
  GX_HIP(hipStreamSynchronize(h->stream));
  return after_tree_installed(h);
}

int galax_fmm_last_build_timings(galax_fmm_t* h, galax_build_timings_t* out) {
  if (!h || !out) return set_error(GALAX_ERR_INVALID, "NULL argument");
  *out = galax_build_timings_t{};
  if (!h->build_timed || h->tree.n_bodies == 0) return 0;
  float ms[5], tot = 0.f;
  for (int i = 0; i < 5; ++i) GX_HIP(hipEventElapsedTime(&ms[i], h->bev[i], h->bev[i + 1]));
  GX_HIP(hipEventElapsedTime(&tot, h->bev[0], h->bev[5]));
  out->sort = ms[0];
  out->permute = ms[1];
  out->tree = ms[2];
  out->lists = ms[3];
  out->gather = ms[4];
  out->total = tot;
  return 0;
}

int galax_fmm_build_tree(galax_fmm_t* h, galax_bodies_t* b, uint64_t* perm_out) {
  if (!h || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  if (b->n > h->max_n) return set_error(GALAX_ERR_INVALID, "n = %zu exceeds max_n = %zu", b->n, h->max_n);
  if (b->n > 0 && (!b->x || !b->y || !b->m)) return set_error(GALAX_ERR_INVALID, "x, y and m are required");
  if (b->n >= 0xFFFFFFF0ull) return set_error(GALAX_ERR_INVALID, "n too large");
  hipStream_t s = h->stream;
  size_t n = b->n;
  GX_HIP(hipSetDevice(h->device));
  GX_CHECK(up(s, h->x, b->x, n));
  GX_CHECK(up(s, h->y, b->y, n));
  GX_CHECK(up(s, h->m, b->m, n));
  // arrays the caller did not pass are not permuted (they stay zero on device)
  DBuf* opt[4] = {&h->vx, &h->vy, &h->ax, &h->ay};
  double* src[4] = {b->vx, b->vy, b->ax, b->ay};
  for (int k = 0; k < 4; ++k) {
    if (src[k])
      GX_CHECK(up(s, *opt[k], src[k], n));
    else if (n)
      GX_HIP(hipMemsetAsync(opt[k]->p, 0, n * 8, s));
  }
  GX_CHECK(build_common(h, n));
  GX_CHECK(down(s, b->x, h->x, n));
  GX_CHECK(down(s, b->y, h->y, n));
  GX_CHECK(down(s, b->m, h->m, n));
  for (int k = 0; k < 4; ++k) GX_CHECK(down(s, src[k], *opt[k], n));
  if (perm_out && n) {
    std::vector<uint32_t> pv(n);
    if (h->cfg.tree_mode == GALAX_TREE_CONSTRAINED) {
      for (size_t i = 0; i < n; ++i) pv[i] = (uint32_t)i;
    } else {
      GX_HIP(hipMemcpyAsync(pv.data(), h->perm.p, n * 4, hipMemcpyDeviceToHost, s));
    }
    GX_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) perm_out[i] = pv[i];
  }
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

int galax_fmm_build_tree_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_vx, double* d_vy, double* d_m,
                                double* d_ax, double* d_ay, size_t n) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (n > h->max_n) return set_error(GALAX_ERR_INVALID, "n = %zu exceeds max_n = %zu", n, h->max_n);
  if (n > 0 && (!d_x || !d_y || !d_m)) return set_error(GALAX_ERR_INVALID, "x, y and m are required");
  hipStream_t s = h->stream;
  double* arr[7] = {d_x, d_y, d_vx, d_vy, d_m, d_ax, d_ay};
  DBuf* own[7] = {&h->x, &h->y, &h->vx, &h->vy, &h->m, &h->ax, &h->ay};
  for (int k = 0; k < 7; ++k) {
    if (arr[k] && n)
      GX_HIP(hipMemcpyAsync(own[k]->p, arr[k], n * 8, hipMemcpyDeviceToDevice, s));
    else if (n)
      GX_HIP(hipMemsetAsync(own[k]->p, 0, n * 8, s));
  }
  GX_CHECK(build_common(h, n));
  for (int k = 0; k < 7; ++k)
    if (arr[k] && n) GX_HIP(hipMemcpyAsync(arr[k], own[k]->p, n * 8, hipMemcpyDeviceToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

int galax_fmm_upload_tree(galax_fmm_t* h, const galax_tree_view_t* tree, const galax_lists_view_t* m2l,
                          const galax_lists_view_t* p2p) {
  if (!h || !tree || !m2l || !p2p) return set_error(GALAX_ERR_INVALID, "NULL argument");
  h->build_timed = false;  // an uploaded tree has no device build stages
  // install_tree_host checks everything on the host before it changes any
  // state; once it starts uploading, the previous tree is gone
  GX_CHECK(install_tree_host(h, tree, m2l, p2p));
  return after_tree_installed(h);
}

int galax_fmm_set_softening(galax_fmm_t* h, double softening) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (!(softening >= 0.0) || !std::isfinite(softening))
    return set_error(GALAX_ERR_INVALID, "softening must be finite and >= 0");
  h->cfg.softening = softening;
  h->keys_valid = false;  // derivative tables depend on eps
  h->mx_valid = h->mx_adj_valid = h->mx_rows32 = false;
  return 0;
}

int galax_fmm_tree_info(galax_fmm_t* h, size_t* n_nodes, size_t* n_leaves, uint32_t* n_levels, size_t* m2l_total,
                        size_t* p2p_total) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (n_nodes) *n_nodes = h->tree.n_nodes;
  if (n_leaves) *n_leaves = h->tree.n_leaves;
  if (n_levels) *n_levels = h->tree.n_levels;
  if (m2l_total) {
    GX_CHECK(ensure_full_m2l_lists(h));
    *m2l_total = h->tree.m2l_total;
  }
  if (p2p_total) *p2p_total = h->tree.p2p_total;
  return 0;
}

int galax_fmm_export_corrected_lists(galax_fmm_t* h, galax_lists_view_t* m2l, galax_lists_view_t* p2p) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (h->cfg.algo != GALAX_ALGO_CORRECTED || !h->corrected_ready)
    return set_error(GALAX_ERR_INVALID, "no corrected-variant lists: evaluate with GALAX_ALGO_CORRECTED first");
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  const size_t nn = t.n_nodes;
  // rows: nn for the M2L lists, the leaves [0, n_leaves) for the P2P lists
  auto ex = [&](galax_lists_view_t* lv, DBuf& off, DBuf& dat, size_t rows, size_t total) -> int {
    if (!lv) return 0;
    if (lv->n_nodes != nn) return set_error(GALAX_ERR_INVALID, "list view holds %zu nodes, tree has %zu", lv->n_nodes, nn);
    if (!lv->offsets) return set_error(GALAX_ERR_INVALID, "NULL offsets");
    std::vector<uint32_t> o(rows + 1, 0);
    if (nn) GX_HIP(hipMemcpyAsync(o.data(), off.p, (rows + 1) * 4, hipMemcpyDeviceToHost, s));
    if (lv->data && total) GX_HIP(hipMemcpyAsync(lv->data, dat.p, total * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i <= nn; ++i) lv->offsets[i] = i <= rows ? o[i] : o[rows];
    return 0;
  };
  GX_CHECK(ex(m2l, h->c_m2l_off, h->c_m2l_dat, nn, h->c_m2l_total));
  GX_CHECK(ex(p2p, h->c_gat_off, h->c_gat_dat, t.n_leaves, h->c_gat_total));
  return 0;
}

int galax_fmm_export_expansions(galax_fmm_t* h, double* multipoles, double* locals, size_t n_nodes) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  const size_t nn = h->tree.n_nodes, S = (size_t)(h->cfg.p + 1) * (h->cfg.p + 2) / 2;
  if (n_nodes != nn) return set_error(GALAX_ERR_INVALID, "arrays hold %zu nodes, tree has %zu", n_nodes, nn);
  if (nn == 0) return 0;
  if (!h->mp.p || h->mp.bytes < nn * S * 8 || !h->loc.p || h->loc.bytes < nn * S * 8)
    return set_error(GALAX_ERR_INVALID, "no expansions yet: evaluate first");
  hipStream_t s = h->stream;
  if (multipoles) GX_HIP(hipMemcpyAsync(multipoles, h->mp.p, nn * S * 8, hipMemcpyDeviceToHost, s));
  if (locals) GX_HIP(hipMemcpyAsync(locals, h->loc.p, nn * S * 8, hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

int galax_fmm_export_tree(galax_fmm_t* h, galax_tree_view_t* tv, galax_lists_view_t* m2l, galax_lists_view_t* p2p) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  DevTree& t = h->tree;
  hipStream_t s = h->stream;
  size_t nn = t.n_nodes;
  if (tv) {
    if (tv->n_nodes != nn) return set_error(GALAX_ERR_INVALID, "tree view holds %zu nodes, tree has %zu", tv->n_nodes, nn);
    std::vector<uint32_t> bs(nn), be(nn), cf(nn), par(nn);
    std::vector<uint8_t> cc(nn);
    GX_HIP(hipMemcpyAsync(tv->center_x, t.center_x.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(tv->center_y, t.center_y.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(tv->half_width, t.half_width.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(tv->mass, t.mass.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(tv->com_x, t.com_x.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(tv->com_y, t.com_y.p, nn * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(bs.data(), t.body_start.p, nn * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(be.data(), t.body_end.p, nn * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(cf.data(), t.child_first.p, nn * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(par.data(), t.parent.p, nn * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(cc.data(), t.child_count.p, nn, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < nn; ++i) {
      tv->body_start[i] = bs[i];
      tv->body_end[i] = be[i];
      for (int k = 0; k < 4; ++k) tv->children[i * 4 + k] = (k < cc[i]) ? (int32_t)(cf[i] + k) : -1;
      tv->parent[i] = par[i] == NONE ? -1 : (int32_t)par[i];
      tv->is_leaf[i] = cc[i] == 0 ? 1 : 0;
    }
  }
  auto ex = [&](galax_lists_view_t* lv, DBuf& off, DBuf& dat, size_t total) -> int {
    if (!lv) return 0;
    if (lv->n_nodes != nn) return set_error(GALAX_ERR_INVALID, "list view size mismatch");
    std::vector<uint32_t> o(nn + 1);
    GX_HIP(hipMemcpyAsync(o.data(), off.p, (nn + 1) * 4, hipMemcpyDeviceToHost, s));
    if (total) GX_HIP(hipMemcpyAsync(lv->data, dat.p, total * 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i <= nn; ++i) lv->offsets[i] = o[i];
    return 0;
  };
  if (nn) {
    if (m2l) GX_CHECK(ensure_full_m2l_lists(h));
    GX_CHECK(ex(m2l, t.m2l_off, t.m2l_dat, t.m2l_total));
    GX_CHECK(ex(p2p, t.p2p_off, t.p2p_dat, t.p2p_total));
  }
  return 0;
}

static int prepare_eval(galax_fmm_t* h, const double* d_x, const double* d_y) {
  if (!h->has_tree) return set_error(GALAX_ERR_INVALID, "no tree: call galax_fmm_build_tree or galax_fmm_upload_tree");
  if (h->cfg.algo == GALAX_ALGO_CORRECTED && !h->corrected_ready) GX_CHECK(build_lists_corrected(h, d_x, d_y));
  return 0;
}

int galax_fmm_evaluate(galax_fmm_t* h, galax_bodies_t* b) {
  if (!h || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  size_t n = h->tree.n_bodies;
  if (b->n != n) return set_error(GALAX_ERR_INVALID, "bodies.n = %zu but the tree holds %zu bodies", b->n, n);
  if (n && (!b->x || !b->y || !b->m || !b->ax || !b->ay))
    return set_error(GALAX_ERR_INVALID, "x, y, m, ax, ay are required");
  hipStream_t s = h->stream;
  GX_CHECK(up(s, h->x, b->x, n));
  GX_CHECK(up(s, h->y, b->y, n));
  GX_CHECK(up(s, h->m, b->m, n));
  GX_CHECK(prepare_eval(h, h->x.as<double>(), h->y.as<double>()));
  GX_CHECK(evaluate(h, h->x.as<double>(), h->y.as<double>(), h->m.as<double>(), h->ax.as<double>(),
                    h->ay.as<double>(), b->phi ? h->phi.as<double>() : nullptr));
  uint32_t changed = 0;
  GX_HIP(hipMemcpyAsync(&changed, h->flag.as<uint32_t>() + 4 + h->san_cur, 4, hipMemcpyDeviceToHost, s));
  GX_CHECK(down(s, b->ax, h->ax, n));
  GX_CHECK(down(s, b->ay, h->ay, n));
  if (b->phi) GX_CHECK(down(s, b->phi, h->phi, n));
  GX_HIP(hipStreamSynchronize(s));
  if (changed) {  // sanitize mutates the caller's bodies, as the reference does
    GX_CHECK(down(s, b->x, h->x, n));
    GX_CHECK(down(s, b->y, h->y, n));
    GX_CHECK(down(s, b->m, h->m, n));
    GX_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

int galax_fmm_crosscheck(galax_fmm_t* h, galax_bodies_t* b, double threshold, double* max_rel, double* mean_rel,
                         size_t* count) {
  if (!h || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  if (std::isnan(threshold)) return set_error(GALAX_ERR_INVALID, "threshold is NaN");
  size_t n = h->tree.n_bodies;
  if (b->n != n) return set_error(GALAX_ERR_INVALID, "bodies.n = %zu but the tree holds %zu bodies", b->n, n);
  if (n && (!b->x || !b->y || !b->m || !b->ax || !b->ay))
    return set_error(GALAX_ERR_INVALID, "x, y, m, ax, ay are required");
  hipStream_t s = h->stream;
  GX_CHECK(up(s, h->x, b->x, n));
  GX_CHECK(up(s, h->y, b->y, n));
  GX_CHECK(up(s, h->m, b->m, n));
  GX_CHECK(prepare_eval(h, h->x.as<double>(), h->y.as<double>()));
  // the context's form ...
  GX_CHECK(evaluate(h, h->x.as<double>(), h->y.as<double>(), h->m.as<double>(), h->ax.as<double>(),
                    h->ay.as<double>(), nullptr));
  std::vector<double> gax(n), gay(n);
  GX_CHECK(down(s, gax.data(), h->ax, n));
  GX_CHECK(down(s, gay.data(), h->ay, n));
  GX_HIP(hipStreamSynchronize(s));
  // ... then galax-core's arithmetic on the same tree (FP64_EXACT uses no
  // matrix-core operators, so the context's cached ones stay valid)
  const uint32_t prec = h->cfg.precision;
  h->cfg.precision = GALAX_PREC_FP64_EXACT;
  const int rc = evaluate(h, h->x.as<double>(), h->y.as<double>(), h->m.as<double>(), h->ax.as<double>(),
                          h->ay.as<double>(), nullptr);
  h->cfg.precision = prec;
  GX_CHECK(rc);
  GX_CHECK(down(s, b->ax, h->ax, n));
  GX_CHECK(down(s, b->ay, h->ay, n));
  GX_HIP(hipStreamSynchronize(s));
  // GpuContext::crosscheck's metric (galax-gpu/src/lib.rs:646-658)
  double mx = 0.0, sum = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double cm = std::sqrt(b->ax[i] * b->ax[i] + b->ay[i] * b->ay[i]);
    const double gm = std::sqrt(gax[i] * gax[i] + gay[i] * gay[i]);
    const double den = std::fmax(std::fmax(cm, gm), 1e-30);
    const double dx = gax[i] - b->ax[i], dy = gay[i] - b->ay[i];
    const double rel = std::sqrt(dx * dx + dy * dy) / den;
    mx = std::fmax(mx, rel);
    sum += rel;
  }
  if (max_rel) *max_rel = mx;
  if (mean_rel) *mean_rel = n ? sum / (double)n : 0.0;
  if (count) *count = n;
  if (mx > threshold)
    return set_error(GALAX_ERR_THRESHOLD, "crosscheck max relative error %.2e exceeds threshold %.2e", mx, threshold);
  return 0;
}

int galax_fmm_evaluate_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_m, double* d_ax, double* d_ay,
                              double* d_phi) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (h->tree.n_bodies && (!d_x || !d_y || !d_m || !d_ax || !d_ay))
    return set_error(GALAX_ERR_INVALID, "x, y, m, ax, ay are required");
  GX_CHECK(prepare_eval(h, d_x, d_y));
  return evaluate(h, d_x, d_y, d_m, d_ax, d_ay, d_phi);
}

int galax_fmm_last_timings(galax_fmm_t* h, galax_timings_t* out) {
  if (!h || !out) return set_error(GALAX_ERR_INVALID, "NULL argument");
  *out = h->timings;
  return 0;
}

int galax_fmm_set_timing(galax_fmm_t* h, int enable) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  h->timing = enable != 0;
  return 0;
}

int galax_fmm_kdk(galax_fmm_t* h, galax_bodies_t* b, double dt, uint64_t n_steps, uint64_t every,
                  galax_diag_t* diag) {
  if (!h || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  size_t n = h->tree.n_bodies;
  if (b->n != n) return set_error(GALAX_ERR_INVALID, "bodies.n = %zu but the tree holds %zu bodies", b->n, n);
  if (n && (!b->x || !b->y || !b->vx || !b->vy || !b->m || !b->ax || !b->ay))
    return set_error(GALAX_ERR_INVALID, "x, y, vx, vy, m, ax, ay are required");
  hipStream_t s = h->stream;
  GX_CHECK(up(s, h->x, b->x, n));
  GX_CHECK(up(s, h->y, b->y, n));
  GX_CHECK(up(s, h->vx, b->vx, n));
  GX_CHECK(up(s, h->vy, b->vy, n));
  GX_CHECK(up(s, h->m, b->m, n));
  GX_CHECK(up(s, h->ax, b->ax, n));
  GX_CHECK(up(s, h->ay, b->ay, n));
  GX_CHECK(prepare_eval(h, h->x.as<double>(), h->y.as<double>()));
  GX_CHECK(kdk_device(h, dt, n_steps, every, diag));
  double* outs[7] = {b->x, b->y, b->vx, b->vy, b->m, b->ax, b->ay};
  DBuf* own[7] = {&h->x, &h->y, &h->vx, &h->vy, &h->m, &h->ax, &h->ay};
  for (int k = 0; k < 7; ++k) GX_CHECK(down(s, outs[k], *own[k], n));
  if (b->phi) GX_CHECK(down(s, b->phi, h->phi, n));
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

int galax_fmm_vjp(galax_fmm_t* h, const galax_bodies_t* b, const double* axb, const double* ayb, double* xb,
                  double* yb, double* mb) {
  if (!h || !b) return set_error(GALAX_ERR_INVALID, "NULL argument");
  size_t n = h->tree.n_bodies;
  if (b->n != n) return set_error(GALAX_ERR_INVALID, "bodies.n = %zu but the tree holds %zu bodies", b->n, n);
  if (n && (!b->x || !b->y || !b->m || !axb || !ayb || !xb || !yb || !mb))
    return set_error(GALAX_ERR_INVALID, "x, y, m, cotangents and outputs are required");
  hipStream_t s = h->stream;
  GX_CHECK(up(s, h->x, b->x, n));
  GX_CHECK(up(s, h->y, b->y, n));
  GX_CHECK(up(s, h->m, b->m, n));
  GX_CHECK(up(s, h->ax, axb, n));  // the body buffers hold the cotangents and adjoints
  GX_CHECK(up(s, h->ay, ayb, n));
  // a buffer of its own: the passes run inside reuse the workspace
  GX_CHECK(dbuf_reserve(h->vj_out, std::max<size_t>(n, 1) * 24, s));
  double* out = h->vj_out.as<double>();
  GX_CHECK(prepare_eval(h, h->x.as<double>(), h->y.as<double>()));
  GX_CHECK(gx::vjp(h, h->x.as<double>(), h->y.as<double>(), h->m.as<double>(), h->ax.as<double>(),
                   h->ay.as<double>(), out, out + n, out + 2 * n));
  if (n) {
    GX_HIP(hipMemcpyAsync(xb, out, n * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(yb, out + n, n * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(mb, out + 2 * n, n * 8, hipMemcpyDeviceToHost, s));
  }
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

int galax_fmm_vjp_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_m, const double* d_axb,
                         const double* d_ayb, double* d_xb, double* d_yb, double* d_mb) {
  if (!h) return set_error(GALAX_ERR_INVALID, "NULL handle");
  if (h->tree.n_bodies && (!d_x || !d_y || !d_m || !d_axb || !d_ayb || !d_xb || !d_yb || !d_mb))
    return set_error(GALAX_ERR_INVALID, "NULL device array");
  GX_CHECK(prepare_eval(h, d_x, d_y));
  return gx::vjp(h, d_x, d_y, d_m, d_axb, d_ayb, d_xb, d_yb, d_mb);
}

int galax_fmm_validate(galax_fmm_t* h, galax_bodies_t* b, const uint64_t* samples, size_t ns, double* max_rel,
                       double* mean_rel, double* rms_vec) {
  if (!h || !b || (!samples && ns)) return set_error(GALAX_ERR_INVALID, "NULL argument");
  size_t n = b->n;
  std::vector<double> ax(n), ay(n), ex(ns), ey(ns);
  galax_bodies_t c = *b;
  std::vector<double> x(b->x, b->x + n), y(b->y, b->y + n), m(b->m, b->m + n);
  c.x = x.data();
  c.y = y.data();
  c.m = m.data();
  c.ax = ax.data();
  c.ay = ay.data();
  c.phi = nullptr;
  GX_CHECK(galax_fmm_evaluate(h, &c));
  GX_CHECK(galax_direct_sum(x.data(), y.data(), m.data(), n, h->cfg.softening, samples, ns, ex.data(), ey.data(),
                            nullptr));
  double mx = 0.0, sum = 0.0, sq = 0.0;
  for (size_t k = 0; k < ns; ++k) {
    size_t i = samples[k];
    double fa = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
    double ea = std::sqrt(ex[k] * ex[k] + ey[k] * ey[k]);
    double den = std::max(std::fabs(ea), 1e-20);
    double r = std::fabs(fa - ea) / den;  // galax-core/src/lib.rs:1554-1557
    mx = std::max(mx, r);
    sum += r;
    double vx = ax[i] - ex[k], vy = ay[i] - ey[k];
    sq += (vx * vx + vy * vy) / (den * den);
  }
  if (max_rel) *max_rel = mx;
  if (mean_rel) *mean_rel = ns ? sum / (double)ns : 0.0;
  if (rms_vec) *rms_vec = ns ? std::sqrt(sq / (double)ns) : 0.0;
  return 0;
}

}  // extern "C"
