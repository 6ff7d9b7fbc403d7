// common.h — internal definitions shared by the HIP translation units of
// libgalax_fmm.so (context, error plumbing, device helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/galax_fmm.h"

namespace gx {

constexpr int MAX_P = 16;          // plans exist for p <= 16 (galax-core/src/lib.rs:1163-1164)
constexpr uint32_t LIST_CAP = 64;  // InteractionLists data: [u32; 64] (galax-core/src/lib.rs:942)
constexpr uint32_t MORTON_BITS = 21;
constexpr uint32_t NONE = 0xFFFFFFFFu;

inline __host__ __device__ constexpr int num_moments(int p) { return (p + 1) * (p + 2) / 2; }
inline __host__ __device__ constexpr int moment_index(int i, int j) { return (i + j) * (i + j + 1) / 2 + j; }
// Derivative table index of G^(k,l) in a (2p+1)^2 buffer (galax-core/src/lib.rs:1033-1036).
inline __host__ __device__ constexpr int deriv_index(int k, int l, int p) { return (k + l) * (2 * p + 1) + l; }

// Thread-local error text; set_error returns the code for `return set_error(...)`.
int set_error(int code, const char* fmt, ...);
const char* last_error();

#define GX_HIP(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return ::gx::set_error(GALAX_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                                 \
  } while (0)

#define GX_CHECK(expr)        \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)

// Grow-only device buffer owned by a handle.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};
int dbuf_reserve(DBuf& b, size_t bytes, hipStream_t s);
void dbuf_free(DBuf& b);
// a temporary buffer freed when the scope ends (early error returns included)
struct ScopedDBuf : DBuf {
  ScopedDBuf() = default;
  ScopedDBuf(const ScopedDBuf&) = delete;
  ScopedDBuf& operator=(const ScopedDBuf&) = delete;
  ~ScopedDBuf() { dbuf_free(*this); }
};

// P2P work items over gather lists: item i covers target leaf tgt[i] and the
// gather entries [g0[i], g1[i]). A target with several items writes partial
// near fields to slots slot[i] and a combine pass sums them in item order;
// a target with one item (slot NONE) writes its result directly.
struct P2PItems {
  DBuf tgt, g0, g1, slot, part_off, multi;
  size_t n_items = 0, n_multi = 0, n_part = 0;
};

// Symmetric near field (the FP32 forms, leaves of at most SYM_LEAF bodies).
// Each unordered leaf pair {t, b} of the (symmetric) gather lists is
// evaluated once. The reference fixes the pair set (b in t's list, t < b,
// galax-core/src/lib.rs:1415-1437), not which side computes a pair; here the
// leaf with the longer gather list owns it (ties: the lower id), so the few
// leaves listed by thousands of others sum their pairs in registers. A work
// item covers one owner leaf t and at most SYM_ITEM_PARTNERS of its owned
// partners own_dat[u0 .. u1): forces on t go to the item's own slot, forces
// on a partner's bodies to the pair's slot u (the position in own_dat). Each
// leaf's results are then its own slots (item order) and the slots of the
// pairs it does not own (ascending owner), listed in refs (own slots tagged
// with SYM_OWN_TAG). They are summed in chunks of SYM_CHUNK refs, in parallel,
// and leaves with several chunks add their chunk sums in chunk order: a
// fixed order, so the result does not depend on scheduling.
constexpr uint32_t SYM_LEAF = 32;
constexpr uint32_t SYM_ITEM_PARTNERS = 16;
constexpr uint32_t SYM_CHUNK = 32;
constexpr uint32_t SYM_OWN_TAG = 0x80000000u;
struct SymP2P {
  bool valid = false;
  size_t n_items = 0, n_pairs = 0, n_leaves = 0, n_chunks = 0, n_heavy = 0, n_cpart = 0;
  DBuf own_off, own_dat;                   // per owner leaf: owned partners (n_leaves + 1 / n_pairs)
  DBuf itgt, iu0, iu1, iown;               // per item, in execution order
  DBuf refs;                               // per leaf (ref_off): own slots then foreign pair slots
  DBuf ch_leaf, ch_r0, ch_r1, ch_dst;      // per chunk of refs; ch_dst NONE = the leaf's final sum
  DBuf heavy_leaf, heavy_p0, heavy_n;      // leaves with several chunks: first chunk sum, count
  DBuf rel;                                // float4 per body: x - com, y - com, m, 0 (leaf COM)
  DBuf own2, ownp, pair2, pairp, cpart2, cpartp;  // partial sums
  // Leaves of up to 2 SYM_LEAF bodies (halves): the structure above is built
  // over halves of at most SYM_LEAF bodies (the first SYM_LEAF bodies, then
  // the rest), each with its body range, its leaf's COM, and its leaf (whose
  // local expansion and parent give its far field).
  bool halves = false;
  DBuf vbs, vbe, vcx, vcy, vleaf;
  DBuf ab2;  // the VJP's cotangents per body in FP32 (its near-field adjoint reuses this structure)
};
int build_sym_p2p(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t n_leaves, SymP2P& sp);
// k_sym_chunk_sum / k_sym_heavy_sum over the structure's leaves (or halves)
// bs, be: each leaf's slots summed into (o0, o1) and, if o2, the third value.
int sym_combine(const SymP2P& sp, const uint32_t* bs, const uint32_t* be, double* o0, double* o1, double* o2,
                hipStream_t s);
// The symmetric near field of the tree's leaves (n_max <= 2 SYM_LEAF): over the
// leaves themselves, or over their halves when n_max > SYM_LEAF.
int build_sym_near(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t n_leaves, SymP2P& sp);
void sym_p2p_free(SymP2P& sp);

// M2L work in tiles of 16 rows (targets, or sources for the adjoint). The live
// pairs are grouped by (row, key) runs; a tile's steps are the distinct keys of
// its groups. step_ids lists, per step, each of its 16 columns' sources
// ([step][q][column], at most MX_GROUP): the other ends of the pairs (node ids),
// or for a run longer than MX_GROUP one pre-summed row (id n_nodes + its index
// in lr_first/lr_len, summed into `virt` each evaluation). Work items cover at
// most MX_ITEM_STEPS steps of one tile; tiles with several items write
// partial results to slots that a combine pass sums in item order.
struct M2LTiles {
  DBuf src, tile_off, step_key, step_first, step_cnt, step_ids;
  DBuf lr_first, lr_len, virt, hubs;  // long runs: source range; pre-summed rows; runs taking a block
  DBuf lr_order;                      // long runs by decreasing length (the pre-sum kernel's order)
  DBuf hub_first, hub_len, hub_coff, hub_part;  // hub runs in chunks: rows and partial sums per chunk
  size_t nhub_chunks = 0;
  DBuf item_tile, item_s0, item_s1, item_slot;  // per work item
  DBuf multi_tile, multi_slot, multi_cnt;       // per tile with several items
  DBuf part;                                    // partial results, 16 x S doubles per slot
  // mixed forms: lr_key: key of each long run (its pre-summed FP32 row, at the
  // key's scale, follows the node rows in the handle's mp32); step_sh: per
  // (step, column) the sources' row-scale shifts as int8 (packed)
  DBuf lr_key, step_sh;
  // mixed forms: per long run in the pre-sum kernel's order (first pair, length,
  // pre-summed row), and per pair of a long run its row's shift to the key's scale
  DBuf lr_tab, lr_psh;
  uint32_t group = 4;  // longest run summed inside the product kernel (1, 2 or 4); longer runs are pre-summed
  size_t n_rows = 0, ntiles = 0, ngroups = 0, nsteps = 0, nitems = 0, nmulti = 0, nslots = 0, nlong = 0, nhub = 0;
};

// Device-side tree in structure-of-arrays form. Node ids are level-ordered
// (leaves first, root last) and each node's children are consecutive ids
// [child_first, child_first + child_count), as Tree::build creates them.
struct DevTree {
  size_t n_nodes = 0, n_leaves = 0, n_bodies = 0;
  uint32_t n_levels = 0;                // levels 0 (leaves) .. n_levels-1 (root)
  std::vector<uint32_t> level_offset;   // host copy, n_levels + 1 entries
  DBuf center_x, center_y, half_width, mass, com_x, com_y;  // double
  DBuf body_start, body_end;            // uint32
  DBuf child_first;                     // uint32 (NONE for leaves)
  DBuf child_count;                     // uint8
  DBuf parent;                          // uint32 (NONE for the root)
  DBuf level;                           // uint8
  DBuf keys;                            // uint64, Morton key per sorted body
  // Lists as CSR (uint32 offsets): the reference's M2L and P2P lists for every
  // node, and per leaf the symmetric P2P gather list (self first).
  DBuf m2l_off, m2l_dat, p2p_off, p2p_dat, gat_off, gat_dat;
  size_t m2l_total = 0, p2p_total = 0, gat_total = 0;
  // false: the M2L rows of levels >= 2, which no evaluation reads (a leaf
  // receives only its own and its parent's M2L), are not built yet
  // (ensure_full_m2l_lists builds them; m2l_total then counts every row)
  bool m2l_full = true;
  uint64_t p2p_interactions = 0;  // directed body-body pairs of the gather lists
  P2PItems items;
  SymP2P sym;  // symmetric near field over the gather lists (built on first use)
  // per level l >= 1: every parent of level l has one child, the parent
  // level_offset[l] + i that of child level_offset[l - 1] + i (chains of such
  // levels are shifted in one launch); empty until checked for this tree
  std::vector<uint8_t> chain_ok;
  size_t capacity = 0;  // node capacity of the node arrays
};

struct Workspace {
  DBuf a, b, c, d, e, f, g, h;  // scratch reused across build stages
  DBuf scan[6];                 // partial sums of the multi-level scans
  DBuf sort_ka, sort_kb, sort_va, sort_vb, sort_hist;  // generic radix sort
};

}  // namespace gx

struct galax_fmm {
  galax_fmm_config_t cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  size_t max_n = 0;
  // bodies in tree order, owned (used by host-array entry points)
  gx::DBuf x, y, m, vx, vy, ax, ay, phi;
  gx::DevTree tree;
  bool has_tree = false;
  // expansions and M2L key tables
  gx::DBuf mp, loc, mp32, loc32, gtab, key_of_pair, perm, hash, tmp_u32, flag;
  gx::DBuf key_dxy;  // representative offset (dx, dy) per M2L key
  // M2L on the matrix cores (m2l_mfma.hip): forward tiles over targets,
  // adjoint tiles over sources, and the per-key operators
  gx::M2LTiles mx_fwd, mx_adj;
  gx::DBuf mx_t64, mx_t64T, mx_that, mx_sig, mx_zero;  // mx_zero: one zero source row (missing ids read it)
  // mixed forms: per node the exponent e of its FP32 multipole row. mp32
  // holds rows of 16 ceil(S/16) floats, zero padded: the nodes' M_ij
  // 2^(-e (i+j)), then the forward M2L's pre-summed long runs
  gx::DBuf mx_nexp;
  bool mx_rows32 = false;
  bool mx_valid = false, mx_adj_valid = false;
  // adjoint (VJP) buffers
  gx::DBuf vj_lbar, vj_mbar, vj_leaf_of, vj_fwd, vj_out;
  size_t n_keys = 0;
  size_t n_live_targets = 0, n_live_pairs = 0;
  bool keys_valid = false;
  gx::Workspace ws;
  // corrected-variant lists (dual tree traversal on the device)
  gx::DBuf c_m2l_off, c_m2l_dat, c_gat_off, c_gat_dat, c_dtt_pairs;  // c_dtt_pairs: visited pairs while building
  gx::DBuf c_m2l_order;  // M2L targets (sources, for the adjoint) by decreasing list length
  size_t c_m2l_total = 0, c_gat_total = 0, c_dtt_visited = 0;
  uint64_t c_p2p_interactions = 0;
  gx::P2PItems c_items;
  gx::SymP2P c_sym;   // symmetric near field over the corrected lists
  bool near_gather = false;  // GALAX_NEAR_GATHER=1 at create: the FP32 forms use the gather kernel
  gx::DBuf p2p_part;  // partial near fields of targets split over several items
  bool corrected_ready = false;
  galax_timings_t timings{};
  hipEvent_t ev[12]{};
  bool timing = true;     // record the per-pass events of evaluate (galax_fmm_set_timing)
  uint32_t san_cur = 0;   // sanitize counter of the last evaluate: flag[4 + san_cur]
  bool san_init = false;  // both sanitize counters zeroed
  hipEvent_t bev[6]{};  // build stages
  bool build_timed = false;
};

namespace gx {

// ---------------------------------------------------------------- scans ----
// Exclusive (or inclusive) prefix sum / max over n uint32 values in d_in,
// written to d_out (may alias). Uses ws scratch. Returns the total in *total
// when non-null (sum only; synchronises the stream).
int scan_u32(galax_fmm* h, const uint32_t* d_in, uint32_t* d_out, size_t n, bool inclusive, bool max_op,
             uint32_t* total);

// ---------------------------------------------------------- tree build ----
int build_tree_plain(galax_fmm* h, size_t n);             // bodies already in h->x/y/... unsorted
int build_tree_constrained(galax_fmm* h, size_t n);       // bodies in h->x/y/m, current order
int build_lists_reference(galax_fmm* h);                  // M2L + P2P lists + gather CSR
int ensure_full_m2l_lists(galax_fmm* h);                 // the M2L rows of levels >= 2, once per tree
// Records build-stage event i (0..5) on the handle's stream when the handle
// has its build events (created by the API entry points).
int build_mark(galax_fmm* h, int i);
int install_tree_host(galax_fmm* h, const galax_tree_view_t* t, const galax_lists_view_t* m2l,
                      const galax_lists_view_t* p2p);
int build_lists_corrected(galax_fmm* h, const double* d_x, const double* d_y);
int ensure_node_capacity(galax_fmm* h, size_t nodes);
int count_interactions(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint64_t* out);
// exact 64-bit sum of n uint32 device values (uses flag[14..15])
int sum_u32(galax_fmm* h, const uint32_t* d_in, size_t n, uint64_t* total);
int radix_sort_bits(galax_fmm* h, uint64_t* ka, uint32_t* va, uint64_t* kb, uint32_t* vb, size_t n, int bits);
// P2P work items: each target leaf's gather list in chunks of at most
// P2P_ITEM_LEAVES source leaves (balances leaves with thousands of sources).
constexpr uint32_t P2P_ITEM_LEAVES = 32;
int build_p2p_items(galax_fmm* h, const uint32_t* goff, const uint32_t* gdat, uint32_t n_leaves, P2PItems& it);
// Sums the partial near-field slots of targets split over several items into
// (o0, o1, o2) per body (o2 may be null); three doubles per body and slot.
hipError_t p2p_combine(const P2PItems& it, const uint32_t* goff, const uint32_t* bs, const uint32_t* be,
                       uint32_t n_max, const double* part, double* o0, double* o1, double* o2, hipStream_t s);

// Body index of element w0 + lane of a P2P work item's source order (w0
// uniform over the wave). The item's source leaf j covers elements
// [offs_j, offs_j + count_j) and starts at body starts_j; lane j of the wave
// holds offs_j and starts_j (j < nlv). jcur (uniform) is a leaf whose range
// starts at or before w0; on return it is the leaf of element w0 + 63, the
// start for the next window. The scan reads each candidate leaf's offset with
// one readlane and stops at the first leaf past the window, so a window of 64
// elements costs a few steps (no per-element search).
__device__ __forceinline__ uint32_t item_source(uint32_t w0, int lane, int nlv, uint32_t offs, uint32_t starts,
                                                uint32_t& jcur) {
  const uint32_t q = w0 + (uint32_t)lane;
  uint32_t jl = jcur;
  for (uint32_t j = jcur + 1; j < (uint32_t)nlv; ++j) {
    const uint32_t oj = __builtin_amdgcn_readlane(offs, j);
    if (oj > w0 + 63u) break;
    if (q >= oj) jl = j;
  }
  jcur = __builtin_amdgcn_readlane(jl, 63);
  return (uint32_t)__shfl((int)starts, (int)jl, 64) + (q - (uint32_t)__shfl((int)offs, (int)jl, 64));
}

// Near-field arithmetic type: FP64 or FP32 (v_rsq_f32) on coordinates
// relative to the target leaf's COM. FP64: the v_rsq_f64 estimate y with one
// second-order refinement, y (1 + e/2 + 3e^2/8) with e = 1 - v y^2; no
// fix-ups for v = 0 or inf (v >= eps^2 > 0 on every unmasked use; with eps = 0
// two coincident bodies give a non-finite force, as 1/sqrt(0) does).
template <typename T>
struct Near;
template <>
struct Near<double> {
  typedef double4 V4;
  static __device__ __forceinline__ double rsq(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    const double e = fma(-v * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
  }
};
template <>
struct Near<float> {
  typedef float4 V4;
  static __device__ __forceinline__ float rsq(float v) { return __builtin_amdgcn_rsqf(v); }
};

// v from lane (q + K) mod 16 or (q - K) mod 16 of the caller's 16-lane row
// (a DPP row rotation; the compiler folds it into the instruction that reads
// v). Which of the two does not matter: rotation 16 - K undoes rotation K.
template <int K>
__device__ __forceinline__ float row_rot(float v) {
  if constexpr ((K & 15) == 0)
    return v;
  else
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + (K & 15), 0xF, 0xF, true));
}

// ---------------------------------------------------------- evaluation ----
// near = false computes the expansions only (P2M .. L2L), for the VJP
int evaluate(galax_fmm* h, double* d_x, double* d_y, double* d_m, double* d_ax, double* d_ay, double* d_phi,
             bool near = true);
int vjp(galax_fmm* h, const double* d_x, const double* d_y, const double* d_m, const double* d_axb,
        const double* d_ayb, double* d_xb, double* d_yb, double* d_mb);
int m2l_mfma_setup(galax_fmm* h);
int m2l_mfma_run(galax_fmm* h, const double* mp, double* loc);
// Adjoint of the reference M2L on the matrix cores (FP64):
// mbar_s = sum over pairs (t, s) of T_K^T lbar_t, for sources s < n_live_targets.
int m2l_adjoint_setup(galax_fmm* h);
int m2l_adjoint_run(galax_fmm* h, const double* lbar, double* mbar);
int kdk_device(galax_fmm* h, double dt, uint64_t n_steps, uint64_t every, galax_diag_t* diag);

// Host restatement of the kernel-derivative numerator polynomials
// N_{k,l}[a][b] (galax-core/src/lib.rs:1059-1130); they depend on eps and p only.
std::vector<double> kernel_polys(double eps, int p);

// ------------------------------------------- corrected M2L, lane per pair ----
// One lane runs one pair's derivative recurrence order by order (fully
// unrolled) and scatters every new entry D(k, l) straight into the
// coefficients it feeds: out(I, J) += in(k - I, l - J) D(k, l) over
// I + J <= p and (k - I) + (l - J) <= p. The M2L (in = signed multipole,
// out = locals) and its adjoint (in = local cotangent, out = multipole
// cotangent) are both this contraction. Only three orders of the table are
// live, so the table, `in` and `out` stay in registers.
// Calls f(std::integral_constant<int, i>) for i = B .. E-1: loop indices that
// are constants to the front end, so `if constexpr` prunes the dead terms.
// Workgroups are dealt to the 8 XCDs in turn (block b to XCD b % 8), and each
// XCD has its own L2. This maps block b to a work item so that runs of
// XCD_RUN consecutive items (neighbouring targets, which read mostly the same
// sources) go to one XCD while the runs rotate over the XCDs, which keeps the
// balance of the round-robin deal. A bijection of [0, n) (the last partial
// group keeps the identity).
constexpr uint32_t XCD_RUN = 32;
__device__ __forceinline__ uint32_t xcd_runs(uint32_t b, uint32_t n) {
  constexpr uint32_t G = 8 * XCD_RUN;
  const uint32_t g = b / G;
  if ((g + 1) * G > n) return b;
  const uint32_t r = b % G;
  return g * G + (r % 8) * XCD_RUN + r / 8;
}

// a! and C(n, k) as compile-time constants (exact in double for n <= 18)
constexpr double ct_fact(int a) { return a <= 1 ? 1.0 : a * ct_fact(a - 1); }
constexpr double ct_binom(int n, int k) { return ct_fact(n) / (ct_fact(k) * ct_fact(n - k)); }

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <int P, typename T>
__device__ __forceinline__ void m2l_pair_scatter(T x, T y, T ir2, T d0, const T (&smv)[num_moments(P)],
                                                 T (&L)[num_moments(P)]) {
  constexpr int Q = 2 * P;
  T D[3][Q + 1];  // order n lives in D[n % 3], entry l = number of y derivatives
  static_for<0, Q + 1>([&](auto nc) {
    constexpr int n = decltype(nc)::value;
    static_for<0, n + 1>([&](auto lc) {
      constexpr int l = decltype(lc)::value;
      constexpr int k = n - l;
      T v;
      if constexpr (n == 0) {
        v = d0;
      } else {
        const T* D1 = D[(n + 2) % 3];
        const T* D2 = D[(n + 1) % 3];
        if constexpr (k >= 1) {
          v = -(T)(2 * k - 1) * x * D1[l];
          if constexpr (k >= 2) v -= (T)((k - 1) * (k - 1)) * D2[l];
          if constexpr (l >= 1) v -= (T)(2 * l) * y * D1[l - 1];
          if constexpr (l >= 2) v -= (T)(l * (l - 1)) * D2[l - 2];
        } else {
          v = -(T)(2 * l - 1) * y * D1[l - 1];
          if constexpr (l >= 2) v -= (T)((l - 1) * (l - 1)) * D2[l - 2];
        }
        v *= ir2;
      }
      D[n % 3][l] = v;
      // the locals this entry feeds: I <= k, J <= l, I + J <= P, i + j <= P
      static_for<0, (k < P ? k : P) + 1>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        constexpr int jlo = (n - P - I > 0 ? n - P - I : 0);  // J >= n - P - I
        constexpr int jhi = (l < P - I ? l : P - I);
        static_for<jlo, jhi + 1>([&](auto Jc) {
          constexpr int J = decltype(Jc)::value;
          L[moment_index(I, J)] = fma(smv[moment_index(k - I, l - J)], v, L[moment_index(I, J)]);
        });
      });
    });
  });
}

// Sums v[0..S) over the wave; lane c returns the sum of coefficient c (0 for
// c >= S). Each of the six steps halves the coefficients a lane carries: the
// lane keeps the half its lane bit selects and adds its partner's copy of it.
template <int S>
__device__ __forceinline__ double wave_sum_scatter(const double (&v)[S], int lane) {
  static_assert(S <= 64, "at most one coefficient per lane");
  double w[32];
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      const double lo_v = c < S ? v[c] : 0.0, hi_v = c + 32 < S ? v[c + 32] : 0.0;
      w[c] = (hi ? hi_v : lo_v) + __shfl_xor(hi ? lo_v : hi_v, 32, 64);
    }
  }
#pragma unroll
  for (int st = 0; st < 5; ++st) {  // a counted loop, so that it unrolls
    const int h = 16 >> st;
    const bool hi = lane & h;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      if (c >= h) continue;
      const double lo_v = w[c], hi_v = w[c + h];
      w[c] = (hi ? hi_v : lo_v) + __shfl_xor(hi ? lo_v : hi_v, h, 64);
    }
  }
  return w[0];
}

}  // namespace gx
