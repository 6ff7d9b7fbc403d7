// galax_oracle.cpp — CPU restatement of galax-core's FMM force path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product is
// compared against; it is never linked into, called by or shipped with the
// product library (galax_amd/). Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.
//
// It restates, function by function, the Rust reference knekvasil/galax
// (galax-core/src/lib.rs, galax-integrate/src/lib.rs, galax-cli/src/main.rs).
// Every function cites the reference lines it follows. Compiled with
// -ffp-contract=off so that every product is rounded before it is added, as
// Rust does (Rust never contracts a*b+c into an FMA).
//
// Two algorithm variants share the code:
//   ALGO_REFERENCE (0): exactly galax-core's semantics — 64-capped id-ordered
//     lists, the `partner > node` P2P rule, the M2L derivative cache keyed by
//     truncated COM offsets with sequential first-touch representatives (what
//     Rayon gives with one thread), ascending-id L2L.
//   ALGO_CORRECTED (1): the same tree; M2L/P2P lists from a dual tree
//     traversal with an explicit multipole acceptance test (theta); exact
//     per-pair kernel derivatives; top-down L2L. This is the physically
//     complete variant; it has no counterpart in the reference.
//
// Parity pinning of each function (what checks it besides this restatement)
// is listed in the tests under tests/test_oracle_*.py.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int MAX_P = 17;            // galax-core/src/lib.rs:6
constexpr uint32_t MORTON_BITS = 21; // galax-core/src/lib.rs:73
constexpr uint32_t LIST_CAP = 64;    // galax-core/src/lib.rs:942 ([u32; 64])
constexpr uint32_t NONE32 = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Rust scalar semantics
// ---------------------------------------------------------------------------

// `f64 as i64`: truncation toward zero, saturating, NaN -> 0.
int64_t rust_f64_as_i64(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)v;
}

// `f64 as u64`: truncation, saturating, NaN and negatives -> 0.
uint64_t rust_f64_as_u64(double v) {
  if (std::isnan(v) || v <= 0.0) return 0;
  if (v >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)v;
}

// `f64::powi` as LLVM lowers a non-constant exponent: compiler-rt's
// __powidf2 (binary exponentiation, r *= a on set bits, a *= a between).
double powi(double a, int b) {
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

// galax-core/src/lib.rs:8-13
void fill_powers(double* arr, double base, int p) {
  arr[0] = 1.0;
  for (int k = 1; k <= p; ++k) arr[k] = arr[k - 1] * base;
}

// galax-core/src/lib.rs:15-28
struct Tables {
  double binom[MAX_P][MAX_P];
  double fact[MAX_P];
  Tables() {
    for (int n = 0; n < MAX_P; ++n)
      for (int k = 0; k < MAX_P; ++k) binom[n][k] = 0.0;
    for (int n = 0; n < MAX_P; ++n) {
      binom[n][0] = 1.0;
      binom[n][n] = 1.0;
      for (int k = 1; k < n; ++k) binom[n][k] = binom[n - 1][k - 1] + binom[n - 1][k];
    }
    // galax-core/src/lib.rs:30-39
    fact[0] = 1.0;
    for (int i = 1; i < MAX_P; ++i) fact[i] = fact[i - 1] * (double)i;
  }
};
const Tables& tables() {
  static Tables t;
  return t;
}

inline size_t num_moments(int p) { return (size_t)(p + 1) * (p + 2) / 2; }  // :825-827
inline size_t moment_index(int i, int j) {                                   // :830-833
  int n = i + j;
  return (size_t)n * (n + 1) / 2 + j;
}

// ---------------------------------------------------------------------------
// Morton keys — galax-core/src/lib.rs:73-115
// ---------------------------------------------------------------------------

uint64_t normalize(double x, double lo, double hi) {
  double t = (x - lo) / (hi - lo);
  // f64::clamp(0.0, 1.0 - 1e-15): NaN passes through
  const double top = 1.0 - 1e-15;
  if (t < 0.0) t = 0.0;
  if (t > top) t = top;
  return rust_f64_as_u64(t * (double)(1ull << MORTON_BITS));
}

double denormalize(uint64_t v, double lo, double hi) {
  double cell = (hi - lo) / (double)(1ull << MORTON_BITS);
  return lo + ((double)v + 0.5) * cell;
}

uint64_t spread_bits(uint64_t v) {
  uint64_t x = v & ((1ull << MORTON_BITS) - 1);
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

uint64_t compact_bits(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
  x = (x | (x >> 16)) & 0x00000000ffffffffull;
  return x;
}

uint64_t morton_encode(double x, double y, double x0, double x1, double y0, double y1) {
  return spread_bits(normalize(x, x0, x1)) | (spread_bits(normalize(y, y0, y1)) << 1);
}

// ---------------------------------------------------------------------------
// Tree — galax-core/src/lib.rs:117-361, 541-694
// ---------------------------------------------------------------------------

struct Node {
  double center_x, center_y, half_width, mass, com_x, com_y;
  uint64_t body_start, body_end;
  int32_t children[4];  // -1 = None
  int32_t parent;       // -1 = None
  uint8_t is_leaf;
};

struct Tree {
  std::vector<Node> nodes;
  uint64_t n_max;
  double x_min, x_max, y_min, y_max;
};

struct Bodies {  // BodiesSoA (galax-core/src/lib.rs:41-49), caller-owned
  double *x, *y, *vx, *vy, *mass, *ax, *ay;
  size_t n;
};

// Leaf geometry: galax-core/src/lib.rs:229-267 (identical at :617-640)
Node make_leaf(const Bodies& b, uint64_t start, uint64_t end) {
  double mass_sum = 0.0, com_x = 0.0, com_y = 0.0;
  double bx_min = 1.7976931348623157e308, bx_max = -1.7976931348623157e308;
  double by_min = 1.7976931348623157e308, by_max = -1.7976931348623157e308;
  for (uint64_t idx = start; idx < end; ++idx) {
    double m = b.mass[idx];
    mass_sum += m;
    com_x += b.x[idx] * m;
    com_y += b.y[idx] * m;
    bx_min = std::fmin(bx_min, b.x[idx]);  // f64::min
    bx_max = std::fmax(bx_max, b.x[idx]);
    by_min = std::fmin(by_min, b.y[idx]);
    by_max = std::fmax(by_max, b.y[idx]);
  }
  Node nd;
  nd.center_x = (bx_min + bx_max) * 0.5;
  nd.center_y = (by_min + by_max) * 0.5;
  nd.half_width = std::fmax(bx_max - bx_min, by_max - by_min) * 0.5 + 1e-15;
  nd.mass = mass_sum;
  nd.com_x = mass_sum > 0.0 ? com_x / mass_sum : nd.center_x;
  nd.com_y = mass_sum > 0.0 ? com_y / mass_sum : nd.center_y;
  nd.body_start = start;
  nd.body_end = end;
  for (int k = 0; k < 4; ++k) nd.children[k] = -1;
  nd.parent = -1;
  nd.is_leaf = 1;
  return nd;
}

// Parent rounds: galax-core/src/lib.rs:269-354. Tree::build compares with
// `if a < b` / `if a > b` (:307-317); build_with_constraints uses f64::min/max
// (:663-667). The two agree for finite values, which is all a tree holds after
// leaf creation from finite bodies; `constrained_style` selects the latter.
void build_parents(std::vector<Node>& nodes, const std::vector<uint64_t>& keys, bool constrained_style) {
  size_t current_start = 0, current_end = nodes.size();
  uint32_t group_shift = 2;
  while (current_end - current_start > 1) {
    size_t i = current_start;
    while (i < current_end) {
      // Rust release builds mask the shift amount to 6 bits.
      uint64_t group_id = keys[nodes[i].body_start] >> (group_shift & 63);
      size_t j = i + 1;
      while (j < current_end && j - i < 4) {
        if ((keys[nodes[j].body_start] >> (group_shift & 63)) != group_id) break;
        ++j;
      }
      double mass_sum = 0.0, com_x = 0.0, com_y = 0.0;
      double bx_min = 1.7976931348623157e308, bx_max = -1.7976931348623157e308;
      double by_min = 1.7976931348623157e308, by_max = -1.7976931348623157e308;
      double max_child_hw = 0.0;
      for (size_t si = i; si < j; ++si) {
        const Node& c = nodes[si];
        mass_sum += c.mass;
        com_x += c.com_x * c.mass;
        com_y += c.com_y * c.mass;
        if (constrained_style) {
          max_child_hw = std::fmax(max_child_hw, c.half_width);
          bx_min = std::fmin(bx_min, c.center_x - c.half_width);
          bx_max = std::fmax(bx_max, c.center_x + c.half_width);
          by_min = std::fmin(by_min, c.center_y - c.half_width);
          by_max = std::fmax(by_max, c.center_y + c.half_width);
        } else {
          if (c.half_width > max_child_hw) max_child_hw = c.half_width;
          double cl = c.center_x - c.half_width, cr = c.center_x + c.half_width;
          if (cl < bx_min) bx_min = cl;
          if (cr > bx_max) bx_max = cr;
          double cb = c.center_y - c.half_width, ct = c.center_y + c.half_width;
          if (cb < by_min) by_min = cb;
          if (ct > by_max) by_max = ct;
        }
      }
      Node pn;
      pn.center_x = (bx_min + bx_max) * 0.5;
      pn.center_y = (by_min + by_max) * 0.5;
      pn.half_width = max_child_hw * 2.0;
      pn.mass = mass_sum;
      pn.com_x = mass_sum > 0.0 ? com_x / mass_sum : pn.center_x;
      pn.com_y = mass_sum > 0.0 ? com_y / mass_sum : pn.center_y;
      pn.body_start = nodes[i].body_start;
      pn.body_end = nodes[j - 1].body_end;
      for (int k = 0; k < 4; ++k) pn.children[k] = -1;
      int32_t pid = (int32_t)nodes.size();
      for (size_t si = i; si < j; ++si) {
        nodes[si].parent = pid;
        pn.children[si - i] = (int32_t)si;
      }
      pn.parent = -1;
      pn.is_leaf = 0;
      nodes.push_back(pn);
      i = j;
    }
    current_start = current_end;
    current_end = nodes.size();
    group_shift += 2;
  }
}

// Tree::permute — galax-core/src/lib.rs:146-161
void permute(Bodies& b, const std::vector<size_t>& idx) {
  std::vector<double> tmp(b.n);
  double* arrs[7] = {b.x, b.y, b.vx, b.vy, b.mass, b.ax, b.ay};
  for (double* a : arrs) {
    if (!a) continue;
    for (size_t k = 0; k < b.n; ++k) tmp[k] = a[idx[k]];
    std::memcpy(a, tmp.data(), b.n * sizeof(double));
  }
}

// Tree::build — galax-core/src/lib.rs:163-361
Tree tree_build(Bodies& b, uint64_t n_max, double x0, double x1, double y0, double y1,
                std::vector<uint64_t>* perm_out) {
  Tree t;
  t.n_max = n_max;
  t.x_min = x0; t.x_max = x1; t.y_min = y0; t.y_max = y1;
  size_t n = b.n;
  if (n == 0) return t;
  std::vector<uint64_t> keys(n);
  for (size_t i = 0; i < n; ++i) keys[i] = morton_encode(b.x[i], b.y[i], x0, x1, y0, y1);
  std::vector<size_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  // sort_by_key is stable (:192)
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return keys[a] < keys[c]; });
  permute(b, idx);
  if (perm_out) perm_out->assign(idx.begin(), idx.end());
  for (size_t i = 0; i < n; ++i) keys[i] = morton_encode(b.x[i], b.y[i], x0, x1, y0, y1);
  for (uint64_t s = 0; s < n;) {
    uint64_t e = std::min<uint64_t>(s + n_max, n);
    t.nodes.push_back(make_leaf(b, s, e));
    s = e;
  }
  build_parents(t.nodes, keys, false);
  return t;
}

// build_with_constraints — galax-core/src/lib.rs:592-694 (no sort, :598-600)
Tree build_with_constraints(Bodies& b, uint64_t n_max, double x0, double x1, double y0, double y1,
                            const std::vector<uint64_t>& splits) {
  Tree t;
  t.n_max = n_max;
  t.x_min = x0; t.x_max = x1; t.y_min = y0; t.y_max = y1;
  size_t n = b.n;
  std::vector<uint64_t> keys(n);
  for (size_t i = 0; i < n; ++i) keys[i] = morton_encode(b.x[i], b.y[i], x0, x1, y0, y1);
  for (uint64_t s = 0; s < n;) {
    uint64_t bound = std::min<uint64_t>(s + n_max, n);
    uint64_t e = bound;
    for (uint64_t sp : splits)
      if (sp > s && sp < bound && sp < e) e = sp;
    t.nodes.push_back(make_leaf(b, s, e));
    s = e;
  }
  build_parents(t.nodes, keys, true);
  return t;
}

// cells_adjacent — galax-core/src/lib.rs:369-376
bool cells_adjacent(const Node& a, const Node& b) {
  double ratio = std::fmax(a.half_width, b.half_width) / std::fmin(a.half_width, b.half_width);
  if (ratio > 8.0) return false;
  double dx = std::fabs(a.center_x - b.center_x);
  double dy = std::fabs(a.center_y - b.center_y);
  return dx <= a.half_width + b.half_width + 1e-15 && dy <= a.half_width + b.half_width + 1e-15;
}

// neighbors_geometric — galax-core/src/lib.rs:382-423
void neighbors_geometric(const Tree& t, size_t node_id, int64_t out[8]) {
  const Node& na = t.nodes[node_id];
  double cx = na.center_x, cy = na.center_y, hw = na.half_width;
  const double pts[8][2] = {
      {cx + hw * 2.0, cy},           {cx + hw * 2.0, cy + hw * 2.0}, {cx, cy + hw * 2.0},
      {cx - hw * 2.0, cy + hw * 2.0}, {cx - hw * 2.0, cy},           {cx - hw * 2.0, cy - hw * 2.0},
      {cx, cy - hw * 2.0},           {cx + hw * 2.0, cy - hw * 2.0}};
  for (int k = 0; k < 8; ++k) {
    double px = pts[k][0], py = pts[k][1];
    int64_t best = -1;
    double best_dist = 1.7976931348623157e308;
    for (size_t o = 0; o < t.nodes.size(); ++o) {
      if (o == node_id) continue;
      const Node& nb = t.nodes[o];
      double dx = std::fabs(px - nb.center_x), dy = std::fabs(py - nb.center_y);
      if (dx <= nb.half_width + 1e-15 && dy <= nb.half_width + 1e-15) {
        if (cells_adjacent(na, nb)) {
          double ex = px - nb.center_x, ey = py - nb.center_y;
          double dist = ex * ex + ey * ey;  // powi(2) == x*x
          if (dist < best_dist) {
            best_dist = dist;
            best = (int64_t)o;
          }
        }
      }
    }
    out[k] = best;
  }
}

// build_constrained — galax-core/src/lib.rs:541-588
Tree build_constrained(Bodies& b, uint64_t n_max, double x0, double x1, double y0, double y1) {
  std::vector<uint64_t> split_points;
  for (int iter = 0; iter < 10; ++iter) {
    Tree t = build_with_constraints(b, n_max, x0, x1, y0, y1, split_points);
    std::vector<uint64_t> new_splits;
    for (size_t lid = 0; lid < t.nodes.size(); ++lid) {
      if (!t.nodes[lid].is_leaf) continue;
      const Node& coarse = t.nodes[lid];
      double hw = coarse.half_width;
      int64_t neigh[8];
      neighbors_geometric(t, lid, neigh);
      for (int k = 0; k < 8; ++k) {
        if (neigh[k] < 0) continue;
        double nhw = t.nodes[neigh[k]].half_width;
        if (hw > nhw * 2.1) {
          uint64_t s = coarse.body_start, e = coarse.body_end;
          if (e - s > 2) {
            uint64_t mid = (s + e) / 2;
            if (std::find(split_points.begin(), split_points.end(), mid) == split_points.end() &&
                std::find(new_splits.begin(), new_splits.end(), mid) == new_splits.end())
              new_splits.push_back(mid);
          }
          break;
        }
      }
    }
    if (new_splits.empty()) return t;
    split_points.insert(split_points.end(), new_splits.begin(), new_splits.end());
    std::sort(split_points.begin(), split_points.end());
    split_points.erase(std::unique(split_points.begin(), split_points.end()), split_points.end());
  }
  return build_with_constraints(b, n_max, x0, x1, y0, y1, split_points);
}

// compute_levels — galax-core/src/lib.rs:426-438 (same loop at :952-961)
std::vector<uint32_t> compute_levels(const Tree& t) {
  std::vector<uint32_t> lv(t.nodes.size(), 0);
  for (size_t id = 0; id < t.nodes.size(); ++id) {
    uint32_t m = 0;
    for (int k = 0; k < 4; ++k)
      if (t.nodes[id].children[k] >= 0) m = std::max(m, lv[t.nodes[id].children[k]] + 1);
    lv[id] = m;
  }
  return lv;
}

// ---------------------------------------------------------------------------
// Interaction lists — galax-core/src/lib.rs:941-1026, 1358-1392
// ---------------------------------------------------------------------------

struct Lists {
  // cap == LIST_CAP reproduces InteractionLists{data:[u32;64], lengths:u8};
  // cap == 0 means uncapped (CORRECTED variant).
  std::vector<std::vector<uint32_t>> data;
};

inline bool touch(const Node& a, const Node& b) {
  // the box test of :990-995, :1000-1004 and :1380-1383 (one symmetric predicate)
  double dx = std::fabs(a.center_x - b.center_x);
  double dy = std::fabs(a.center_y - b.center_y);
  double h = a.half_width + b.half_width;
  return dx <= h + 1e-15 && dy <= h + 1e-15;
}

// InteractionLists::build — galax-core/src/lib.rs:947-1025
Lists m2l_lists_build(const Tree& t) {
  size_t nn = t.nodes.size();
  Lists L;
  L.data.assign(nn, {});
  std::vector<uint32_t> level = compute_levels(t);
  uint32_t max_level = 0;
  for (uint32_t v : level) max_level = std::max(max_level, v);
  std::vector<std::vector<size_t>> by_level(max_level + 1);
  for (size_t id = 0; id < nn; ++id) by_level[level[id]].push_back(id);
  for (uint32_t lv = 0; lv < max_level; ++lv) {
    const auto& at_lv = by_level[lv];
    const auto& at_par = by_level[lv + 1];
    for (size_t node_id : at_lv) {
      const Node& node = t.nodes[node_id];
      if (node.parent < 0) continue;
      size_t parent_id = (size_t)node.parent;
      const Node& parent = t.nodes[parent_id];
      auto& out = L.data[node_id];
      for (size_t pn_id : at_par) {
        if (pn_id == parent_id) continue;
        const Node& pn = t.nodes[pn_id];
        double dx = std::fabs(pn.center_x - parent.center_x);
        double dy = std::fabs(pn.center_y - parent.center_y);
        double h_sum = parent.half_width + pn.half_width;
        if (dx > h_sum + 1e-15 || dy > h_sum + 1e-15) continue;
        for (int k = 0; k < 4; ++k) {
          int32_t c = pn.children[k];
          if (c < 0) continue;
          if (touch(t.nodes[c], node)) continue;
          if (out.size() < LIST_CAP) out.push_back((uint32_t)c);
        }
      }
    }
  }
  return L;
}

// ALGO_CORRECTED lists (no counterpart in the reference): a dual tree
// traversal over the same tree with the acceptance test
//   R_A + R_B < theta * |com_A - com_B|,
// R_n = max distance from node n's COM to its bodies. Accepted pairs go to
// both nodes' M2L lists; unaccepted leaf pairs to both leaves' P2P lists
// (self first). The larger-radius non-leaf node of a pair is split.
std::vector<double> node_radii(const Tree& t, const double* x, const double* y) {
  std::vector<double> R(t.nodes.size());
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const Node& n = t.nodes[i];
    double r2 = 0.0;
    for (uint64_t b = n.body_start; b < n.body_end; ++b) {
      double dx = x[b] - n.com_x, dy = y[b] - n.com_y;
      r2 = std::fmax(r2, dx * dx + dy * dy);
    }
    R[i] = std::sqrt(r2);
  }
  return R;
}

void dtt(const Tree& t, const std::vector<double>& R, double theta, uint32_t a, uint32_t b, Lists& m2l_l,
         Lists& p2p_l) {
  const Node& A = t.nodes[a];
  const Node& B = t.nodes[b];
  if (a == b) {
    if (A.is_leaf) {
      p2p_l.data[a].insert(p2p_l.data[a].begin(), a);
      return;
    }
    for (int i = 0; i < 4; ++i)
      for (int j = i; j < 4; ++j)
        if (A.children[i] >= 0 && A.children[j] >= 0)
          dtt(t, R, theta, (uint32_t)A.children[i], (uint32_t)A.children[j], m2l_l, p2p_l);
    return;
  }
  double dx = A.com_x - B.com_x, dy = A.com_y - B.com_y;
  double d = std::sqrt(dx * dx + dy * dy);
  if (R[a] + R[b] < theta * d) {
    m2l_l.data[a].push_back(b);
    m2l_l.data[b].push_back(a);
    return;
  }
  if (A.is_leaf && B.is_leaf) {
    p2p_l.data[a].push_back(b);
    p2p_l.data[b].push_back(a);
    return;
  }
  bool split_a = B.is_leaf || (!A.is_leaf && R[a] >= R[b]);
  if (split_a) {
    for (int i = 0; i < 4; ++i)
      if (A.children[i] >= 0) dtt(t, R, theta, (uint32_t)A.children[i], b, m2l_l, p2p_l);
  } else {
    for (int i = 0; i < 4; ++i)
      if (B.children[i] >= 0) dtt(t, R, theta, a, (uint32_t)B.children[i], m2l_l, p2p_l);
  }
}

void corrected_lists_build(const Tree& t, const double* x, const double* y, double theta, Lists& m2l_l,
                           Lists& p2p_l) {
  m2l_l.data.assign(t.nodes.size(), {});
  p2p_l.data.assign(t.nodes.size(), {});
  if (t.nodes.empty()) return;
  std::vector<double> R = node_radii(t, x, y);
  uint32_t root = (uint32_t)t.nodes.size() - 1;
  dtt(t, R, theta, root, root, m2l_l, p2p_l);
  // canonical order: self first, then ascending ids
  for (auto& v : p2p_l.data)
    if (!v.empty()) std::sort(v.begin() + 1, v.end());
  for (auto& v : m2l_l.data) std::sort(v.begin(), v.end());
}

// build_p2p_lists — galax-core/src/lib.rs:1358-1392
Lists p2p_lists_build(const Tree& t) {
  size_t nn = t.nodes.size();
  Lists L;
  L.data.assign(nn, {});
  std::vector<size_t> leaf_ids;
  for (size_t id = 0; id < nn; ++id)
    if (t.nodes[id].is_leaf) leaf_ids.push_back(id);
  for (size_t a : leaf_ids) {
    auto& out = L.data[a];
    out.push_back((uint32_t)a);
    const Node& na = t.nodes[a];
    for (size_t b : leaf_ids) {
      if (a == b) continue;
      const Node& nb = t.nodes[b];
      double dx = std::fabs(na.center_x - nb.center_x);
      double dy = std::fabs(na.center_y - nb.center_y);
      if (dx <= na.half_width + nb.half_width + 1e-15 && dy <= na.half_width + nb.half_width + 1e-15) {
        if (out.size() < LIST_CAP) out.push_back((uint32_t)b);
      }
    }
  }
  return L;
}

// ---------------------------------------------------------------------------
// FMM operators — galax-core/src/lib.rs:848-1355
// ---------------------------------------------------------------------------

// p2m — galax-core/src/lib.rs:848-887
void p2m(const double* x, const double* y, const double* mass, const Tree& t, int p, double* mp) {
  size_t S = num_moments(p);
  const double* fact = tables().fact;
  std::fill(mp, mp + t.nodes.size() * S, 0.0);
  for (size_t id = 0; id < t.nodes.size(); ++id) {
    const Node& nd = t.nodes[id];
    if (!nd.is_leaf) continue;
    double* base = mp + id * S;
    for (uint64_t b = nd.body_start; b < nd.body_end; ++b) {
      double dx = x[b] - nd.com_x, dy = y[b] - nd.com_y, m = mass[b];
      double dxp[MAX_P], dyp[MAX_P];
      fill_powers(dxp, dx, p);
      fill_powers(dyp, dy, p);
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p - i; ++j) base[moment_index(i, j)] += m * dxp[i] * dyp[j] / (fact[i] * fact[j]);
    }
  }
  for (size_t k = 0; k < t.nodes.size() * S; ++k)
    if (!std::isfinite(mp[k])) mp[k] = 0.0;
}

// m2m — galax-core/src/lib.rs:890-938
void m2m(double* mp, const Tree& t, int p) {
  size_t S = num_moments(p);
  const double* fact = tables().fact;
  for (size_t id = 0; id < t.nodes.size(); ++id) {
    const Node& nd = t.nodes[id];
    if (nd.is_leaf) continue;
    double* pb = mp + id * S;
    for (int k = 0; k < 4; ++k) {
      int32_t c = nd.children[k];
      if (c < 0) continue;
      const Node& ch = t.nodes[c];
      double dx = ch.com_x - nd.com_x, dy = ch.com_y - nd.com_y;
      const double* cb = mp + (size_t)c * S;
      double dxp[MAX_P], dyp[MAX_P];
      fill_powers(dxp, dx, p);
      fill_powers(dyp, dy, p);
      for (int oi = 0; oi <= p; ++oi)
        for (int oj = 0; oj <= p - oi; ++oj) {
          double sum = 0.0;
          for (int i = 0; i <= oi; ++i)
            for (int j = 0; j <= oj; ++j) {
              if (i + j > p) continue;
              sum += cb[moment_index(i, j)] * dxp[oi - i] * dyp[oj - j] / (fact[oi - i] * fact[oj - j]);
            }
          pb[moment_index(oi, oj)] += sum;
        }
    }
  }
  for (size_t k = 0; k < t.nodes.size() * S; ++k)
    if (!std::isfinite(mp[k])) mp[k] = 0.0;
}

// kernel_deriv_idx — galax-core/src/lib.rs:1033-1036
inline size_t kernel_deriv_idx(int k, int l, int p) { return (size_t)(k + l) * (2 * p + 1) + l; }

// compute_kernel_derivs — galax-core/src/lib.rs:1040-1151. Numerator
// polynomials N_{k,l} by the recurrence N_{k+1,l} = R^2 dN/dx - (1+2k+2l) x N
// (and the y analogue when k == 0), evaluated over R^{1+2(k+l)}.
void compute_kernel_derivs(double dx, double dy, double eps, int p, double* deriv) {
  int order = 2 * p;
  int stride = order + 1;
  std::fill(deriv, deriv + (size_t)stride * stride, 0.0);
  double x = dx, y = dy;
  double eps2 = eps * eps;
  double r2 = x * x + y * y + eps2;
  double r = std::sqrt(r2);
  std::vector<double> inv_r(2 + 2 * order, 0.0);
  inv_r[1] = 1.0 / r;
  for (int n = 3; n <= 1 + 2 * order; n += 2) inv_r[n] = inv_r[n - 2] / r2;
  auto pidx = [&](int k, int l) { return (size_t)k * stride + l; };
  std::vector<std::vector<double>> poly((size_t)stride * stride, std::vector<double>((size_t)stride * stride, 0.0));
  auto at = [&](std::vector<double>& c, int a, int b) -> double& { return c[(size_t)a * stride + b]; };
  at(poly[pidx(0, 0)], 0, 0) = -1.0;
  for (int n = 1; n <= order; ++n) {
    for (int k = 0; k <= n; ++k) {
      int l = n - k;
      std::vector<double> coeff((size_t)stride * stride, 0.0);
      if (k > 0) {
        std::vector<double>& prev = poly[pidx(k - 1, l)];
        int m = 1 + 2 * (k - 1) + 2 * l;
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = at(prev, a, b);
            if (c == 0.0) continue;
            if (a > 0) {
              double dc = c * (double)a;
              if (a + 1 < stride) at(coeff, a + 1, b) += dc;
              if (b + 2 < stride) at(coeff, a - 1, b + 2) += dc;
              at(coeff, a - 1, b) += dc * eps2;
            }
          }
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = at(prev, a, b);
            if (c == 0.0) continue;
            if (a + 1 < stride) at(coeff, a + 1, b) -= (double)m * c;
          }
      } else if (l > 0) {
        std::vector<double>& prev = poly[pidx(k, l - 1)];
        int m = 1 + 2 * k + 2 * (l - 1);
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = at(prev, a, b);
            if (c == 0.0) continue;
            if (b > 0) {
              double dc = c * (double)b;
              if (a + 2 < stride) at(coeff, a + 2, b - 1) += dc;
              if (b + 1 < stride) at(coeff, a, b + 1) += dc;
              at(coeff, a, b - 1) += dc * eps2;
            }
          }
        for (int a = 0; a < stride; ++a)
          for (int b = 0; b < stride - a; ++b) {
            double c = at(prev, a, b);
            if (c == 0.0) continue;
            if (b + 1 < stride) at(coeff, a, b + 1) -= (double)m * c;
          }
      }
      poly[pidx(k, l)] = std::move(coeff);
    }
  }
  for (int k = 0; k <= order; ++k)
    for (int l = 0; l <= order - k; ++l) {
      double val = 0.0;
      std::vector<double>& c = poly[pidx(k, l)];
      for (int a = 0; a < stride; ++a)
        for (int b = 0; b < stride - a; ++b) {
          double cc = at(c, a, b);
          if (cc != 0.0) val += cc * powi(x, a) * powi(y, b);
        }
      deriv[kernel_deriv_idx(k, l, p)] = val * inv_r[1 + 2 * k + 2 * l];
    }
}

// m2l_plan — galax-core/src/lib.rs:1153-1196
struct M2lTerm {
  size_t src, tgt, deriv;
  double weight;
};
std::vector<M2lTerm> m2l_plan(int p) {
  const double* fact = tables().fact;
  std::vector<M2lTerm> plan;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= p - i; ++j) {
      double sign = ((i + j) % 2 == 0) ? 1.0 : -1.0;
      size_t src = moment_index(i, j);
      for (int oi = 0; oi <= p; ++oi)
        for (int oj = 0; oj <= p - oi; ++oj) {
          double l_factor = 1.0 / (fact[oi] * fact[oj]);
          int dk = i + oi, dl = j + oj;
          if (dk + dl <= 2 * p) plan.push_back({src, moment_index(oi, oj), kernel_deriv_idx(dk, dl, p), sign * l_factor});
        }
    }
  return plan;
}

// m2l — galax-core/src/lib.rs:1199-1259.
//   cache_mode 0: the reference's derivative cache keyed by
//     (dx as i64, dy as i64), sequential first-touch (1-thread Rayon).
//   cache_mode 1: no cache, exact per-pair derivatives (CORRECTED).
//   cache_mode 2: the reference's cache under `threads` OpenMP threads with a
//     mutex, first writer wins as under Rayon (timing mode; racy by design).
// min_level/max_level restrict targets by tree level (UINT32_MAX = all).
void m2l(double* locals, const double* mp, const Lists& lists, const Tree& t, int p, double eps,
         int cache_mode, int threads, const std::vector<uint32_t>* level, uint32_t max_target_level) {
  size_t S = num_moments(p);
  const std::vector<M2lTerm> plan = m2l_plan(p);
  size_t dsz = (size_t)(2 * p + 1) * (2 * p + 1);
  std::map<std::pair<int64_t, int64_t>, std::vector<double>> cache;
  std::mutex mu;
  long nn = (long)t.nodes.size();
  auto body = [&](long node_id) {
    const auto& lst = lists.data[node_id];
    if (lst.empty()) return;
    if (level && (*level)[node_id] > max_target_level) return;
    const Node& tg = t.nodes[node_id];
    double* lc = locals + (size_t)node_id * S;
    std::vector<double> deriv(dsz);
    for (uint32_t src_id : lst) {
      const Node& sr = t.nodes[src_id];
      const double* sb = mp + (size_t)src_id * S;
      double dx = tg.com_x - sr.com_x, dy = tg.com_y - sr.com_y;
      if (cache_mode == 1) {
        compute_kernel_derivs(dx, dy, eps, p, deriv.data());
      } else {
        auto key = std::make_pair(rust_f64_as_i64(dx), rust_f64_as_i64(dy));
        bool hit = false;
        {
          std::lock_guard<std::mutex> g(mu);
          auto it = cache.find(key);
          if (it != cache.end()) {
            deriv = it->second;
            hit = true;
          }
        }
        if (!hit) {
          compute_kernel_derivs(dx, dy, eps, p, deriv.data());
          std::lock_guard<std::mutex> g(mu);
          cache.emplace(key, deriv);  // or_insert: first writer wins
        }
      }
      for (const M2lTerm& tm : plan) {
        double mv = sb[tm.src];
        if (mv == 0.0) continue;
        lc[tm.tgt] += mv * tm.weight * deriv[tm.deriv];
      }
    }
  };
  if (cache_mode == 2 && threads > 1) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (long id = 0; id < nn; ++id) body(id);
  } else {
    for (long id = 0; id < nn; ++id) body(id);
  }
}

// l2l — galax-core/src/lib.rs:1262-1302. descending == false is the
// reference's ascending-id order; true is top-down (CORRECTED).
void l2l(double* locals, const Tree& t, int p, bool descending) {
  size_t S = num_moments(p);
  const auto& binom = tables().binom;
  size_t nn = t.nodes.size();
  for (size_t k = 0; k < nn; ++k) {
    size_t id = descending ? nn - 1 - k : k;
    const Node& nd = t.nodes[id];
    if (nd.is_leaf) continue;
    const double* pb = locals + id * S;
    for (int c = 0; c < 4; ++c) {
      int32_t ch = nd.children[c];
      if (ch < 0) continue;
      const Node& cn = t.nodes[ch];
      double dx = cn.com_x - nd.com_x, dy = cn.com_y - nd.com_y;
      double* cb = locals + (size_t)ch * S;
      double dxp[MAX_P], dyp[MAX_P];
      fill_powers(dxp, dx, p);
      fill_powers(dyp, dy, p);
      for (int oi = 0; oi <= p; ++oi)
        for (int oj = 0; oj <= p - oi; ++oj) {
          double sum = 0.0;
          for (int i = oi; i <= p; ++i)
            for (int j = oj; j <= p - i; ++j)
              sum += pb[moment_index(i, j)] * binom[i][oi] * binom[j][oj] * dxp[i - oi] * dyp[j - oj];
          cb[moment_index(oi, oj)] += sum;
        }
    }
  }
}

// l2p — galax-core/src/lib.rs:1308-1355. Also evaluates the local expansion
// itself into phi when phi != nullptr (new output; the reference has none).
void l2p(const double* x, const double* y, double* ax, double* ay, double* phi, const double* locals,
         const Tree& t, int p, int threads) {
  size_t S = num_moments(p);
  long nn = (long)t.nodes.size();
  auto body = [&](long id) {
    const Node& nd = t.nodes[id];
    if (!nd.is_leaf) return;
    const double* base = locals + (size_t)id * S;
    for (uint64_t b = nd.body_start; b < nd.body_end; ++b) {
      double dux = x[b] - nd.com_x, duy = y[b] - nd.com_y;
      double xp[MAX_P], yp[MAX_P];
      fill_powers(xp, dux, p);
      fill_powers(yp, duy, p);
      double sax = 0.0, say = 0.0, sphi = 0.0;
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p - i; ++j) {
          double c = base[moment_index(i, j)];
          if (i >= 1) sax += c * (double)i * xp[i - 1] * yp[j];
          if (j >= 1) say += c * (double)j * xp[i] * yp[j - 1];
          sphi += c * xp[i] * yp[j];
        }
      ax[b] -= sax;
      ay[b] -= say;
      if (phi) phi[b] += sphi;
    }
  };
  // The reference collects per-leaf results in parallel then applies them
  // sequentially (:1311-1354); leaves own disjoint bodies, so a parallel apply
  // gives identical bits.
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
  for (long id = 0; id < nn; ++id) body(id);
}

// p2p — galax-core/src/lib.rs:1395-1462. phi (new output) accumulates the
// softened pair potential -m/sqrt(r^2+eps^2) with the same pair rule.
void p2p(const double* x, const double* y, const double* mass, double* ax, double* ay, double* phi,
         const Tree& t, const Lists& lists, double softening) {
  double soft2 = softening * softening;
  for (size_t id = 0; id < t.nodes.size(); ++id) {
    if (!t.nodes[id].is_leaf) continue;
    const auto& lst = lists.data[id];
    for (uint32_t partner : lst) {
      if (partner == id) {
        uint64_t s = t.nodes[id].body_start, e = t.nodes[id].body_end;
        for (uint64_t i = s; i < e; ++i) {
          double axi = 0.0, ayi = 0.0, phii = 0.0;
          for (uint64_t j = i + 1; j < e; ++j) {
            double dx = x[j] - x[i], dy = y[j] - y[i];
            double r2 = dx * dx + dy * dy + soft2;
            double inv_r = 1.0 / std::sqrt(r2);
            double inv_r3 = powi(inv_r, 3);
            double f = mass[j] * inv_r3;
            axi += f * dx;
            ayi += f * dy;
            ax[j] -= f * dx * mass[i] / mass[j];
            ay[j] -= f * dy * mass[i] / mass[j];
            if (phi) {
              phii -= mass[j] * inv_r;
              phi[j] -= mass[i] * inv_r;
            }
          }
          ax[i] += axi;
          ay[i] += ayi;
          if (phi) phi[i] += phii;
        }
      } else if (partner < id) {
        continue;
      } else {
        uint64_t pb = t.nodes[partner].body_start, pe = t.nodes[partner].body_end;
        uint64_t ma = t.nodes[id].body_start, mb = t.nodes[id].body_end;
        for (uint64_t i = ma; i < mb; ++i) {
          double axi = 0.0, ayi = 0.0, phii = 0.0;
          for (uint64_t j = pb; j < pe; ++j) {
            double dx = x[j] - x[i], dy = y[j] - y[i];
            double r2 = dx * dx + dy * dy + soft2;
            double inv_r = 1.0 / std::sqrt(r2);
            double inv_r3 = powi(inv_r, 3);
            double f = mass[j] * inv_r3;
            axi += f * dx;
            ayi += f * dy;
            ax[j] -= f * dx * mass[i] / mass[j];
            ay[j] -= f * dy * mass[i] / mass[j];
            if (phi) {
              phii -= mass[j] * inv_r;
              phi[j] -= mass[i] * inv_r;
            }
          }
          ax[i] += axi;
          ay[i] += ayi;
          if (phi) phi[i] += phii;
        }
      }
    }
  }
}

// compute_fmm_force — galax-core/src/lib.rs:1466-1499 (sanitize, zero,
// P2M -> M2M -> M2L -> L2L -> L2P -> P2P). Returns the number of M2L
// derivative-cache keys (0 for the uncached variant).
struct Pass {
  double t[8];
};

size_t compute_fmm_force(double* x, double* y, double* mass, double* ax, double* ay, double* phi, size_t n,
                         const Tree& t, const Lists& m2l_l, const Lists& p2p_l, int p, double eps, int algo,
                         int threads, double* pass_times) {
  for (size_t i = 0; i < n; ++i) {
    if (std::fabs(x[i]) > 1e100 || std::fabs(y[i]) > 1e100 || !std::isfinite(x[i]) || !std::isfinite(y[i])) {
      x[i] = 0.0;
      y[i] = 0.0;
    }
    if (!std::isfinite(mass[i]) || mass[i] <= 0.0) mass[i] = 1.0;
  }
  std::fill(ax, ax + n, 0.0);
  std::fill(ay, ay + n, 0.0);
  if (phi) std::fill(phi, phi + n, 0.0);
  size_t S = num_moments(p);
  std::vector<double> mp(t.nodes.size() * S), loc(t.nodes.size() * S, 0.0);
  auto now = []() {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
  };
  double t0 = now();
  p2m(x, y, mass, t, p, mp.data());
  double t1 = now();
  m2m(mp.data(), t, p);
  double t2 = now();
  int cache_mode = algo == 0 ? (threads > 1 ? 2 : 0) : 1;
  if (algo == 1 && threads > 1) {
    // the CORRECTED variant has no shared cache; parallel over targets
    size_t Sx = S;
    const std::vector<M2lTerm> plan = m2l_plan(p);
    size_t dsz = (size_t)(2 * p + 1) * (2 * p + 1);
    long nn = (long)t.nodes.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (long id = 0; id < nn; ++id) {
      std::vector<double> deriv(dsz);
      const Node& tg = t.nodes[id];
      double* lc = loc.data() + (size_t)id * Sx;
      for (uint32_t s : m2l_l.data[id]) {
        const Node& sr = t.nodes[s];
        compute_kernel_derivs(tg.com_x - sr.com_x, tg.com_y - sr.com_y, eps, p, deriv.data());
        const double* sb = mp.data() + (size_t)s * Sx;
        for (const M2lTerm& tm : plan) {
          double mv = sb[tm.src];
          if (mv == 0.0) continue;
          lc[tm.tgt] += mv * tm.weight * deriv[tm.deriv];
        }
      }
    }
  } else {
    m2l(loc.data(), mp.data(), m2l_l, t, p, eps, cache_mode, threads, nullptr, UINT32_MAX);
  }
  double t3 = now();
  l2l(loc.data(), t, p, algo == 1);
  double t4 = now();
  l2p(x, y, ax, ay, phi, loc.data(), t, p, threads);
  double t5 = now();
  p2p(x, y, mass, ax, ay, phi, t, p2p_l, eps);
  double t6 = now();
  if (pass_times) {
    pass_times[0] = t1 - t0;
    pass_times[1] = t2 - t1;
    pass_times[2] = t3 - t2;
    pass_times[3] = t4 - t3;
    pass_times[4] = t5 - t4;
    pass_times[5] = t6 - t5;
  }
  return 0;
}

}  // namespace

// ===========================================================================
// C ABI of the oracle (ctypes from tests/ and bench.py's cpu_baseline)
// ===========================================================================

struct orc_tree {
  Tree t;
};
struct orc_lists {
  Lists l;
};

ORC_API uint64_t orc_morton_encode(double x, double y, double x0, double x1, double y0, double y1) {
  return morton_encode(x, y, x0, x1, y0, y1);
}

ORC_API void orc_morton_decode(uint64_t key, double x0, double x1, double y0, double y1, double* ox, double* oy) {
  *ox = denormalize(compact_bits(key), x0, x1);
  *oy = denormalize(compact_bits(key >> 1), y0, y1);
}

ORC_API int64_t orc_f64_as_i64(double v) { return rust_f64_as_i64(v); }

// Tree::build on caller arrays (permuted in place). Any of vx, vy, ax, ay may be
// NULL. perm (optional, n entries) receives the sorted order: new[i] = old[perm[i]].
ORC_API orc_tree* orc_tree_build(double* x, double* y, double* vx, double* vy, double* m, double* ax, double* ay,
                                 size_t n, uint64_t n_max, double x0, double x1, double y0, double y1,
                                 uint64_t* perm) {
  if (n_max == 0) return nullptr;
  Bodies b{x, y, vx, vy, m, ax, ay, n};
  std::vector<uint64_t> pv;
  orc_tree* h = new orc_tree;
  h->t = tree_build(b, n_max, x0, x1, y0, y1, perm ? &pv : nullptr);
  if (perm) std::copy(pv.begin(), pv.end(), perm);
  return h;
}

ORC_API orc_tree* orc_tree_build_constrained(double* x, double* y, double* m, size_t n, uint64_t n_max, double x0,
                                             double x1, double y0, double y1) {
  if (n_max == 0) return nullptr;
  Bodies b{x, y, nullptr, nullptr, m, nullptr, nullptr, n};
  orc_tree* h = new orc_tree;
  h->t = build_constrained(b, n_max, x0, x1, y0, y1);
  return h;
}

ORC_API size_t orc_tree_num_nodes(const orc_tree* h) { return h->t.nodes.size(); }

ORC_API void orc_tree_export(const orc_tree* h, double* cx, double* cy, double* hw, double* mass, double* comx,
                             double* comy, uint64_t* bstart, uint64_t* bend, int32_t* children, int32_t* parent,
                             uint8_t* is_leaf) {
  for (size_t i = 0; i < h->t.nodes.size(); ++i) {
    const Node& n = h->t.nodes[i];
    cx[i] = n.center_x; cy[i] = n.center_y; hw[i] = n.half_width; mass[i] = n.mass;
    comx[i] = n.com_x; comy[i] = n.com_y; bstart[i] = n.body_start; bend[i] = n.body_end;
    for (int k = 0; k < 4; ++k) children[i * 4 + k] = n.children[k];
    parent[i] = n.parent;
    is_leaf[i] = n.is_leaf;
  }
}

// Builds a tree handle from arrays (so a GPU-built tree can be checked by the
// oracle's own passes).
ORC_API orc_tree* orc_tree_import(size_t nn, const double* cx, const double* cy, const double* hw, const double* mass,
                                  const double* comx, const double* comy, const uint64_t* bstart,
                                  const uint64_t* bend, const int32_t* children, const int32_t* parent,
                                  const uint8_t* is_leaf, uint64_t n_max, double x0, double x1, double y0,
                                  double y1) {
  orc_tree* h = new orc_tree;
  h->t.n_max = n_max;
  h->t.x_min = x0; h->t.x_max = x1; h->t.y_min = y0; h->t.y_max = y1;
  h->t.nodes.resize(nn);
  for (size_t i = 0; i < nn; ++i) {
    Node& n = h->t.nodes[i];
    n.center_x = cx[i]; n.center_y = cy[i]; n.half_width = hw[i]; n.mass = mass[i];
    n.com_x = comx[i]; n.com_y = comy[i]; n.body_start = bstart[i]; n.body_end = bend[i];
    for (int k = 0; k < 4; ++k) n.children[k] = children[i * 4 + k];
    n.parent = parent[i];
    n.is_leaf = is_leaf[i];
  }
  return h;
}

ORC_API void orc_tree_free(orc_tree* h) { delete h; }

ORC_API void orc_tree_levels(const orc_tree* h, uint32_t* out) {
  std::vector<uint32_t> lv = compute_levels(h->t);
  std::copy(lv.begin(), lv.end(), out);
}

ORC_API void orc_neighbors_geometric(const orc_tree* h, size_t node, int64_t* out8) {
  neighbors_geometric(h->t, node, out8);
}

ORC_API orc_lists* orc_m2l_lists_build(const orc_tree* h) {
  orc_lists* l = new orc_lists;
  l->l = m2l_lists_build(h->t);
  return l;
}

ORC_API orc_lists* orc_p2p_lists_build(const orc_tree* h) {
  orc_lists* l = new orc_lists;
  l->l = p2p_lists_build(h->t);
  return l;
}

ORC_API void orc_corrected_lists_build(const orc_tree* h, const double* x, const double* y, double theta,
                                       orc_lists** m2l_out, orc_lists** p2p_out) {
  orc_lists* a = new orc_lists;
  orc_lists* b = new orc_lists;
  corrected_lists_build(h->t, x, y, theta, a->l, b->l);
  *m2l_out = a;
  *p2p_out = b;
}

ORC_API void orc_node_radii(const orc_tree* h, const double* x, const double* y, double* out) {
  std::vector<double> R = node_radii(h->t, x, y);
  std::copy(R.begin(), R.end(), out);
}

ORC_API size_t orc_lists_total(const orc_lists* l) {
  size_t s = 0;
  for (auto& v : l->l.data) s += v.size();
  return s;
}

// CSR export: offsets[nn+1], data[total]
ORC_API void orc_lists_export_csr(const orc_lists* l, uint64_t* offsets, uint32_t* data) {
  size_t s = 0;
  offsets[0] = 0;
  for (size_t i = 0; i < l->l.data.size(); ++i) {
    for (uint32_t v : l->l.data[i]) data[s++] = v;
    offsets[i + 1] = s;
  }
}

// CSR import (lists built elsewhere, e.g. exported from the device build)
ORC_API orc_lists* orc_lists_import_csr(size_t nn, const uint64_t* offsets, const uint32_t* data) {
  orc_lists* h = new orc_lists;
  h->l.data.resize(nn);
  for (size_t i = 0; i < nn; ++i) h->l.data[i].assign(data + offsets[i], data + offsets[i + 1]);
  return h;
}

ORC_API void orc_lists_free(orc_lists* l) { delete l; }

ORC_API void orc_kernel_derivs(double dx, double dy, double eps, int p, double* out) {
  compute_kernel_derivs(dx, dy, eps, p, out);
}

ORC_API void orc_p2m(const double* x, const double* y, const double* m, const orc_tree* h, int p, double* mp) {
  p2m(x, y, m, h->t, p, mp);
}

ORC_API void orc_m2m(double* mp, const orc_tree* h, int p) { m2m(mp, h->t, p); }

ORC_API void orc_m2l(double* loc, const double* mp, const orc_lists* l, const orc_tree* h, int p, double eps,
                     int cache_mode) {
  m2l(loc, mp, l->l, h->t, p, eps, cache_mode, 1, nullptr, UINT32_MAX);
}

ORC_API void orc_l2l(double* loc, const orc_tree* h, int p, int descending) { l2l(loc, h->t, p, descending != 0); }

ORC_API void orc_l2p(const double* x, const double* y, double* ax, double* ay, double* phi, const double* loc,
                     const orc_tree* h, int p) {
  l2p(x, y, ax, ay, phi, loc, h->t, p, 1);
}

ORC_API void orc_p2p(const double* x, const double* y, const double* m, double* ax, double* ay, double* phi,
                     const orc_tree* h, const orc_lists* l, double eps) {
  p2p(x, y, m, ax, ay, phi, h->t, l->l, eps);
}

// compute_fmm_force. threads > 1 runs the reference's three Rayon regions
// (M2L over targets, L2P over leaves) under OpenMP; pass_times (6 doubles,
// seconds) receives P2M, M2M, M2L, L2L, L2P, P2P wall times when non-NULL.
ORC_API void orc_compute_fmm_force(double* x, double* y, double* m, double* ax, double* ay, double* phi, size_t n,
                                   const orc_tree* h, const orc_lists* m2l_l, const orc_lists* p2p_l, int p,
                                   double eps, int algo, int threads, double* pass_times) {
  compute_fmm_force(x, y, m, ax, ay, phi, n, h->t, m2l_l->l, p2p_l->l, p, eps, algo, threads, pass_times);
}

// compute_n2_force — galax-core/src/lib.rs:1685-1721, restricted to the
// targets listed in idx (all bodies when idx == NULL). phi: -sum m_j/r.
ORC_API void orc_n2_force(const double* x, const double* y, const double* m, size_t n, double eps,
                          const uint64_t* idx, size_t n_targets, double* ax, double* ay, double* phi, int threads) {
  double soft2 = eps * eps;
  long nt = idx ? (long)n_targets : (long)n;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
  for (long k = 0; k < nt; ++k) {
    size_t i = idx ? idx[k] : (size_t)k;
    double axi = 0.0, ayi = 0.0, ph = 0.0, xi = x[i], yi = y[i];
    for (size_t j = 0; j < n; ++j) {
      if (i == j) continue;
      double dx = x[j] - xi, dy = y[j] - yi;
      double r2 = dx * dx + dy * dy + soft2;
      double inv_r = 1.0 / std::sqrt(r2);
      double inv_r3 = inv_r * inv_r * inv_r;
      double f = m[j] * inv_r3;
      axi += f * dx;
      ayi += f * dy;
      ph -= m[j] * inv_r;
    }
    ax[k] = axi;
    ay[k] = ayi;
    if (phi) phi[k] = ph;
  }
}

// Adjoint of compute_n2_force (no reference counterpart; the closed form of
// the derivative of the direct sum): for cotangents (axb, ayb),
//   xb_j = sum_{i!=j} K(x_i - x_j) (m_j a_i - m_i a_j),  K = d/dd [d / (|d|^2+eps^2)^{3/2}]
//   mb_j = -sum_{i!=j} a_i . (x_i - x_j) / (|x_i - x_j|^2 + eps^2)^{3/2}
ORC_API void orc_n2_vjp(const double* x, const double* y, const double* m, size_t n, double eps, const double* axb,
                        const double* ayb, double* xb, double* yb, double* mb, int threads) {
  double eps2 = eps * eps;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
  for (long jj = 0; jj < (long)n; ++jj) {
    size_t j = (size_t)jj;
    double gx = 0.0, gy = 0.0, gm = 0.0;
    for (size_t i = 0; i < n; ++i) {
      if (i == j) continue;
      double dx = x[i] - x[j], dy = y[i] - y[j];
      double r2 = dx * dx + dy * dy + eps2;
      double inv = 1.0 / std::sqrt(r2);
      double inv3 = inv * inv * inv, inv5 = inv3 * inv * inv;
      double vx = m[j] * axb[i] - m[i] * axb[j], vy = m[j] * ayb[i] - m[i] * ayb[j];
      double dv = dx * vx + dy * vy;
      gx += vx * inv3 - 3.0 * dx * dv * inv5;
      gy += vy * inv3 - 3.0 * dy * dv * inv5;
      gm -= (axb[i] * dx + ayb[i] * dy) * inv3;
    }
    xb[j] = gx;
    yb[j] = gy;
    mb[j] = gm;
  }
}

// Diagnostics — galax-core/src/lib.rs:1637-1683. out[6] = {KE, PE heuristic,
// px, py, Lz, -} computed sequentially in index order.
ORC_API void orc_diagnostics(const double* x, const double* y, const double* vx, const double* vy, const double* m,
                             const double* ax, const double* ay, size_t n, double* out) {
  double ke = 0.0, pe = 0.0, px = 0.0, py = 0.0, lz = 0.0;
  for (size_t i = 0; i < n; ++i) ke += 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i]);
  for (size_t i = 0; i < n; ++i) {
    double a = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
    double r = std::fmax(std::sqrt(x[i] * x[i] + y[i] * y[i]), 1.0);
    pe += m[i] * a * r;
  }
  for (size_t i = 0; i < n; ++i) {
    px += m[i] * vx[i];
    py += m[i] * vy[i];
  }
  for (size_t i = 0; i < n; ++i) lz += m[i] * (x[i] * vy[i] - y[i] * vx[i]);
  out[0] = ke; out[1] = -0.5 * pe; out[2] = px; out[3] = py; out[4] = lz; out[5] = 0.0;
}

// simulate / leapfrog_kdk — galax-integrate/src/lib.rs:158-237. The tree and
// lists stay frozen across steps (as in the reference). ax/ay must hold the
// current accelerations on entry.
ORC_API void orc_simulate(double* x, double* y, double* vx, double* vy, double* m, double* ax, double* ay, size_t n,
                          const orc_tree* h, const orc_lists* m2l_l, const orc_lists* p2p_l, int p, double eps,
                          double dt, uint64_t n_steps, int algo, int threads) {
  double half = 0.5 * dt;
  for (uint64_t s = 0; s < n_steps; ++s) {
    for (size_t i = 0; i < n; ++i) {
      vx[i] += ax[i] * half;
      vy[i] += ay[i] * half;
    }
    for (size_t i = 0; i < n; ++i) {
      x[i] += vx[i] * dt;
      y[i] += vy[i] * dt;
    }
    compute_fmm_force(x, y, m, ax, ay, nullptr, n, h->t, m2l_l->l, p2p_l->l, p, eps, algo, threads, nullptr);
    for (size_t i = 0; i < n; ++i) {
      vx[i] += ax[i] * half;
      vy[i] += ay[i] * half;
    }
  }
}

// galax-cli/src/main.rs:182-190: 0.1 x mean |b_i - b_{i+1 mod n}| over the
// first min(n,1000) bodies of the (Morton-sorted) array.
ORC_API double orc_default_softening(const double* x, const double* y, size_t n) {
  double sum = 0.0;
  size_t k = n < 1000 ? n : 1000;
  for (size_t i = 0; i < k; ++i) {
    double dx = x[i] - x[(i + 1) % n], dy = y[i] - y[(i + 1) % n];
    sum += std::sqrt(dx * dx + dy * dy);
  }
  return sum / (double)k * 0.1;
}

// galax-cli/src/main.rs:232-244
ORC_API double orc_default_dt(const double* ax, const double* ay, size_t n) {
  double mx = 0.0;
  for (size_t i = 0; i < n; ++i) mx = std::fmax(mx, std::sqrt(ax[i] * ax[i] + ay[i] * ay[i]));
  return mx > 1e-30 ? 0.5 / std::sqrt(mx) : 0.01;
}
