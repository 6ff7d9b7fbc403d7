/* galax_fmm.h — C ABI of the MI355X (gfx950) FMM force backend.
 *
 * Drop-in for galax's force backends (knekvasil/galax):
 *   galax-core  compute_fmm_force(&mut BodiesSoA, &Tree, &InteractionLists m2l,
 *               &InteractionLists p2p, p, softening)   galax-core/src/lib.rs:1466-1499
 *   galax-gpu   GpuContext::{new, upload_tree_data, step, read_accelerations,
 *               crosscheck, n2_debug}                  galax-gpu/src/lib.rs:230-718
 * Bodies in (structure of arrays, caller-owned), accelerations out, with the
 * same p, n_max, softening and tree bounds. Potentials are an extra output.
 *
 * Conventions: every entry point returns 0 on success or a negative
 * galax_status_t; nothing aborts or throws across the ABI. The text of the
 * last error on the calling thread is galax_fmm_last_error(). A handle owns
 * one HIP stream and its device buffers; calls on one handle are not
 * reentrant. Host-array entry points are synchronous on return.
 * Only C types appear here, so a Rust `extern "C"` block (repr(C) structs)
 * binds it directly.
 */
#ifndef GALAX_FMM_H
#define GALAX_FMM_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define GALAX_API __attribute__((visibility("default")))
#else
#define GALAX_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes; the reference's GpuError variants (galax-gpu/src/lib.rs:42-59)
 * and its panics (map failure :556, crosscheck threshold :663-665) map here. */
typedef enum {
  GALAX_OK = 0,
  GALAX_ERR_INIT = -1,        /* GpuError::InitFailed: no HIP device / stream */
  GALAX_ERR_COMPILE = -2,     /* GpuError::ShaderCompilation: no gfx950 code object */
  GALAX_ERR_ALLOC = -3,       /* GpuError::BufferAllocation */
  GALAX_ERR_INVALID = -4,     /* bad argument, size above capacity, no tree yet */
  GALAX_ERR_DEVICE = -5,      /* a HIP call or kernel failed */
  GALAX_ERR_UNSUPPORTED = -6, /* a combination this build does not provide */
  GALAX_ERR_THRESHOLD = -7,   /* crosscheck error above its threshold */
  GALAX_ERR_IO = -8,          /* SnapshotError::Io (galax-io/src/lib.rs:10-11) */
  GALAX_ERR_FORMAT = -9       /* SnapshotError::BadMagic / UnexpectedEof (galax-io/src/lib.rs:12-17) */
} galax_status_t;

/* Algorithm variants (same tree, same kernels). */
enum {
  /* galax-core's semantics: 64-capped id-ordered lists, P2P pair rule
   * `partner > node`, M2L derivative cache keyed by (dx as i64, dy as i64)
   * with sequential first-touch representatives, ascending-id L2L. */
  GALAX_ALGO_REFERENCE = 0,
  /* dual-tree-traversal lists with acceptance test R_A+R_B < theta*d, exact
   * per-pair derivatives, top-down L2L (no counterpart in the reference). */
  GALAX_ALGO_CORRECTED = 1
};

enum {
  GALAX_TREE_PLAIN = 0,      /* Tree::build: stable Morton sort + n_max chunks (lib.rs:163-361) */
  GALAX_TREE_CONSTRAINED = 1 /* build_constrained on the current order (lib.rs:541-588) */
};

enum {
  GALAX_PREC_FP64 = 0,          /* everything in FP64; M2L on FP64 MFMA (grouped by cache key) */
  GALAX_PREC_MIXED_FP32 = 1,    /* M2L products on FP32 MFMA, P2P in FP32 on leaf-relative coordinates;
                                   expansions, geometry and sums in FP64 (GALAX_ALGO_CORRECTED, p <= 8:
                                   its per-pair M2L in FP32 on pairs rescaled by powers of two) */
  GALAX_PREC_MIXED_BF16X3 = 2,  /* as MIXED_FP32 with M2L as a 3-term split BF16 MFMA product */
  GALAX_PREC_FP64_EXACT = 3     /* FP64 with galax-core's M2L summation order (per pair, on the VALU):
                                   expansions bit-identical to the reference's arithmetic */
};

/* Mirrors GpuConfig (galax-gpu/src/lib.rs:22-40) plus the tree bounds that
 * galax-cli hard-codes (galax-cli/src/main.rs:159) and the variant selectors. */
typedef struct {
  uint32_t p;          /* expansion order, 0..16 (MAX_P = 17, lib.rs:6) */
  uint32_t n_max;      /* bodies per leaf */
  double softening;    /* eps in 1/sqrt(r^2+eps^2) */
  double x_min, x_max, y_min, y_max; /* Morton bounds */
  uint32_t tree_mode;  /* GALAX_TREE_* */
  uint32_t algo;       /* GALAX_ALGO_* */
  uint32_t precision;  /* GALAX_PREC_* */
  double theta;        /* acceptance parameter of GALAX_ALGO_CORRECTED (galax-cli --theta) */
} galax_fmm_config_t;

/* Caller-owned structure of arrays; mirrors BodiesSoA (lib.rs:41-49) plus phi.
 * Pointers not needed by an entry point may be NULL (phi == NULL: not written). */
typedef struct {
  double *x, *y, *vx, *vy, *m, *ax, *ay, *phi;
  size_t n;
} galax_bodies_t;

/* Caller-owned view of a tree in galax-core's node order (Node, lib.rs:117-129).
 * Arrays have n_nodes entries (children: 4 per node, -1 = None; parent -1 = None). */
typedef struct {
  size_t n_nodes;
  double *center_x, *center_y, *half_width, *mass, *com_x, *com_y;
  uint64_t *body_start, *body_end;
  int32_t *children;
  int32_t *parent;
  uint8_t *is_leaf;
} galax_tree_view_t;

/* Interaction lists as CSR: partners of node i are data[offsets[i] .. offsets[i+1]),
 * in the reference's order (InteractionLists{data:[u32;64], lengths:u8}, lib.rs:941-944). */
typedef struct {
  size_t n_nodes;
  uint64_t *offsets; /* n_nodes + 1 */
  uint32_t *data;    /* offsets[n_nodes] */
} galax_lists_view_t;

/* Energy / momentum diagnostics (galax-integrate/src/lib.rs:192-199). potential
 * is 1/2 sum_i m_i phi_i with phi_i = -sum_{j != i} m_j / sqrt(r_ij^2 + eps^2)
 * (direct_potential's sign, galax-core/src/lib.rs:1627-1635) from the FMM
 * potentials; potential_heuristic is galax-core's compute_fmm_potential_energy
 * (lib.rs:1671-1683). */
typedef struct {
  uint64_t step;
  double kinetic, potential, potential_heuristic, momentum_x, momentum_y, angular_momentum;
} galax_diag_t;

typedef struct galax_fmm galax_fmm_t;

/* Per-pass device times of the last evaluate, milliseconds (HIP events on the
 * handle's stream). The sanitize of compute_fmm_force runs inside the P2M
 * kernel: its time is part of p2m and `sanitize` is the gap before it. */
typedef struct {
  double sanitize, p2m, m2m, m2l_setup, m2l, l2l, l2p_p2p, total;
  /* m2l_pairs_reference: every M2L pair of the reference's lists, 0 while
   * the rows of levels >= 2 are not built (see galax_fmm_tree_info) */
  uint64_t m2l_pairs_executed, m2l_pairs_reference, m2l_keys, p2p_interactions;
  /* matrix-core M2L work units: (target, key) groups (runs of more than 4
   * pairs are pre-summed into one row each evaluation),
   * 16-column products (steps) and work items; 0 for FP64_EXACT / CORRECTED */
  uint64_t m2l_groups, m2l_steps, m2l_items;
  /* the main kernel of the M2L and of the L2P+P2P pass alone (no combine) */
  double m2l_kernel, p2p_kernel;
} galax_timings_t;

/* Device times of the last tree build, milliseconds (HIP events on the
 * handle's stream): Morton keys + radix sort, permutation of the body arrays,
 * leaves + parent rounds, the reference M2L/P2P lists, and the gather lists +
 * P2P work items. sort and permute are 0 for GALAX_TREE_CONSTRAINED. */
typedef struct {
  double sort, permute, tree, lists, gather, total;
} galax_build_timings_t;

/* ~GpuContext::new (galax-gpu/src/lib.rs:268-390): validates the config,
 * selects the device current on the calling thread, creates the stream and
 * preallocates body buffers for max_n bodies. */
GALAX_API int galax_fmm_create(const galax_fmm_config_t* cfg, size_t max_n, galax_fmm_t** out);

/* ~Tree::build + InteractionLists::build + build_p2p_lists
 * (galax-core/src/lib.rs:163-361, 947-1025, 1358-1392; galax-cli/src/main.rs:159-193):
 * builds the tree and both lists on the device. Permutes all non-NULL arrays of
 * `bodies` in place exactly as Tree::build does (stable Morton order; no sort in
 * GALAX_TREE_CONSTRAINED). perm_out (optional, n entries): new[i] = old[perm_out[i]].
 * The tree is then static, as in the reference (built once, galax-cli/src/main.rs:159). */
GALAX_API int galax_fmm_build_tree(galax_fmm_t* h, galax_bodies_t* bodies, uint64_t* perm_out);

/* As galax_fmm_build_tree with device-resident, already-allocated arrays of
 * n bodies (x, y, m required; vx, vy, ax, ay optional) permuted in place. */
GALAX_API int galax_fmm_build_tree_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_vx, double* d_vy, double* d_m,
                                double* d_ax, double* d_ay, size_t n);

/* ~GpuContext::upload_tree_data (galax-gpu/src/lib.rs:392-462): installs a tree
 * and its two lists built elsewhere (e.g. by galax-core). Node ids must be
 * level-ordered with consecutive children, as Tree::build produces. */
GALAX_API int galax_fmm_upload_tree(galax_fmm_t* h, const galax_tree_view_t* tree, const galax_lists_view_t* m2l,
                          const galax_lists_view_t* p2p);

/* Changes the softening of an existing context (galax-cli derives its default
 * from the Morton-sorted bodies after Tree::build, galax-cli/src/main.rs:182-190). */
GALAX_API int galax_fmm_set_softening(galax_fmm_t* h, double softening);

/* Sizes of the installed tree: nodes, leaves, levels, M2L and P2P list lengths.
 * galax_fmm_build_tree walks the M2L rows of levels 0 and 1 only, the rows
 * an evaluation reads; asking for m2l_total (or exporting the M2L lists)
 * first builds the rows of the other levels, once per tree. */
GALAX_API int galax_fmm_tree_info(galax_fmm_t* h, size_t* n_nodes, size_t* n_leaves, uint32_t* n_levels, size_t* m2l_total,
                        size_t* p2p_total);

/* Copies the installed tree and lists out (for bit-exactness checks).
 * Views must be sized from galax_fmm_tree_info; any view may be NULL. An M2L
 * view gets every row (the rows of levels >= 2 are built first if needed). */
GALAX_API int galax_fmm_export_tree(galax_fmm_t* h, galax_tree_view_t* tree, galax_lists_view_t* m2l,
                          galax_lists_view_t* p2p);

/* No reference counterpart: copies the multipole and local coefficients of
 * the last evaluation out for checks, n_nodes x S doubles each (node-major,
 * galax-core's coefficient order, S = (p+1)(p+2)/2). Only the nodes an
 * evaluation uses hold values: with GALAX_ALGO_REFERENCE the multipoles of
 * levels 0-1 and the M2L locals of levels 0-1 (the leaves' share of their
 * parents' locals is added inside the force kernel, not stored); with
 * GALAX_ALGO_CORRECTED every node, the locals after L2L except at the leaves.
 * Either pointer may be NULL. */
GALAX_API int galax_fmm_export_expansions(galax_fmm_t* h, double* multipoles, double* locals, size_t n_nodes);

/* No reference counterpart (GALAX_ALGO_CORRECTED has none): copies the
 * corrected variant's lists out for checks. They are built on the device by a
 * dual tree traversal at the first evaluation after a tree build. M2L: each
 * node's accepted partners, ascending; P2P: each leaf itself, then its
 * unaccepted leaf partners, ascending (internal nodes empty). Views hold
 * n_nodes rows; a view with data == NULL receives its offsets only (the
 * total is offsets[n_nodes]). Either view may be NULL. */
GALAX_API int galax_fmm_export_corrected_lists(galax_fmm_t* h, galax_lists_view_t* m2l, galax_lists_view_t* p2p);

/* ~compute_fmm_force (galax-core/src/lib.rs:1466-1499) and
 * GpuContext::step + read_accelerations (galax-gpu/src/lib.rs:464-566):
 * reads x, y, m in the tree's body order, sanitizes them like :1475-1482
 * (writing the sanitized values back to the caller's arrays when any changed),
 * writes ax, ay and, when non-NULL, phi. Synchronous. */
GALAX_API int galax_fmm_evaluate(galax_fmm_t* h, galax_bodies_t* bodies);

/* ~GpuContext::crosscheck (galax-gpu/src/lib.rs:623-667): evaluates the
 * bodies in the context's precision form, then in GALAX_PREC_FP64_EXACT
 * (galax-core's arithmetic: expansions bit-identical to compute_fmm_force's)
 * on the same tree, and compares them per body as the reference does:
 * |a_form - a_exact| / max(|a_exact|, |a_form|, 1e-30). Writes the
 * FP64_EXACT accelerations to ax, ay (the reference leaves the CPU result in
 * bodies.ax/ay) and max, mean and count to the optional outputs. Returns
 * GALAX_ERR_THRESHOLD (outputs written) where the reference panics: when the
 * maximum exceeds threshold. */
GALAX_API int galax_fmm_crosscheck(galax_fmm_t* h, galax_bodies_t* bodies, double threshold, double* max_rel,
                                   double* mean_rel, size_t* count);

/* As galax_fmm_evaluate on device-resident arrays of the installed tree's n
 * bodies, enqueued on the handle's stream and synchronised before return.
 * d_phi may be NULL. Inputs are sanitized in place. */
GALAX_API int galax_fmm_evaluate_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_m, double* d_ax, double* d_ay,
                              double* d_phi);

/* Device times and work counts of the last evaluate. No reference counterpart
 * (galax-bench times whole calls on the host, galax-bench/src/main.rs:7-17).
 * The times are 0 when timing is off (galax_fmm_set_timing). */
GALAX_API int galax_fmm_last_timings(galax_fmm_t* h, galax_timings_t* out);

/* Turns the per-pass timing events of evaluate on (the default) or off. Each
 * event separates two passes on the stream and costs a few microseconds of
 * device time; a caller that does not read galax_fmm_last_timings can turn
 * them off. No reference counterpart. */
GALAX_API int galax_fmm_set_timing(galax_fmm_t* h, int enable);

/* Device times of the last galax_fmm_build_tree[_device]. No reference
 * counterpart (as above). */
GALAX_API int galax_fmm_last_build_timings(galax_fmm_t* h, galax_build_timings_t* out);

/* ~galax-integrate simulate / leapfrog_kdk_gpu (galax-integrate/src/lib.rs:158-237,
 * 292-368): n_steps kick-drift-kick steps on the device (positions and
 * velocities in FP64, tree static). ax, ay must hold the current
 * accelerations on entry and hold the final ones on return. When diag != NULL
 * and every > 0, diagnostics after every `every` steps are written to
 * diag[0 .. n_steps/every). */
GALAX_API int galax_fmm_kdk(galax_fmm_t* h, galax_bodies_t* bodies, double dt, uint64_t n_steps, uint64_t every,
                  galax_diag_t* diag);

/* Vector-Jacobian product of the accelerations: given cotangents ax_bar, ay_bar
 * (tree body order), returns x_bar, y_bar, m_bar = J^T (ax_bar, ay_bar) for
 * the map (x, y, m) -> (ax, ay) on the installed static tree (topology, lists,
 * M2L keys and node centres frozen, as the reference freezes them after
 * Tree::build). Recomputes the forward expansions first, in the context's
 * precision form. The adjoint passes run in FP64, except the near field of the
 * mixed forms, which runs in FP32 on leaf-relative coordinates like their
 * forward near field. No reference counterpart; for differentiable simulation
 * and for checking through torch.autograd. */
GALAX_API int galax_fmm_vjp(galax_fmm_t* h, const galax_bodies_t* bodies, const double* ax_bar, const double* ay_bar,
                            double* x_bar, double* y_bar, double* m_bar);

/* As galax_fmm_vjp on device arrays (x, y, m are sanitized in place). */
GALAX_API int galax_fmm_vjp_device(galax_fmm_t* h, double* d_x, double* d_y, double* d_m, const double* d_ax_bar,
                                   const double* d_ay_bar, double* d_x_bar, double* d_y_bar, double* d_m_bar);

/* ~compute_n2_force (galax-core/src/lib.rs:1685-1721) / GpuContext::n2_debug
 * (galax-gpu/src/lib.rs:670-714) in FP64 on the device, for the targets listed
 * in targets (all bodies when NULL). */
GALAX_API int galax_direct_sum(const double* x, const double* y, const double* m, size_t n, double softening,
                     const uint64_t* targets, size_t n_targets, double* ax, double* ay, double* phi);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): galax_direct_sum
 * without any allocation. Its device memory (staging for the inputs and
 * outputs, partial sums) is a fixed array of 64 MiB inside the library, one
 * per GPU, held until the process ends; galax_direct_sum_noalloc_bytes(n,
 * n_targets) is what a call needs, and a call needing more is rejected with
 * GALAX_ERR_INVALID (e.g. n = 2^20 with 10,000 targets fits; all n targets
 * fit up to ~600,000 bodies). Runs on the calling thread's per-thread
 * default stream of the current device and returns synchronised. Results
 * are bit-identical to galax_direct_sum's. All calls on one GPU share the
 * array: calls that overlap in time on one GPU (from several threads) race. */
GALAX_API int galax_direct_sum_noalloc(const double* x, const double* y, const double* m, size_t n, double softening,
                                       const uint64_t* targets, size_t n_targets, double* ax, double* ay, double* phi);
GALAX_API size_t galax_direct_sum_noalloc_bytes(size_t n, size_t n_targets);

/* ~validate_fmm (galax-core/src/lib.rs:1503-1564) with caller-chosen sample
 * indices: evaluates the FMM, computes the direct sum on the device for the
 * samples and returns (max, mean) of | |a_fmm| - |a_exact| | / max(|a_exact|, 1e-20)
 * plus the RMS of the vector error |a_fmm - a_exact| / |a_exact|. */
GALAX_API int galax_fmm_validate(galax_fmm_t* h, galax_bodies_t* bodies, const uint64_t* samples, size_t n_samples,
                       double* max_rel, double* mean_rel, double* rms_vec);

GALAX_API void galax_fmm_destroy(galax_fmm_t* h);

/* ~galax-io write_snapshot (galax-io/src/lib.rs:21-42): a GALX v1 file,
 * little-endian: "GALX", u32 version 1, u64 n, f64 timestamp, then n f64 each
 * of x, y, vx, vy, mass, ax, ay. The reference stamps the current time; here
 * the caller passes it. NULL arrays are written as zeros (phi is not stored). */
GALAX_API int galax_snapshot_write(const char* path, const galax_bodies_t* bodies, double timestamp);

/* The header of a GALX file: body count and timestamp (either may be NULL).
 * Errors as galax_snapshot_read. */
GALAX_API int galax_snapshot_info(const char* path, size_t* n, double* timestamp);

/* ~galax-io read_snapshot (galax-io/src/lib.rs:53-84): fills caller arrays of
 * bodies->n entries, which must equal the file's count (galax_snapshot_info);
 * NULL arrays are skipped, phi is untouched. GALAX_ERR_FORMAT for a wrong magic
 * or a short file, GALAX_ERR_UNSUPPORTED for a version other than 1,
 * GALAX_ERR_IO when the file cannot be opened or read. */
GALAX_API int galax_snapshot_read(const char* path, galax_bodies_t* bodies);

/* Text of the last error raised on the calling thread ("" if none). */
GALAX_API const char* galax_fmm_last_error(void);

/* Library version string. */
GALAX_API const char* galax_fmm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GALAX_FMM_H */
