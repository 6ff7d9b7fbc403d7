"""ctypes front end of the CPU oracle (liboracle.so, built from galax_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP product. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.

Each wrapper follows the reference function named in its docstring
(knekvasil/galax, galax-core/src/lib.rs unless stated otherwise).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

ALGO_REFERENCE = 0
ALGO_CORRECTED = 1

_d = C.c_double
_u64 = C.c_uint64
_sz = C.c_size_t
_vp = C.c_void_p


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        build()
    L = C.CDLL(_SO)
    L.orc_morton_encode.restype = _u64
    L.orc_morton_encode.argtypes = [_d] * 6
    L.orc_morton_decode.argtypes = [_u64, _d, _d, _d, _d, _vp, _vp]
    L.orc_f64_as_i64.restype = C.c_int64
    L.orc_f64_as_i64.argtypes = [_d]
    L.orc_tree_build.restype = _vp
    L.orc_tree_build.argtypes = [_vp] * 7 + [_sz, _u64, _d, _d, _d, _d, _vp]
    L.orc_tree_build_constrained.restype = _vp
    L.orc_tree_build_constrained.argtypes = [_vp, _vp, _vp, _sz, _u64, _d, _d, _d, _d]
    L.orc_tree_num_nodes.restype = _sz
    L.orc_tree_num_nodes.argtypes = [_vp]
    L.orc_tree_export.argtypes = [_vp] * 12
    L.orc_tree_import.restype = _vp
    L.orc_tree_import.argtypes = [_sz] + [_vp] * 11 + [_u64, _d, _d, _d, _d]
    L.orc_tree_free.argtypes = [_vp]
    L.orc_tree_levels.argtypes = [_vp, _vp]
    L.orc_neighbors_geometric.argtypes = [_vp, _sz, _vp]
    L.orc_m2l_lists_build.restype = _vp
    L.orc_m2l_lists_build.argtypes = [_vp]
    L.orc_p2p_lists_build.restype = _vp
    L.orc_p2p_lists_build.argtypes = [_vp]
    L.orc_corrected_lists_build.argtypes = [_vp, _vp, _vp, _d, _vp, _vp]
    L.orc_node_radii.argtypes = [_vp, _vp, _vp, _vp]
    L.orc_lists_total.restype = _sz
    L.orc_lists_total.argtypes = [_vp]
    L.orc_lists_export_csr.argtypes = [_vp, _vp, _vp]
    L.orc_lists_free.argtypes = [_vp]
    L.orc_lists_import_csr.restype = _vp
    L.orc_lists_import_csr.argtypes = [_sz, _vp, _vp]
    L.orc_kernel_derivs.argtypes = [_d, _d, _d, C.c_int, _vp]
    L.orc_p2m.argtypes = [_vp, _vp, _vp, _vp, C.c_int, _vp]
    L.orc_m2m.argtypes = [_vp, _vp, C.c_int]
    L.orc_m2l.argtypes = [_vp, _vp, _vp, _vp, C.c_int, _d, C.c_int]
    L.orc_l2l.argtypes = [_vp, _vp, C.c_int, C.c_int]
    L.orc_l2p.argtypes = [_vp] * 7 + [C.c_int]
    L.orc_p2p.argtypes = [_vp] * 8 + [_d]
    L.orc_compute_fmm_force.argtypes = [_vp] * 6 + [_sz, _vp, _vp, _vp, C.c_int, _d, C.c_int, C.c_int, _vp]
    L.orc_n2_force.argtypes = [_vp, _vp, _vp, _sz, _d, _vp, _sz, _vp, _vp, _vp, C.c_int]
    L.orc_diagnostics.argtypes = [_vp] * 7 + [_sz, _vp]
    L.orc_n2_vjp.argtypes = [_vp, _vp, _vp, _sz, _d, _vp, _vp, _vp, _vp, _vp, C.c_int]
    L.orc_simulate.argtypes = [_vp] * 7 + [_sz, _vp, _vp, _vp, C.c_int, _d, _d, _u64, C.c_int, C.c_int]
    L.orc_default_softening.restype = _d
    L.orc_default_softening.argtypes = [_vp, _vp, _sz]
    L.orc_default_dt.restype = _d
    L.orc_default_dt.argtypes = [_vp, _vp, _sz]
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def num_moments(p: int) -> int:
    """galax-core/src/lib.rs:825-827"""
    return (p + 1) * (p + 2) // 2


def moment_index(i: int, j: int) -> int:
    """galax-core/src/lib.rs:830-833"""
    n = i + j
    return n * (n + 1) // 2 + j


def kernel_deriv_idx(k: int, l: int, p: int) -> int:
    """galax-core/src/lib.rs:1033-1036"""
    return (k + l) * (2 * p + 1) + l


def morton_encode(x, y, x0, x1, y0, y1) -> int:
    """galax-core/src/lib.rs:105-109"""
    return int(lib().orc_morton_encode(x, y, x0, x1, y0, y1))


def morton_decode(key, x0, x1, y0, y1):
    """galax-core/src/lib.rs:111-115"""
    ox, oy = C.c_double(), C.c_double()
    lib().orc_morton_decode(key, x0, x1, y0, y1, C.byref(ox), C.byref(oy))
    return ox.value, oy.value


class Tree:
    """Owns an oracle tree handle (galax-core Tree, :131-138)."""

    def __init__(self, handle, n_max, bounds):
        if not handle:
            raise ValueError("oracle tree build failed (n_max == 0?)")
        self.h = handle
        self.n_max = n_max
        self.bounds = tuple(bounds)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tree_free(self.h)
            self.h = None

    def __len__(self):
        return int(lib().orc_tree_num_nodes(self.h))

    def export(self) -> dict:
        nn = len(self)
        out = {k: np.zeros(nn) for k in ("center_x", "center_y", "half_width", "mass", "com_x", "com_y")}
        out["body_start"] = np.zeros(nn, np.uint64)
        out["body_end"] = np.zeros(nn, np.uint64)
        out["children"] = np.zeros((nn, 4), np.int32)
        out["parent"] = np.zeros(nn, np.int32)
        out["is_leaf"] = np.zeros(nn, np.uint8)
        lib().orc_tree_export(self.h, *[_p(out[k]) for k in ("center_x", "center_y", "half_width", "mass", "com_x",
                                                             "com_y", "body_start", "body_end", "children",
                                                             "parent", "is_leaf")])
        return out

    def levels(self) -> np.ndarray:
        """compute_levels, galax-core/src/lib.rs:426-438"""
        out = np.zeros(len(self), np.uint32)
        lib().orc_tree_levels(self.h, _p(out))
        return out

    @classmethod
    def from_arrays(cls, d: dict, n_max, bounds):
        nn = len(d["center_x"])
        arrs = [_f64(d[k]) for k in ("center_x", "center_y", "half_width", "mass", "com_x", "com_y")]
        bs = np.ascontiguousarray(d["body_start"], np.uint64)
        be = np.ascontiguousarray(d["body_end"], np.uint64)
        ch = np.ascontiguousarray(d["children"], np.int32)
        pa = np.ascontiguousarray(d["parent"], np.int32)
        il = np.ascontiguousarray(d["is_leaf"], np.uint8)
        h = lib().orc_tree_import(nn, *[_p(a) for a in arrs], _p(bs), _p(be), _p(ch), _p(pa), _p(il), n_max,
                                  *bounds)
        return cls(h, n_max, bounds)


class Lists:
    """Owns oracle interaction lists (galax-core InteractionLists, :941-944)."""

    def __init__(self, handle, n_nodes):
        self.h = handle
        self.n_nodes = n_nodes

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_lists_free(self.h)
            self.h = None

    def csr(self):
        tot = int(lib().orc_lists_total(self.h))
        off = np.zeros(self.n_nodes + 1, np.uint64)
        data = np.zeros(max(tot, 1), np.uint32)
        lib().orc_lists_export_csr(self.h, _p(off), _p(data))
        return off, data[:tot]

    @classmethod
    def from_csr(cls, offsets, data):
        off = np.ascontiguousarray(offsets, np.uint64)
        dat = np.ascontiguousarray(data if len(data) else np.zeros(1), np.uint32)
        return cls(lib().orc_lists_import_csr(len(off) - 1, _p(off), _p(dat)), len(off) - 1)

    def as_lists(self):
        off, data = self.csr()
        return [data[off[i]:off[i + 1]].tolist() for i in range(self.n_nodes)]


def tree_build(bodies: dict, n_max: int, bounds, return_perm=False):
    """Tree::build (galax-core/src/lib.rs:163-361). Permutes the float64 arrays of
    `bodies` (keys x, y, m and optionally vx, vy, ax, ay) in place."""
    n = len(bodies["x"])
    arrs = [bodies.get(k) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")]
    for a in arrs:
        if a is not None:
            assert a.dtype == np.float64 and a.flags.c_contiguous and len(a) == n
    perm = np.zeros(n, np.uint64)
    h = lib().orc_tree_build(*[_p(a) for a in arrs], n, n_max, *bounds, _p(perm))
    t = Tree(h, n_max, bounds)
    return (t, perm) if return_perm else t


def build_constrained(bodies: dict, n_max: int, bounds):
    """build_constrained (galax-core/src/lib.rs:541-588); no sort, current order."""
    n = len(bodies["x"])
    h = lib().orc_tree_build_constrained(_p(bodies["x"]), _p(bodies["y"]), _p(bodies["m"]), n, n_max, *bounds)
    return Tree(h, n_max, bounds)


def neighbors_geometric(tree: Tree, node: int):
    """neighbors_geometric (galax-core/src/lib.rs:382-423); -1 = None."""
    out = np.zeros(8, np.int64)
    lib().orc_neighbors_geometric(tree.h, node, _p(out))
    return out


def m2l_lists(tree: Tree) -> Lists:
    """InteractionLists::build (galax-core/src/lib.rs:947-1025)"""
    return Lists(lib().orc_m2l_lists_build(tree.h), len(tree))


def p2p_lists(tree: Tree) -> Lists:
    """build_p2p_lists (galax-core/src/lib.rs:1358-1392)"""
    return Lists(lib().orc_p2p_lists_build(tree.h), len(tree))


def corrected_lists(tree: Tree, x, y, theta=0.5):
    """ALGO_CORRECTED lists (no reference counterpart): dual tree traversal with
    the acceptance test R_A + R_B < theta |com_A - com_B|. Returns (m2l, p2p)."""
    a, b = C.c_void_p(), C.c_void_p()
    x, y = _f64(x), _f64(y)
    lib().orc_corrected_lists_build(tree.h, _p(x), _p(y), theta, C.byref(a), C.byref(b))
    return Lists(a.value, len(tree)), Lists(b.value, len(tree))


def node_radii(tree: Tree, x, y):
    out = np.zeros(len(tree))
    x, y = _f64(x), _f64(y)
    lib().orc_node_radii(tree.h, _p(x), _p(y), _p(out))
    return out


def kernel_derivs(dx, dy, eps, p) -> np.ndarray:
    """compute_kernel_derivs (galax-core/src/lib.rs:1040-1151); (2p+1)^2 buffer."""
    out = np.zeros((2 * p + 1) ** 2)
    lib().orc_kernel_derivs(dx, dy, eps, p, _p(out))
    return out


def p2m(x, y, m, tree: Tree, p) -> np.ndarray:
    """p2m (galax-core/src/lib.rs:848-887)"""
    out = np.zeros(len(tree) * num_moments(p))
    x, y, m = _f64(x), _f64(y), _f64(m)
    lib().orc_p2m(_p(x), _p(y), _p(m), tree.h, p, _p(out))
    return out


def m2m(mp, tree: Tree, p):
    """m2m (galax-core/src/lib.rs:890-938), in place"""
    lib().orc_m2m(_p(mp), tree.h, p)
    return mp


def m2l(locals_, mp, lists: Lists, tree: Tree, p, eps, cache_mode=0):
    """m2l (galax-core/src/lib.rs:1199-1259), in place on locals_"""
    lib().orc_m2l(_p(locals_), _p(mp), lists.h, tree.h, p, eps, cache_mode)
    return locals_


def l2l(locals_, tree: Tree, p, descending=False):
    """l2l (galax-core/src/lib.rs:1262-1302), in place"""
    lib().orc_l2l(_p(locals_), tree.h, p, int(descending))
    return locals_


def l2p(x, y, locals_, tree: Tree, p, ax=None, ay=None, phi=None):
    """l2p (galax-core/src/lib.rs:1308-1355); subtracts into ax, ay"""
    n = len(x)
    ax = np.zeros(n) if ax is None else ax
    ay = np.zeros(n) if ay is None else ay
    x, y = _f64(x), _f64(y)
    lib().orc_l2p(_p(x), _p(y), _p(ax), _p(ay), _p(phi), _p(locals_), tree.h, p)
    return ax, ay


def p2p(x, y, m, tree: Tree, lists: Lists, eps, ax=None, ay=None, phi=None):
    """p2p (galax-core/src/lib.rs:1395-1462); adds into ax, ay"""
    n = len(x)
    ax = np.zeros(n) if ax is None else ax
    ay = np.zeros(n) if ay is None else ay
    x, y, m = _f64(x), _f64(y), _f64(m)
    lib().orc_p2p(_p(x), _p(y), _p(m), _p(ax), _p(ay), _p(phi), tree.h, lists.h, eps)
    return ax, ay


def compute_fmm_force(x, y, m, tree: Tree, m2l_l: Lists, p2p_l: Lists, p, eps, algo=ALGO_REFERENCE,
                      threads=1, want_phi=False, pass_times=None):
    """compute_fmm_force (galax-core/src/lib.rs:1466-1499). Sanitizes copies of
    x, y, m (the reference mutates its input). Returns (ax, ay, phi|None)."""
    x = np.array(x, np.float64)
    y = np.array(y, np.float64)
    m = np.array(m, np.float64)
    n = len(x)
    ax, ay = np.zeros(n), np.zeros(n)
    phi = np.zeros(n) if want_phi else None
    pt = np.zeros(8) if pass_times is not None else None
    lib().orc_compute_fmm_force(_p(x), _p(y), _p(m), _p(ax), _p(ay), _p(phi), n, tree.h, m2l_l.h, p2p_l.h, p, eps,
                                algo, threads, _p(pt))
    if pass_times is not None:
        pass_times[:] = pt[:6]
    return ax, ay, phi


def n2_force(x, y, m, eps, targets=None, threads=1, want_phi=False):
    """compute_n2_force (galax-core/src/lib.rs:1685-1721) on a subset of targets."""
    x, y, m = _f64(x), _f64(y), _f64(m)
    n = len(x)
    idx = None if targets is None else np.ascontiguousarray(targets, np.uint64)
    nt = n if idx is None else len(idx)
    ax, ay = np.zeros(nt), np.zeros(nt)
    phi = np.zeros(nt) if want_phi else None
    lib().orc_n2_force(_p(x), _p(y), _p(m), n, eps, _p(idx), nt, _p(ax), _p(ay), _p(phi), threads)
    return ax, ay, phi


def n2_vjp(x, y, m, eps, axb, ayb, threads=1):
    """Closed-form adjoint of the direct sum: returns (x_bar, y_bar, m_bar)."""
    x, y, m, axb, ayb = (_f64(a) for a in (x, y, m, axb, ayb))
    n = len(x)
    xb, yb, mb = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_n2_vjp(_p(x), _p(y), _p(m), n, eps, _p(axb), _p(ayb), _p(xb), _p(yb), _p(mb), threads)
    return xb, yb, mb


def diagnostics(x, y, vx, vy, m, ax, ay):
    """compute_kinetic_energy, compute_fmm_potential_energy, compute_momentum,
    compute_angular_momentum (galax-core/src/lib.rs:1637-1683)."""
    out = np.zeros(6)
    arrs = [_f64(a) for a in (x, y, vx, vy, m, ax, ay)]
    lib().orc_diagnostics(*[_p(a) for a in arrs], len(x), _p(out))
    return {"kinetic": out[0], "potential": out[1], "momentum_x": out[2], "momentum_y": out[3],
            "angular_momentum": out[4]}


def simulate(b: dict, tree: Tree, m2l_l: Lists, p2p_l: Lists, p, eps, dt, n_steps, algo=ALGO_REFERENCE,
             threads=1):
    """simulate without diagnostics (galax-integrate/src/lib.rs:158-237), in place."""
    n = len(b["x"])
    lib().orc_simulate(*[_p(b[k]) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")], n, tree.h, m2l_l.h, p2p_l.h,
                       p, eps, dt, n_steps, algo, threads)


def default_softening(x, y) -> float:
    """galax-cli/src/main.rs:182-190"""
    x, y = _f64(x), _f64(y)
    return float(lib().orc_default_softening(_p(x), _p(y), len(x)))


def default_dt(ax, ay) -> float:
    """galax-cli/src/main.rs:232-244"""
    ax, ay = _f64(ax), _f64(ay)
    return float(lib().orc_default_dt(_p(ax), _p(ay), len(ax)))
