"""Python front end of libgalax_fmm.so (the gfx950 FMM force backend).

Mirrors the reference's force-backend surface (knekvasil/galax):
  * `BodiesSoA`                 galax-core/src/lib.rs:41-71 (plus `phi`)
  * `GpuConfig`, `GpuContext`   galax-gpu/src/lib.rs:22-40, 230-718
      new -> GpuContext(cfg); upload_tree_data; step; read_accelerations;
      crosscheck; n2_debug
  * `build_tree`                Tree::build / build_constrained + both list
                                builders (galax-core/src/lib.rs:163-361, 541-588,
                                947-1025, 1358-1392), on the device
  * `compute_fmm_force`         galax-core/src/lib.rs:1466-1499
  * `simulate`                  galax-integrate/src/lib.rs:204-237 (on the device)
  * `compute_n2_force`          galax-core/src/lib.rs:1685-1721 (on the device)
  * `validate_fmm`              galax-core/src/lib.rs:1503-1564

Everything runs in libgalax_fmm.so; this module only marshals arrays. It
raises GalaxError when the library is missing or a call fails: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgalax_fmm.so")

ALGO_REFERENCE, ALGO_CORRECTED = 0, 1
TREE_PLAIN, TREE_CONSTRAINED = 0, 1
PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16X3, PREC_FP64_EXACT = 0, 1, 2, 3

STATUS = {0: "OK", -1: "InitFailed", -2: "ShaderCompilation", -3: "BufferAllocation", -4: "InvalidArgument",
          -5: "DeviceError", -6: "Unsupported", -7: "ThresholdExceeded", -8: "Io", -9: "Format"}


class GalaxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("p", C.c_uint32), ("n_max", C.c_uint32), ("softening", C.c_double), ("x_min", C.c_double),
                ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double), ("tree_mode", C.c_uint32),
                ("algo", C.c_uint32), ("precision", C.c_uint32), ("theta", C.c_double)]


class _Bodies(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "vx", "vy", "m", "ax", "ay", "phi")] + [("n", C.c_size_t)]


class _TreeView(C.Structure):
    _fields_ = [("n_nodes", C.c_size_t)] + [(k, C.c_void_p) for k in (
        "center_x", "center_y", "half_width", "mass", "com_x", "com_y", "body_start", "body_end", "children",
        "parent", "is_leaf")]


class _ListsView(C.Structure):
    _fields_ = [("n_nodes", C.c_size_t), ("offsets", C.c_void_p), ("data", C.c_void_p)]


class _Diag(C.Structure):
    _fields_ = [("step", C.c_uint64)] + [(k, C.c_double) for k in (
        "kinetic", "potential", "potential_heuristic", "momentum_x", "momentum_y", "angular_momentum")]


class _Timings(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("sanitize", "p2m", "m2m", "m2l_setup", "m2l", "l2l", "l2p_p2p",
                                          "total")] + \
               [(k, C.c_uint64) for k in ("m2l_pairs_executed", "m2l_pairs_reference", "m2l_keys",
                                          "p2p_interactions", "m2l_groups", "m2l_steps", "m2l_items")] + \
               [(k, C.c_double) for k in ("m2l_kernel", "p2p_kernel")]


class _BuildTimings(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("sort", "permute", "tree", "lists", "gather", "total")]


_lib = None


def lib():
    """Loads libgalax_fmm.so (built by galax_amd.build); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GalaxError(-2, f"{LIB_PATH} is missing: build it with `python -m galax_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, sz, d = C.c_void_p, C.c_size_t, C.c_double
    sig = {
        "galax_fmm_create": [C.POINTER(_Config), sz, C.POINTER(vp)],
        "galax_fmm_build_tree": [vp, C.POINTER(_Bodies), vp],
        "galax_fmm_build_tree_device": [vp] + [vp] * 7 + [sz],
        "galax_fmm_upload_tree": [vp, C.POINTER(_TreeView), C.POINTER(_ListsView), C.POINTER(_ListsView)],
        "galax_fmm_tree_info": [vp, vp, vp, vp, vp, vp],
        "galax_fmm_set_softening": [vp, d],
        "galax_fmm_export_tree": [vp, vp, vp, vp],
        "galax_fmm_export_corrected_lists": [vp, vp, vp],
        "galax_fmm_export_expansions": [vp, vp, vp, C.c_size_t],
        "galax_fmm_evaluate": [vp, C.POINTER(_Bodies)],
        "galax_fmm_evaluate_device": [vp] * 7,
        "galax_fmm_crosscheck": [vp, C.POINTER(_Bodies), d, vp, vp, vp],
        "galax_fmm_last_timings": [vp, C.POINTER(_Timings)],
        "galax_fmm_last_build_timings": [vp, C.POINTER(_BuildTimings)],
        "galax_fmm_set_timing": [vp, C.c_int],
        "galax_fmm_kdk": [vp, C.POINTER(_Bodies), d, C.c_uint64, C.c_uint64, vp],
        "galax_fmm_vjp": [vp, C.POINTER(_Bodies), vp, vp, vp, vp, vp],
        "galax_fmm_vjp_device": [vp] * 9,
        "galax_direct_sum": [vp, vp, vp, sz, d, vp, sz, vp, vp, vp],
        "galax_fmm_validate": [vp, C.POINTER(_Bodies), vp, sz, vp, vp, vp],
        "galax_snapshot_write": [C.c_char_p, C.POINTER(_Bodies), d],
        "galax_snapshot_info": [C.c_char_p, vp, vp],
        "galax_snapshot_read": [C.c_char_p, C.POINTER(_Bodies)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.galax_fmm_destroy.argtypes = [vp]
    L.galax_fmm_destroy.restype = None
    L.galax_fmm_last_error.restype = C.c_char_p
    L.galax_fmm_version.restype = C.c_char_p
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise GalaxError(rc, lib().galax_fmm_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a, n=None):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("arrays must be C-contiguous float64 numpy arrays")
    if n is not None and len(a) != n:
        raise ValueError(f"array length {len(a)} != n = {n}")
    return a


@dataclass
class BodiesSoA:
    """galax-core BodiesSoA (lib.rs:41-71) plus the potential `phi`."""
    x: np.ndarray
    y: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    mass: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    phi: np.ndarray | None = None

    @classmethod
    def new(cls, n: int, with_phi=True):
        z = lambda: np.zeros(n)
        return cls(z(), z(), z(), z(), z(), z(), z(), z() if with_phi else None)

    @classmethod
    def from_dict(cls, d: dict, with_phi=True):
        n = len(d["x"])
        g = lambda k: np.array(d[k], np.float64, copy=True) if k in d else np.zeros(n)
        return cls(g("x"), g("y"), g("vx"), g("vy"), g("m"), g("ax"), g("ay"), np.zeros(n) if with_phi else None)

    def __len__(self):
        return len(self.x)

    def _c(self) -> _Bodies:
        n = len(self.x)
        arrs = [self.x, self.y, self.vx, self.vy, self.mass, self.ax, self.ay, self.phi]
        for a in arrs:
            _f64(a, n)
        return _Bodies(*[_ptr(a) for a in arrs], n)


@dataclass
class GpuConfig:
    """galax-gpu GpuConfig (lib.rs:22-40) plus tree bounds and variant selectors."""
    max_n: int = 262_144
    p: int = 4
    eps: float = 0.1
    n_max: int = 32
    bounds: tuple = (-50.0, 50.0, -50.0, 50.0)
    tree_mode: int = TREE_PLAIN
    algo: int = ALGO_REFERENCE
    precision: int = PREC_FP64
    theta: float = 0.5

    def _c(self) -> _Config:
        x0, x1, y0, y1 = self.bounds
        return _Config(self.p, self.n_max, self.eps, x0, x1, y0, y1, self.tree_mode, self.algo, self.precision,
                       self.theta)


@dataclass
class Timings:
    passes_ms: dict = field(default_factory=dict)
    m2l_pairs_executed: int = 0
    m2l_pairs_reference: int = 0
    m2l_keys: int = 0
    p2p_interactions: int = 0
    m2l_groups: int = 0
    m2l_steps: int = 0
    m2l_items: int = 0


class GpuContext:
    """One device context (galax-gpu GpuContext, lib.rs:230-718): owns a HIP
    stream, device buffers sized for cfg.max_n bodies and, once built or
    uploaded, a static tree with its interaction lists."""

    def __init__(self, cfg: GpuConfig):
        self.cfg = cfg
        h = C.c_void_p()
        c = cfg._c()
        _check(lib().galax_fmm_create(C.byref(c), cfg.max_n, C.byref(h)))
        self.h = h
        self._acc = None
        self._n = None  # bodies of the installed tree
        self._device = None  # torch device index of the handle (the current one at creation)
        try:
            import torch
            if torch.cuda.is_available():
                self._device = torch.cuda.current_device()
        except ImportError:
            pass

    def close(self):
        if getattr(self, "h", None):
            lib().galax_fmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- tree ---------------------------------------------------------------
    def build_tree(self, b: BodiesSoA) -> np.ndarray:
        """Tree::build (or build_constrained) + InteractionLists::build +
        build_p2p_lists on the device. Permutes b's arrays in place like
        Tree::build; returns perm with new[i] = old[perm[i]]."""
        perm = np.zeros(len(b), np.uint64)
        cb = b._c()
        self._n = None
        _check(lib().galax_fmm_build_tree(self.h, C.byref(cb), _ptr(perm)))
        self._n = len(b)
        return perm

    def upload_tree_data(self, tree: dict, m2l_csr, p2p_csr):
        """GpuContext::upload_tree_data (galax-gpu/src/lib.rs:392-462): a tree
        (dict of node arrays as galax-core's Node fields) and two CSR lists
        (offsets, data) built elsewhere."""
        nn = len(tree["center_x"])
        keep = {k: np.ascontiguousarray(tree[k], np.float64) for k in
                ("center_x", "center_y", "half_width", "mass", "com_x", "com_y")}
        keep["body_start"] = np.ascontiguousarray(tree["body_start"], np.uint64)
        keep["body_end"] = np.ascontiguousarray(tree["body_end"], np.uint64)
        keep["children"] = np.ascontiguousarray(tree["children"], np.int32)
        keep["parent"] = np.ascontiguousarray(tree["parent"], np.int32)
        keep["is_leaf"] = np.ascontiguousarray(tree["is_leaf"], np.uint8)
        tv = _TreeView(nn, *[_ptr(keep[k]) for k in ("center_x", "center_y", "half_width", "mass", "com_x", "com_y",
                                                     "body_start", "body_end", "children", "parent", "is_leaf")])
        lv = []
        for off, dat in (m2l_csr, p2p_csr):
            off = np.ascontiguousarray(off, np.uint64)
            dat = np.ascontiguousarray(dat if len(dat) else np.zeros(1), np.uint32)
            keep[f"o{len(lv)}"], keep[f"d{len(lv)}"] = off, dat
            lv.append(_ListsView(nn, _ptr(off), _ptr(dat)))
        self._n = None
        _check(lib().galax_fmm_upload_tree(self.h, C.byref(tv), C.byref(lv[0]), C.byref(lv[1])))
        leaf = np.asarray(tree["is_leaf"]) != 0
        self._n = int(np.asarray(tree["body_end"])[leaf].max()) if leaf.any() else 0

    def set_softening(self, eps: float):
        _check(lib().galax_fmm_set_softening(self.h, eps))
        self.cfg.eps = eps

    def tree_info(self, m2l_total: bool = True) -> dict:
        """Sizes of the installed tree. With m2l_total, also the M2L list
        length: the first such call on a tree builds the M2L rows of levels
        >= 2 (no evaluation reads them; build_tree leaves them for later)."""
        nn, nl, m2l, p2p = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
        nlev = C.c_uint32()
        _check(lib().galax_fmm_tree_info(self.h, C.byref(nn), C.byref(nl), C.byref(nlev),
                                         C.byref(m2l) if m2l_total else None, C.byref(p2p)))
        out = {"n_nodes": nn.value, "n_leaves": nl.value, "n_levels": nlev.value, "p2p_total": p2p.value}
        if m2l_total:
            out["m2l_total"] = m2l.value
        return out

    def export_tree(self):
        """Returns (tree dict, m2l (offsets, data), p2p (offsets, data))."""
        info = self.tree_info()
        nn = info["n_nodes"]
        t = {k: np.zeros(nn) for k in ("center_x", "center_y", "half_width", "mass", "com_x", "com_y")}
        t["body_start"] = np.zeros(nn, np.uint64)
        t["body_end"] = np.zeros(nn, np.uint64)
        t["children"] = np.zeros((nn, 4), np.int32)
        t["parent"] = np.zeros(nn, np.int32)
        t["is_leaf"] = np.zeros(nn, np.uint8)
        tv = _TreeView(nn, *[_ptr(t[k]) for k in ("center_x", "center_y", "half_width", "mass", "com_x", "com_y",
                                                  "body_start", "body_end", "children", "parent", "is_leaf")])
        mo, md = np.zeros(nn + 1, np.uint64), np.zeros(max(info["m2l_total"], 1), np.uint32)
        po, pd = np.zeros(nn + 1, np.uint64), np.zeros(max(info["p2p_total"], 1), np.uint32)
        lm, lp = _ListsView(nn, _ptr(mo), _ptr(md)), _ListsView(nn, _ptr(po), _ptr(pd))
        _check(lib().galax_fmm_export_tree(self.h, C.byref(tv), C.byref(lm), C.byref(lp)))
        return t, (mo, md[:info["m2l_total"]]), (po, pd[:info["p2p_total"]])

    def export_expansions(self):
        """(multipoles, locals) of the last evaluation, each (n_nodes, S)."""
        nn = self.tree_info(m2l_total=False)["n_nodes"]
        S = (self.cfg.p + 1) * (self.cfg.p + 2) // 2
        mp, loc = np.zeros((nn, S)), np.zeros((nn, S))
        _check(lib().galax_fmm_export_expansions(self.h, _ptr(mp), _ptr(loc), nn))
        return mp, loc

    def export_corrected_lists(self):
        """The corrected variant's lists, built on the device at the first
        evaluation: (m2l (offsets, data), p2p (offsets, data)), one row per
        node (P2P rows: the leaf itself first, internal nodes empty)."""
        nn = self.tree_info(m2l_total=False)["n_nodes"]
        out = []
        for which in (0, 1):
            off = np.zeros(nn + 1, np.uint64)
            lv = _ListsView(nn, _ptr(off), None)
            views = (C.byref(lv), None) if which == 0 else (None, C.byref(lv))
            _check(lib().galax_fmm_export_corrected_lists(self.h, *views))
            dat = np.zeros(max(int(off[-1]), 1), np.uint32)
            lv = _ListsView(nn, _ptr(off), _ptr(dat))
            views = (C.byref(lv), None) if which == 0 else (None, C.byref(lv))
            _check(lib().galax_fmm_export_corrected_lists(self.h, *views))
            out.append((off, dat[:int(off[-1])]))
        return out[0], out[1]

    # -- evaluation ---------------------------------------------------------
    def evaluate(self, b: BodiesSoA):
        """compute_fmm_force semantics: sanitizes b's x, y, mass in place,
        writes b.ax, b.ay (and b.phi when present)."""
        cb = b._c()
        _check(lib().galax_fmm_evaluate(self.h, C.byref(cb)))

    def step(self, b: BodiesSoA):
        """GpuContext::step (galax-gpu/src/lib.rs:464-544) + keep the result for
        read_accelerations."""
        n = len(b)
        tmp = BodiesSoA(b.x.copy(), b.y.copy(), b.vx, b.vy, b.mass.copy(), np.zeros(n), np.zeros(n), np.zeros(n))
        self.evaluate(tmp)
        self._acc = (tmp.ax, tmp.ay, tmp.phi)

    def read_accelerations(self, ax: np.ndarray, ay: np.ndarray):
        """GpuContext::read_accelerations (galax-gpu/src/lib.rs:546-566)."""
        if self._acc is None:
            raise GalaxError(-4, "read_accelerations before step")
        ax[:] = self._acc[0]
        ay[:] = self._acc[1]

    def _dev_ptrs(self, n, **arrays):
        """Device pointers of torch tensors checked against what the library
        reads and writes: float64, contiguous, on the handle's device, n
        elements (None passes through). Raw integers are taken as device
        pointers to n doubles, unchecked. Torch's pending work on the
        tensors' device is finished first: the library runs on its own stream."""
        out, dev = [], None
        for name, t in arrays.items():
            if t is None:
                out.append(None)
                continue
            if not hasattr(t, "data_ptr"):
                out.append(int(t))
                continue
            import torch
            if t.dtype != torch.float64:
                raise TypeError(f"{name}: expected float64, got {t.dtype}")
            if not t.is_cuda:
                raise ValueError(f"{name}: expected a device (CUDA/HIP) tensor")
            if self._device is not None and t.device.index != self._device:
                raise ValueError(f"{name}: tensor on device {t.device.index}, context on {self._device}")
            if not t.is_contiguous():
                raise ValueError(f"{name}: tensor is not contiguous")
            if t.numel() != n:
                raise ValueError(f"{name}: {t.numel()} elements, expected {n}")
            dev = t.device
            out.append(t.data_ptr())
        if dev is not None:
            import torch
            torch.cuda.current_stream(dev).synchronize()
        return out

    def _tree_n(self):
        if self._n is None:
            raise GalaxError(-4, "no tree: call build_tree, build_tree_device or upload_tree_data")
        return self._n

    def evaluate_device(self, x, y, m, ax, ay, phi=None):
        """Device-resident arrays (torch CUDA tensors, checked, or raw pointers)
        of the installed tree's bodies."""
        ptrs = self._dev_ptrs(self._tree_n(), x=x, y=y, m=m, ax=ax, ay=ay, phi=phi)
        _check(lib().galax_fmm_evaluate_device(self.h, *ptrs))

    def build_tree_device(self, x, y, m, vx=None, vy=None, ax=None, ay=None):
        n = int(x.numel()) if hasattr(x, "numel") else len(x)
        ptrs = self._dev_ptrs(n, x=x, y=y, vx=vx, vy=vy, m=m, ax=ax, ay=ay)
        self._n = None
        _check(lib().galax_fmm_build_tree_device(self.h, *ptrs, n))
        self._n = n

    def timings(self) -> Timings:
        t = _Timings()
        _check(lib().galax_fmm_last_timings(self.h, C.byref(t)))
        names = ("sanitize", "p2m", "m2m", "m2l_setup", "m2l", "l2l", "l2p_p2p", "total", "m2l_kernel", "p2p_kernel")
        return Timings({k: getattr(t, k) for k in names}, t.m2l_pairs_executed, t.m2l_pairs_reference, t.m2l_keys,
                       t.p2p_interactions, t.m2l_groups, t.m2l_steps, t.m2l_items)

    def set_timing(self, enable: bool) -> None:
        """Per-pass device timing events of evaluate on (default) or off; with
        them off, timings() reports work counts and zero times."""
        _check(lib().galax_fmm_set_timing(self.h, 1 if enable else 0))

    def build_timings(self) -> dict:
        """Device milliseconds of the last tree build per stage (sort, permute,
        tree, lists, gather, total)."""
        t = _BuildTimings()
        _check(lib().galax_fmm_last_build_timings(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _BuildTimings._fields_}

    def simulate(self, b: BodiesSoA, dt: float, n_steps: int, energy_every: int = 0):
        """simulate / leapfrog_kdk (galax-integrate/src/lib.rs:158-237) on the
        device; b.ax/ay must hold the current accelerations. Returns diagnostics."""
        nd = n_steps // energy_every if energy_every else 0
        diags = (_Diag * max(nd, 1))()
        cb = b._c()
        _check(lib().galax_fmm_kdk(self.h, C.byref(cb), dt, n_steps, energy_every, diags if nd else None))
        return [{k: getattr(diags[i], k) for k, _ in _Diag._fields_} for i in range(nd)]

    def vjp(self, b: BodiesSoA, ax_bar: np.ndarray, ay_bar: np.ndarray):
        n = len(b)
        xb, yb, mb = np.zeros(n), np.zeros(n), np.zeros(n)
        cb = b._c()
        _check(lib().galax_fmm_vjp(self.h, C.byref(cb), _ptr(_f64(ax_bar, n)), _ptr(_f64(ay_bar, n)), _ptr(xb),
                                   _ptr(yb), _ptr(mb)))
        return xb, yb, mb

    def vjp_device(self, x, y, m, ax_bar, ay_bar, x_bar, y_bar, m_bar):
        ptrs = self._dev_ptrs(self._tree_n(), x=x, y=y, m=m, ax_bar=ax_bar, ay_bar=ay_bar, x_bar=x_bar, y_bar=y_bar,
                              m_bar=m_bar)
        _check(lib().galax_fmm_vjp_device(self.h, *ptrs))

    def validate(self, b: BodiesSoA, samples: np.ndarray):
        """validate_fmm (galax-core/src/lib.rs:1503-1564) with given samples:
        returns (max, mean) of the magnitude metric and the vector RMS error."""
        s = np.ascontiguousarray(samples, np.uint64)
        mx, mn, rv = C.c_double(), C.c_double(), C.c_double()
        cb = b._c()
        _check(lib().galax_fmm_validate(self.h, C.byref(cb), _ptr(s), len(s), C.byref(mx), C.byref(mn),
                                        C.byref(rv)))
        return mx.value, mn.value, rv.value

    def crosscheck(self, b: BodiesSoA, threshold: float = 0.5):
        """GpuContext::crosscheck (galax-gpu/src/lib.rs:623-667): the context's
        form against FP64_EXACT (galax-core's arithmetic) on the same tree.
        Returns (max_rel, mean_rel, n); leaves the FP64_EXACT accelerations in
        b.ax, b.ay; raises GalaxError(ThresholdExceeded) where the reference
        panics (max_rel > threshold)."""
        mx, mn, cnt = C.c_double(), C.c_double(), C.c_size_t()
        cb = b._c()
        _check(lib().galax_fmm_crosscheck(self.h, C.byref(cb), threshold, C.byref(mx), C.byref(mn), C.byref(cnt)))
        return mx.value, mn.value, cnt.value

    async def read_accelerations_async(self, ax: np.ndarray, ay: np.ndarray):
        """GpuContext::read_accelerations_async (galax-gpu/src/lib.rs:571-621):
        as on the reference's native target, the blocking read_accelerations."""
        self.read_accelerations(ax, ay)

    def n2_debug(self, b: BodiesSoA):
        """GpuContext::n2_debug (galax-gpu/src/lib.rs:670-714), in FP64."""
        ax, ay, _ = compute_n2_force(b.x, b.y, b.mass, self.cfg.eps)
        return ax, ay


def compute_n2_force(x, y, m, eps, targets=None, want_phi=False):
    """Direct softened sum on the device (galax-core/src/lib.rs:1685-1721)."""
    x, y, m = (np.ascontiguousarray(a, np.float64) for a in (x, y, m))
    n = len(x)
    t = None if targets is None else np.ascontiguousarray(targets, np.uint64)
    nt = n if t is None else len(t)
    ax, ay = np.zeros(nt), np.zeros(nt)
    phi = np.zeros(nt) if want_phi else None
    _check(lib().galax_direct_sum(_ptr(x), _ptr(y), _ptr(m), n, eps, _ptr(t), nt, _ptr(ax), _ptr(ay), _ptr(phi)))
    return ax, ay, phi


def compute_fmm_force(b: BodiesSoA, ctx: GpuContext):
    """compute_fmm_force (galax-core/src/lib.rs:1466-1499) through a context
    whose tree was built or uploaded for these bodies."""
    ctx.evaluate(b)


def default_softening(x: np.ndarray, y: np.ndarray) -> float:
    """galax-cli's default softening (galax-cli/src/main.rs:182-190): 0.1 x the
    mean distance between consecutive bodies (i, i+1 mod n) over the first
    min(n, 1000) bodies of the Morton-sorted arrays, summed in order."""
    n = len(x)
    k = min(n, 1000)
    s = 0.0
    for i in range(k):
        dx = float(x[i]) - float(x[(i + 1) % n])
        dy = float(y[i]) - float(y[(i + 1) % n])
        s += math.sqrt(dx * dx + dy * dy)
    return s / k * 0.1


def default_dt(ax: np.ndarray, ay: np.ndarray) -> float:
    """galax-cli's default time step (galax-cli/src/main.rs:232-244):
    0.5 / sqrt(max |a|), or 0.01 when every |a| <= 1e-30."""
    mx = float(np.max(np.sqrt(ax * ax + ay * ay))) if len(ax) else 0.0
    return 0.5 / math.sqrt(mx) if mx > 1e-30 else 0.01


def version() -> str:
    return lib().galax_fmm_version().decode()


# ------------------------------------------------------------ snapshots ----

def write_snapshot(path, bodies: BodiesSoA, timestamp: float | None = None) -> None:
    """galax-io write_snapshot (galax-io/src/lib.rs:21-42): a GALX v1 file with
    x, y, vx, vy, mass, ax, ay (phi is not stored). The reference stamps the
    current time; pass `timestamp` to fix it."""
    ts = time.time() if timestamp is None else float(timestamp)
    cb = bodies._c()
    _check(lib().galax_snapshot_write(os.fsencode(path), C.byref(cb), ts))


def snapshot_info(path) -> tuple[int, float]:
    """(body count, timestamp) of a GALX file."""
    n, ts = C.c_size_t(0), C.c_double(0.0)
    _check(lib().galax_snapshot_info(os.fsencode(path), C.byref(n), C.byref(ts)))
    return n.value, ts.value


def read_snapshot(path) -> BodiesSoA:
    """galax-io read_snapshot (galax-io/src/lib.rs:53-84). Raises GalaxError with
    code -9 (Format: wrong magic, short file), -6 (Unsupported version) or -8 (Io)."""
    n, _ = snapshot_info(path)
    b = BodiesSoA.new(n, with_phi=False)
    cb = b._c()
    _check(lib().galax_snapshot_read(os.fsencode(path), C.byref(cb)))
    return b
