"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU oracle (prc_oracle.cpp).

The oracle restates the reference solver (prc.cu of jlherraiz/PRC_MLEM) on the CPU in
FP32 (reference semantics) and FP64 (accuracy oracle). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product never does.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
ISOTOPES = ["11C", "13N", "15O", "18F", "64Cu", "68Ga", "82Rb", "124I"]  # prc.cu:47-57
RANGE_BINS = 4096  # prc.cu:61
MIN_DENSITY = np.float32(1e-3)  # prc.cu:102

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile the oracle (g++, OpenMP) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "orc_isotope_index": (C.c_int, [C.c_char_p]),
        "orc_model": (C.c_int, [C.c_int, _f32p]),
        "orc_survival_f32": (C.c_float, [C.c_int, C.c_float]),
        "orc_survival_f64": (C.c_double, [C.c_int, C.c_double]),
        "orc_range_cutoff": (C.c_float, [C.c_int, C.c_float]),
        "orc_range_table": (C.c_float, [C.c_int, C.c_float, _f32p]),
        "orc_survival_at": (C.c_float, [_f32p, C.c_float, C.c_float]),
        "orc_ray_f32": (None, [C.c_int, C.c_int, _f32p, _f32p]),
        "orc_ray_f64": (None, [C.c_int, C.c_int, _f64p, _f64p]),
        "orc_trace_f32": (C.c_long, [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_float, _f32p, C.c_int, C.c_void_p, C.c_float, C.c_int,
                                     C.c_int, C.c_int, _f32p, _f32p, C.c_int, _i32p, _f32p, C.c_int,
                                     C.POINTER(C.c_int)]),
        "orc_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _f32p, C.c_void_p]),
        "orc_destroy": (None, [C.c_void_p]),
        "orc_cutoff": (C.c_float, [C.c_void_p]),
        "orc_n_active": (C.c_int, [C.c_void_p]),
        "orc_table": (None, [C.c_void_p, _f32p]),
        "orc_norm": (None, [C.c_void_p, _f64p]),
        "orc_near_weights": (None, [C.c_void_p, _f64p]),
        "orc_vrs": (None, [C.c_void_p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
        "orc_forward_f32": (None, [C.c_void_p, _f32p, _f32p]),
        "orc_adjoint_f32": (None, [C.c_void_p, _f32p, _f32p]),
        "orc_forward_f64": (None, [C.c_void_p, _f64p, _f64p]),
        "orc_adjoint_f64": (None, [C.c_void_p, _f64p, _f64p]),
        "orc_source_column": (C.c_long, [C.c_void_p, C.c_int, _i32p, _f64p, C.c_int]),
        "orc_solve_f32": (None, [C.c_void_p, _f32p, _f32p, C.c_int, C.c_float, C.c_float, C.c_int,
                                 C.c_void_p, C.c_void_p]),
        "orc_solve_f64": (None, [C.c_void_p, _f64p, _f64p, C.c_int, C.c_double, C.c_double, C.c_int,
                                 C.c_void_p, C.c_void_p]),
        "orc_rdp_update_f64": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                            _f64p, _f64p, _u8p, C.c_long, C.c_double, C.c_double, C.c_double]),
        "orc_rdp_update_f32": (C.c_float, [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           _f32p, _f32p, _u8p, C.c_long, C.c_float, C.c_float, C.c_float]),
        "orc_threads": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    _lib = L
    return L


def isotope_index(name):
    idx = lib().orc_isotope_index(name.encode())
    if idx < 0:
        raise ValueError(f"unknown isotope {name}")
    return idx


def model(name):
    out = np.zeros(15, np.float32)
    n = lib().orc_model(isotope_index(name), out)
    return out[:5 * n].reshape(n, 5)


def survival_f32(name, r):
    return lib().orc_survival_f32(isotope_index(name), float(r))


def survival_f64(name, r):
    return lib().orc_survival_f64(isotope_index(name), float(r))


def range_cutoff(name, fraction=0.9995):
    return lib().orc_range_cutoff(isotope_index(name), fraction)


def range_table(name, cutoff=None):
    if cutoff is None:
        cutoff = range_cutoff(name)
    table = np.zeros(RANGE_BINS + 1, np.float32)
    scale = lib().orc_range_table(isotope_index(name), cutoff, table)
    return table, np.float32(scale)


def survival_at(table, scale, w):
    return lib().orc_survival_at(np.ascontiguousarray(table, np.float32), float(scale), float(w))


def ray_f32(k, n):
    s, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_ray_f32(k, n, s, d)
    return s, d


def ray_f64(k, n):
    s, d = np.zeros(3, np.float64), np.zeros(3, np.float64)
    lib().orc_ray_f64(k, n, s, d)
    return s, d


def trace_f32(rho, spacing, voxel, start, direction, isotope="68Ga", near=False, support_rho=0.05,
              air_cap=50.0, cutoff=None, table=None, cap=4096):
    """One ray of prc.cu:134-165. rho has shape (nz, ny, nx); spacing = (dx, dy, dz) mm;
    voxel = (i, j, k). Returns (visits, targets, weights)."""
    nz, ny, nx = rho.shape
    if cutoff is None:
        cutoff = range_cutoff(isotope)
    tab_ptr, scale = None, 0.0
    if table is not None:
        tab, scale = table
        tab = np.ascontiguousarray(tab, np.float32)
        tab_ptr = tab.ctypes.data
    idx, wts, nd = np.zeros(cap, np.int32), np.zeros(cap, np.float32), C.c_int(0)
    steps = lib().orc_trace_f32(nx, ny, nz, *map(float, spacing), float(cutoff), float(support_rho), float(air_cap),
                                np.ascontiguousarray(rho, np.float32), isotope_index(isotope), tab_ptr, float(scale),
                                *map(int, voxel), np.asarray(start, np.float32), np.asarray(direction, np.float32),
                                int(near), idx, wts, cap, C.byref(nd))
    return steps, idx[:min(nd.value, cap)].copy(), wts[:min(nd.value, cap)].copy()


def preprocess(rho, img, support_rho=0.05):
    """prc.cu:591-613: clamp density to >= 1e-3 and activity to >= 0, zero activity off support.
    Returns (rho, img, support mask, support_sum, input_sum)."""
    rho = np.maximum(np.asarray(rho, np.float32), MIN_DENSITY)
    img = np.maximum(np.asarray(img, np.float32), np.float32(0))
    support = rho >= np.float32(support_rho)
    input_sum = float(np.sum(img, dtype=np.float64))
    img = np.where(support, img, np.float32(0)).astype(np.float32)
    support_sum = float(np.sum(img, dtype=np.float64))
    return rho, img, support, support_sum, input_sum


_CB32 = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_float), C.c_void_p)
_CB64 = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.c_void_p)


class Projector:
    """Reference projector (prc.cu:403-450) on the CPU; rho shape (nz, ny, nx), preprocessed."""

    def __init__(self, rho, spacing, isotope="68Ga", n_rays=128, prec=32, support_rho=0.05, air_cap=50.0,
                 exact=False, table=None):
        self.rho = np.ascontiguousarray(rho, np.float32)
        self.shape = self.rho.shape
        self.n = self.rho.size
        self.prec = prec
        self.dtype = np.float64 if prec == 64 else np.float32
        nz, ny, nx = self.shape
        tab = None if table is None else np.ascontiguousarray(table, np.float32)
        self._tab = tab
        self.h = lib().orc_create(prec, isotope_index(isotope), nx, ny, nz, *map(float, spacing),
                                  float(support_rho), float(air_cap), int(n_rays), int(exact), self.rho,
                                  None if tab is None else tab.ctypes.data)
        self.support = self.rho >= np.float32(support_rho)
        self.active = np.flatnonzero(self.support.ravel()).astype(np.int32)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def cutoff(self):
        return lib().orc_cutoff(self.h)

    @property
    def table(self):
        t = np.zeros(RANGE_BINS + 1, np.float32)
        lib().orc_table(self.h, t)
        return t

    @property
    def W(self):
        out = np.zeros(len(self.active), np.float64)
        lib().orc_norm(self.h, out)
        return out

    @property
    def near_weights(self):
        out = np.zeros(len(self.active) * 125, np.float64)
        lib().orc_near_weights(self.h, out)
        return out.reshape(-1, 125)

    @property
    def vrs(self):
        out = np.zeros(2, np.int64)
        lib().orc_vrs(self.h, out)
        return int(out[0]), int(out[1])

    def column(self, a, cap=1 << 20):
        idx, w = np.zeros(cap, np.int32), np.zeros(cap, np.float64)
        c = lib().orc_source_column(self.h, int(a), idx, w, cap)
        return idx[:c].copy(), w[:c].copy()

    def forward(self, x):
        x = np.ascontiguousarray(x, self.dtype).ravel()
        out = np.zeros(self.n, self.dtype)
        (lib().orc_forward_f64 if self.prec == 64 else lib().orc_forward_f32)(self.h, x, out)
        return out.reshape(self.shape)

    def adjoint(self, y, out=None):
        y = np.ascontiguousarray(y, self.dtype).ravel()
        out = np.zeros(self.n, self.dtype) if out is None else np.ascontiguousarray(out, self.dtype).ravel().copy()
        (lib().orc_adjoint_f64 if self.prec == 64 else lib().orc_adjoint_f32)(self.h, y, out)
        return out.reshape(self.shape)

    def solve(self, g, iterations, beta, gamma=20.0, accel=True, history=False):
        """MAP-EM loop (prc.cu:659-710) on preprocessed g; returns the estimate (and the
        per-iteration estimates when history=True)."""
        g = np.ascontiguousarray(g, self.dtype).ravel()
        out = np.zeros(self.n, self.dtype)
        hist = []
        if history:
            ptr_t = C.POINTER(C.c_double if self.prec == 64 else C.c_float)

            def cb(it, p, _):
                hist.append(np.ctypeslib.as_array(C.cast(p, ptr_t), (self.n,)).copy().reshape(self.shape))

            cfn = (_CB64 if self.prec == 64 else _CB32)(cb)
            cb_ptr = C.cast(cfn, C.c_void_p)
        else:
            cfn, cb_ptr = None, None
        fn = lib().orc_solve_f64 if self.prec == 64 else lib().orc_solve_f32
        fn(self.h, g, out, int(iterations), float(beta), float(gamma), int(bool(accel)), cb_ptr, None)
        out = out.reshape(self.shape)
        return (out, hist) if history else out


def threads():
    return lib().orc_threads()


def set_threads(n):
    """Threads of the oracle's parallel loops from now on."""
    lib().orc_set_threads(int(n))
