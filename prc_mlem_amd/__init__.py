"""MI355X-native positron-range correction: Python front end of libprc_amd.so.

Mirrors the reference's operator interface (prc.cu:403-450 `Projector`: construction runs
the normalisation, `forward` overwrites its output, `adjoint` writes active voxels only)
and its solver loop (prc.cu:659-710), over the C ABI declared in include/prc.h. Arrays are
numpy float32 volumes shaped (NZ, NY, NX), X fastest, as the reference's raw files.

There is no CPU fallback: every call goes to the HIP library, and importing a function that
needs it raises if libprc_amd.so has not been built.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_DIR = HERE / "_lib"
LIB_PATH = LIB_DIR / "libprc_amd.so"
PRC_EXE = LIB_DIR / "prc"

PRECISION = {"fp32": 0, "mixed": 1}
FLAG_EXACT_RANGE = 1
FLAG_NO_CACHE_NEAR = 2
FLAG_NO_NEAR_MEMO = 4
FLAG_MARCH_PASSES = 8
RANGE_BINS = 4096
NEAR_RAYS = 2048


class PrcError(RuntimeError):
    """A failed call of the C ABI; .code is the negative status."""

    def __init__(self, code, message):
        super().__init__(f"{message} (status {code})")
        self.code = code


class Geometry(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("sx", C.c_float), ("sy", C.c_float),
                ("sz", C.c_float), ("isotope", C.c_int), ("n_rays", C.c_int), ("support_rho", C.c_float),
                ("air_cap_mm", C.c_float), ("precision", C.c_int), ("flags", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("beta", C.c_float), ("gamma", C.c_float), ("accelerate", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n_voxels", C.c_int64), ("n_active", C.c_int64), ("n_rows", C.c_int64),
                ("weight_entries", C.c_int64), ("target_entries", C.c_int64), ("weight_bytes", C.c_int64),
                ("vrs_near", C.c_int64),
                ("vrs_far", C.c_int64), ("deposits", C.c_int64), ("near_uniform", C.c_int64),
                ("norm_seconds", C.c_double),
                ("cutoff_mm", C.c_float), ("weight_planes", C.c_int64), ("target_planes", C.c_int64),
                ("far_marched", C.c_int64), ("nonzero_entries", C.c_int64)]


# Every symbol include/prc.h declares, with its ctypes signature.
_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
SIGNATURES = {
    "prc_version": (C.c_char_p, []),
    "prc_last_error": (C.c_char_p, []),
    "prc_isotope_count": (C.c_int, []),
    "prc_isotope_name": (C.c_char_p, [C.c_int]),
    "prc_isotope_find": (C.c_int, [C.c_char_p]),
    "prc_range_cutoff": (C.c_float, [C.c_int, C.c_float]),
    "prc_ray_set": (C.c_int, [C.c_int, _F32P, _F32P]),
    "prc_projector_create": (C.c_int, [C.POINTER(Geometry), _F32P, C.c_void_p, C.POINTER(C.c_void_p)]),
    "prc_projector_destroy": (None, [C.c_void_p]),
    "prc_projector_normalise": (C.c_int, [C.c_void_p]),
    "prc_projectors_normalise": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "prc_projector_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "prc_projector_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "prc_projector_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "prc_projector_forward_host": (C.c_int, [C.c_void_p, _F32P, _F32P]),
    "prc_projector_adjoint_host": (C.c_int, [C.c_void_p, _F32P, _F32P]),
    "prc_projector_norm": (C.c_int, [C.c_void_p, _F32P]),
    "prc_projector_table": (C.c_int, [C.c_void_p, _F32P, C.POINTER(C.c_float)]),
    "prc_projector_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "prc_projector_norm_device": (C.c_void_p, [C.c_void_p]),
    "prc_solver_create": (C.c_int, [C.c_void_p, _F32P, C.POINTER(SolveParams), C.POINTER(C.c_void_p)]),
    "prc_solver_destroy": (None, [C.c_void_p]),
    "prc_solver_reset": (C.c_int, [C.c_void_p]),
    "prc_solver_iterate": (C.c_int, [C.c_void_p, C.c_int]),
    "prc_solver_sync": (C.c_int, [C.c_void_p]),
    "prc_solver_estimate": (C.c_int, [C.c_void_p, _F32P]),
    "prc_solver_estimate_device": (C.c_void_p, [C.c_void_p]),
    "prc_solver_reproject": (C.c_int, [C.c_void_p, _F32P]),
    "prc_solver_sums": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "prc_batch_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(SolveParams), C.POINTER(C.c_void_p)]),
    "prc_batch_normalise": (C.c_int, [C.c_void_p]),
    "prc_batch_run": (C.c_int, [C.c_void_p, C.c_int]),
    "prc_batch_sync": (C.c_int, [C.c_void_p]),
    "prc_batch_estimate": (C.c_int, [C.c_void_p, C.c_int, _F32P]),
    "prc_batch_destroy": (None, [C.c_void_p]),
    "prc_batch_correct": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                    C.POINTER(SolveParams), C.c_int, C.POINTER(C.c_void_p)]),
    "prc_profile_enable": (C.c_int, [C.c_int]),
    "prc_profile_report": (C.c_int, [C.c_char_p, C.c_size_t]),
    "prc_tuning_report": (C.c_int, [C.c_char_p, C.c_size_t]),
}

_lib = None


def build(arch="gfx950"):
    """Compile libprc_amd.so and the prc executable in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", str(HERE / "csrc"), f"ARCH={arch}"], check=True)


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with prc_mlem_amd.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PrcError(rc, lib().prc_last_error().decode())


def isotopes():
    L = lib()
    return [L.prc_isotope_name(i).decode() for i in range(L.prc_isotope_count())]


def isotope_index(name):
    idx = lib().prc_isotope_find(name.encode())
    if idx < 0:
        raise ValueError(f"unknown isotope {name!r}; supported: {' '.join(isotopes())}")
    return idx


def range_cutoff(isotope, fraction=0.9995):
    """Water-equivalent range cutoff in mm (prc.cu:83-91)."""
    return lib().prc_range_cutoff(isotope_index(isotope), fraction)


def ray_set(n):
    """Start offsets (voxels) and unit directions of the n-ray set as the device computes them,
    each (n, 3) float32 (prc.cu:118-129)."""
    start, direction = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    _check(lib().prc_ray_set(int(n), start, direction))
    return start, direction


def _volume(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


class Projector:
    """System matrix of one volume: A_ts = w_st / W_s (prc.cu:167-267).

    rho: density (NZ, NY, NX) in g/cm3; spacing: (DX, DY, DZ) mm. Construction runs the
    normalisation pass on the GPU and stores every weight once.
    """

    def __init__(self, rho, spacing, isotope="68Ga", n_rays=128, support_rho=0.05, air_cap_mm=50.0,
                 precision="fp32", exact_range=False, table=None, near_memo=True, march_passes=False):
        rho = _volume(rho)
        if rho.ndim != 3:
            raise ValueError("rho must be a 3D (NZ, NY, NX) volume")
        self.shape = rho.shape
        nz, ny, nx = rho.shape
        self.geometry = Geometry(nx, ny, nz, *map(float, spacing), isotope_index(isotope), int(n_rays),
                                 float(support_rho), float(air_cap_mm), PRECISION[precision],
                                 (FLAG_EXACT_RANGE if exact_range else 0) | (0 if near_memo else FLAG_NO_NEAR_MEMO)
                                 | (FLAG_MARCH_PASSES if march_passes else 0))
        self.precision = precision
        h = C.c_void_p()
        tab = None if table is None else np.ascontiguousarray(table, np.float32)
        _check(lib().prc_projector_create(C.byref(self.geometry), rho, None if tab is None else tab.ctypes.data,
                                          C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().prc_projector_destroy(self.handle)
            self.handle = None

    __del__ = close

    def forward(self, x):
        """A x (every voxel written), host arrays."""
        out = np.zeros(self.shape, np.float32)
        _check(lib().prc_projector_forward_host(self.handle, _volume(x, self.shape), out))
        return out

    def adjoint(self, y, out=None):
        """A^T y on active voxels; other voxels keep `out` (zeros by default)."""
        out = np.zeros(self.shape, np.float32) if out is None else _volume(out, self.shape).copy()
        _check(lib().prc_projector_adjoint_host(self.handle, _volume(y, self.shape), out))
        return out

    def forward_device(self, x_ptr, out_ptr):
        """A x on device pointers (e.g. torch tensors' data_ptr()); stream-ordered on the null stream."""
        _check(lib().prc_projector_forward(self.handle, C.c_void_p(x_ptr), C.c_void_p(out_ptr)))

    def forward_device_noalloc(self, x_ptr, out_ptr):
        """A x on device pointers, allocating nothing (the staging vector is a static device array)."""
        _check(lib().prc_projector_forward_noalloc(self.handle, C.c_void_p(x_ptr), C.c_void_p(out_ptr)))

    def adjoint_device(self, y_ptr, out_ptr):
        _check(lib().prc_projector_adjoint(self.handle, C.c_void_p(y_ptr), C.c_void_p(out_ptr)))

    def normalise(self):
        """Rebuild the stored weights from the resident density; returns when the build is done."""
        _check(lib().prc_projector_normalise(self.handle))

    @property
    def W(self):
        """Per-voxel normalisation W (0 off support)."""
        w = np.zeros(self.shape, np.float32)
        _check(lib().prc_projector_norm(self.handle, w))
        return w

    @property
    def table(self):
        t = np.zeros(RANGE_BINS + 1, np.float32)
        scale = C.c_float()
        _check(lib().prc_projector_table(self.handle, t, C.byref(scale)))
        return t, scale.value

    @property
    def stats(self):
        s = Stats()
        _check(lib().prc_projector_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}


class Solver:
    """MAP-EM with the RDP prior and Biggs-Andrews acceleration (prc.cu:659-710)."""

    def __init__(self, projector, g, beta=0.2, gamma=20.0, accelerate=True):
        self.projector = projector
        self.params = SolveParams(float(beta), float(gamma), int(bool(accelerate)))
        h = C.c_void_p()
        _check(lib().prc_solver_create(projector.handle, _volume(g, projector.shape), C.byref(self.params),
                                       C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().prc_solver_destroy(self.handle)
            self.handle = None

    __del__ = close

    def reset(self):
        _check(lib().prc_solver_reset(self.handle))
        return self

    def iterate(self, n=1):
        _check(lib().prc_solver_iterate(self.handle, int(n)))
        return self

    def sync(self):
        _check(lib().prc_solver_sync(self.handle))

    def estimate(self):
        out = np.zeros(self.projector.shape, np.float32)
        _check(lib().prc_solver_estimate(self.handle, out))
        return out

    def reproject(self):
        out = np.zeros(self.projector.shape, np.float32)
        _check(lib().prc_solver_reproject(self.handle, out))
        return out

    def sums(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().prc_solver_sums(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value


def correct(projector, g, iterations=20, beta=0.2, gamma=20.0, accelerate=True):
    """One correction run (the reference's --mode correct) on host arrays."""
    s = Solver(projector, g, beta, gamma, accelerate)
    try:
        s.iterate(iterations)
        return s.estimate()
    finally:
        s.close()


class Batch:
    """Several volumes corrected together: one launch per stage over all volumes and a
    graph-captured iteration loop. beta/gamma may be scalars or per-volume sequences."""

    def __init__(self, projectors, data, beta=0.2, gamma=20.0, accelerate=True):
        n = len(projectors)
        self.projectors = list(projectors)
        betas = np.broadcast_to(np.asarray(beta, np.float32), (n,))
        gammas = np.broadcast_to(np.asarray(gamma, np.float32), (n,))
        self.params = (SolveParams * n)(*[SolveParams(float(b), float(g), int(bool(accelerate)))
                                          for b, g in zip(betas, gammas)])
        ins = [_volume(d, p.shape) for d, p in zip(data, projectors)]
        hs = (C.c_void_p * n)(*[p.handle.value for p in projectors])
        ip = (C.c_void_p * n)(*[a.ctypes.data for a in ins])
        h = C.c_void_p()
        _check(lib().prc_batch_create(n, hs, ip, self.params, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().prc_batch_destroy(self.handle)
            self.handle = None

    __del__ = close

    def normalise(self):
        """Rebuild every projector's weights on the batch's stream (queued before the next run)."""
        _check(lib().prc_batch_normalise(self.handle))
        return self

    def run(self, iterations):
        _check(lib().prc_batch_run(self.handle, int(iterations)))
        return self

    def sync(self):
        _check(lib().prc_batch_sync(self.handle))

    def estimate(self, v):
        out = np.zeros(self.projectors[v].shape, np.float32)
        _check(lib().prc_batch_estimate(self.handle, int(v), out))
        return out


def batch_correct(projectors, data, iterations, beta=0.2, gamma=20.0, accelerate=True):
    """One batched correction; returns the estimates."""
    b = Batch(projectors, data, beta, gamma, accelerate)
    try:
        b.run(iterations)
        return [b.estimate(v) for v in range(len(projectors))]
    finally:
        b.close()


def profile_enable(on=True):
    """Time every kernel launch with HIP events (graph replay is bypassed while on)."""
    _check(lib().prc_profile_enable(int(bool(on))))


def profile_report():
    """Per-kernel {count, total_ms, mean_ms, min_ms, max_ms} since the last report."""
    import json
    buf = C.create_string_buffer(1 << 16)
    _check(lib().prc_profile_report(buf, len(buf)))
    return json.loads(buf.value.decode())


def normalise_all(projectors):
    """Normalise several projectors at once (each on its own stream; their kernels overlap)."""
    arr = (C.c_void_p * len(projectors))(*[p.handle for p in projectors])
    _check(lib().prc_projectors_normalise(len(projectors), arr))


def tuning_report():
    """Projector-pass tuning decisions: per solver or batch, the mean launch time of each schedule
    of the forward and the adjoint and the one kept."""
    import json
    buf = C.create_string_buffer(1 << 16)
    _check(lib().prc_tuning_report(buf, len(buf)))
    return json.loads(buf.value.decode())


def run_cli(*args, env=None, timeout=600):
    """Run the prc executable (the drop-in CLI) and return the CompletedProcess."""
    return subprocess.run([str(PRC_EXE), *map(str, args)], capture_output=True, text=True, timeout=timeout,
                          env=env if env is not None else os.environ.copy())
