"""ctypes bindings of the oracle (oracle/_build/liboracle.so).

Test infrastructure only: the checker of the HIP solver. Only tests/, the
smoke check and the CPU-baseline leg of bench.py load it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

REF_CUDA, PD, EXACT = 0, 1, 2
RESTART_PREV, RESTART_SEQ_MIN, RESTART_NONE = 0, 1, 2

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrMesh(C.Structure):
    _fields_ = [("N", C.c_int), ("T", C.c_int), ("E", C.c_int), ("P", C.c_int),
                ("mass", _fp), ("tris", _ip), ("Dm_inv", _fp), ("rest_area", _fp),
                ("stretch_k", _fp), ("diag", _fp), ("quads", _ip), ("bend_w", _fp),
                ("bend_n", _fp), ("bend_k", _fp), ("pin_idx", _ip), ("pin_target", _fp)]


class OrParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("iterations", C.c_int), ("use_chebyshev", C.c_int),
                ("rho", C.c_float), ("cheb_delay", C.c_int), ("gravity", C.c_float),
                ("dt", C.c_float), ("damping", C.c_float), ("ext_accel", C.c_float * 3),
                ("tolerance", C.c_float), ("check_every", C.c_int), ("exact_bend", C.c_int),
                ("exact_pins", C.c_int), ("threads", C.c_int), ("relax", C.c_float),
                ("restart_rule", C.c_int), ("restart_factor", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run make -C oracle")
        L = C.CDLL(LIB)
        L.or_params_default.argtypes = [C.POINTER(OrParams)]
        L.or_params_default.restype = None
        L.or_step_f32.argtypes = [C.POINTER(OrMesh), C.POINTER(OrParams), _fp, _fp, _fp, _fp, _dp,
                                  _ip]
        L.or_step_f32.restype = C.c_int
        L.or_step_f64.argtypes = [C.POINTER(OrMesh), C.POINTER(OrParams), _dp, _dp, _dp, _dp, _dp,
                                  _ip]
        L.or_step_f64.restype = C.c_int
        L.or_ref_cuda_rotation_f32.argtypes = [_fp, _fp]
        L.or_ref_cuda_rotation_f32.restype = None
        L.or_polar_rotation_f64.argtypes = [_dp, _dp]
        L.or_polar_rotation_f64.restype = None
        L.or_polar_rotation_f32.argtypes = [_fp, _fp]
        L.or_polar_rotation_f32.restype = None
        L.or_rest_state_f64.argtypes = [C.c_int, C.c_int, _fp, _ip, C.c_float, _dp, _dp, _dp]
        L.or_rest_state_f64.restype = None
        L.or_bend_weights_f64.argtypes = [C.c_int, _fp, _ip, _dp, _dp]
        L.or_bend_weights_f64.restype = None
        L.or_jacobi_diag_f64.argtypes = [C.POINTER(OrMesh), C.c_float, _dp]
        L.or_jacobi_diag_f64.restype = None
        L.or_residual_f64.argtypes = [C.POINTER(OrMesh), C.c_float, _dp, _dp, _dp, _dp, _dp]
        L.or_residual_f64.restype = None
        L.or_residual_disp_f32.argtypes = [C.POINTER(OrMesh), C.c_float, _fp, _dp, _fp, _dp, _dp,
                                           _dp]
        L.or_residual_disp_f32.restype = None
        _lib = L
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Mesh:
    """Flat arrays of one cloth (the inputs the HIP solver reads)."""

    def __init__(self, arrays: dict, stretch=True, bend=True):
        c = lambda k, t: np.ascontiguousarray(arrays[k], dtype=t)
        self.mass = c("mass", np.float32)
        self.tris = c("tris", np.int32)
        self.Dm_inv = c("Dm_inv", np.float32)
        self.rest_area = c("rest_area", np.float32)
        self.stretch_k = c("stretch_k", np.float32) if stretch else None
        self.diag = c("jacobi_diag", np.float32)
        has_bend = bend and arrays.get("bend_quads") is not None and len(arrays["bend_quads"]) > 0
        self.quads = c("bend_quads", np.int32) if has_bend else None
        self.bend_w = c("bend_w", np.float32) if has_bend else None
        self.bend_n = c("bend_n", np.float32) if has_bend else None
        self.bend_k = c("bend_k", np.float32) if has_bend else None
        self.pin_idx = c("pin_idx", np.int32)
        self.pin_target = c("pin_target", np.float32)
        self.N = self.mass.shape[0]
        self.T = self.tris.shape[0]
        self.E = self.quads.shape[0] if has_bend else 0
        self.P = self.pin_idx.shape[0]

    def struct(self) -> OrMesh:
        return OrMesh(self.N, self.T, self.E, self.P, _f(self.mass), _i(self.tris),
                      _f(self.Dm_inv), _f(self.rest_area), _f(self.stretch_k), _f(self.diag),
                      _i(self.quads), _f(self.bend_w), _f(self.bend_n), _f(self.bend_k),
                      _i(self.pin_idx), _f(self.pin_target))


def params(**kw) -> OrParams:
    p = OrParams()
    lib().or_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "ext_accel":
            for i in range(3):
                p.ext_accel[i] = float(v[i])
        else:
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
    return p


def step(mesh: Mesh, p: OrParams, pos, vel, precision="f32", trace=False, resid=False,
         init=None):
    """One frame; returns (pos, vel, trace or None, resid or None, iterations).
    init: optional first iterate (default: pos)."""
    dt = np.float32 if precision == "f32" else np.float64
    pos = np.array(pos, dtype=dt, copy=True).reshape(-1, 3)
    vel = np.array(vel, dtype=dt, copy=True).reshape(-1, 3)
    ini = None if init is None else np.ascontiguousarray(init, dtype=dt).reshape(-1, 3)
    tr = np.zeros((p.iterations, mesh.N, 3), dt) if trace else None
    rs = np.zeros((p.iterations, 2), np.float64) if resid else None
    it = C.c_int(0)
    m = mesh.struct()
    if precision == "f32":
        st = lib().or_step_f32(C.byref(m), C.byref(p), _f(pos), _f(vel), _f(ini), _f(tr),
                               _d(rs), C.byref(it))
    else:
        st = lib().or_step_f64(C.byref(m), C.byref(p), _d(pos), _d(vel), _d(ini), _d(tr),
                               _d(rs), C.byref(it))
    if st != 0:
        raise RuntimeError(f"oracle step failed ({st})")
    n = it.value
    return pos, vel, (tr[:n] if trace else None), (rs[:n] if resid else None), n


def ref_cuda_rotation(F: np.ndarray) -> np.ndarray:
    F = np.ascontiguousarray(F, np.float32).reshape(6)
    R = np.zeros(6, np.float32)
    lib().or_ref_cuda_rotation_f32(_f(F), _f(R))
    return R


def polar_rotation(F: np.ndarray, precision="f64") -> np.ndarray:
    if precision == "f64":
        F = np.ascontiguousarray(F, np.float64).reshape(6)
        R = np.zeros(6, np.float64)
        lib().or_polar_rotation_f64(_d(F), _d(R))
    else:
        F = np.ascontiguousarray(F, np.float32).reshape(6)
        R = np.zeros(6, np.float32)
        lib().or_polar_rotation_f32(_f(F), _f(R))
    return R


def rest_state_f64(rest_pos, tris, density):
    rest_pos = np.ascontiguousarray(rest_pos, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    N, T = rest_pos.shape[0], tris.shape[0]
    dm = np.zeros((T, 4)); ar = np.zeros(T); ms = np.zeros(N)
    lib().or_rest_state_f64(N, T, _f(rest_pos), _i(tris), density, _d(dm), _d(ar), _d(ms))
    return dm, ar, ms


def bend_weights_f64(rest_pos, quads):
    rest_pos = np.ascontiguousarray(rest_pos, np.float32)
    quads = np.ascontiguousarray(quads, np.int32)
    E = quads.shape[0]
    w = np.zeros((E, 4)); n = np.zeros(E)
    lib().or_bend_weights_f64(E, _f(rest_pos), _i(quads), _d(w), _d(n))
    return w, n


def jacobi_diag_f64(mesh: Mesh, dt: float):
    out = np.zeros(mesh.N)
    m = mesh.struct()
    lib().or_jacobi_diag_f64(C.byref(m), dt, _d(out))
    return out


def residual_f64(mesh: Mesh, dt: float, x, y):
    """FP64 (r, ||r||^2, ||b||^2) of state x for prediction y."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 3)
    r = np.zeros_like(x)
    r2, b2 = C.c_double(0), C.c_double(0)
    m = mesh.struct()
    lib().or_residual_f64(C.byref(m), dt, _d(x), _d(y), _d(r), C.byref(r2), C.byref(b2))
    return r, r2.value, b2.value


def prediction(pos, vel, dt, gravity, ext=(0.0, 0.0, 0.0)):
    """y = x + h v + h^2 a in FP64 (gravity along y)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    vel = np.asarray(vel, np.float64).reshape(-1, 3)
    a = np.array([ext[0], gravity + ext[1], ext[2]], np.float64)
    return pos + dt * vel + dt * dt * a


def residual_disp_f32(mesh: Mesh, dt: float, base, disp, dy):
    """FP32 (r, ||r||^2, ||b||^2) of base + disp in the iteration's form."""
    base = np.ascontiguousarray(base, np.float32).reshape(-1, 3)
    disp = np.ascontiguousarray(disp, np.float64).reshape(-1, 3)
    dy = np.ascontiguousarray(dy, np.float32).reshape(-1, 3)
    r = np.zeros(base.shape, np.float64)
    r2, b2 = C.c_double(0), C.c_double(0)
    m = mesh.struct()
    lib().or_residual_disp_f32(C.byref(m), dt, _f(base), _d(disp), _f(dy), _d(r), C.byref(r2),
                               C.byref(b2))
    return r, r2.value, b2.value
