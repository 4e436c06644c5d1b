"""ctypes bindings of libclothhip.so (C ABI, include/clothhip.h) and of the
scene / PDSolver C API of libcloth_core.so (host/cloth_scene.h).

This is the Python-side caller of the drop-in boundary: it is what a
reference-side ctypes user would write. Loading fails loudly when the built
libraries are missing; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "libclothhip.so")
LIB_CORE = os.path.join(HERE, "libcloth_core.so")

# status / enums (include/clothhip.h)
CLH_OK = 0
CLH_H2H, CLH_H2D, CLH_D2H, CLH_D2D = 0, 1, 2, 3
MODE_REF_CUDA, MODE_PD = 0, 1
KERNEL_AUTO, KERNEL_UNFUSED, KERNEL_FUSED = 0, 1, 2
OK, ERR_INVALID, ERR_DEVICE, ERR_ALLOC, ERR_UNSUPPORTED, ERR_NUMERIC = range(6)
CLOTH_RUNNING, CLOTH_CONVERGED, CLOTH_STALLED, CLOTH_NONFINITE = range(4)
PERSIST_OFF, PERSIST_ON = 0, 1
RESTART_PREV, RESTART_SEQ_MIN, RESTART_NONE = 0, 1, 2
TUNE_MAX = 40  # CLH_TUNE_MAX_CANDIDATES
RELAX_AUTO = -1.0
PIN_NONE, PIN_TOP_ROW, PIN_CORNERS, PIN_RANDOM_BOUNDARY = 0, 1, 2, 3
MESH_IRREGULAR = 4

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int),
                ("clock_khz", C.c_int), ("mem_clock_khz", C.c_int), ("mem_bus_width", C.c_int),
                ("total_mem", C.c_size_t), ("l2_cache_bytes", C.c_size_t)]


class Config(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_iterations", C.c_int), ("use_chebyshev", C.c_int),
                ("rho", C.c_float), ("gravity", C.c_float), ("dt", C.c_float),
                ("damping", C.c_float), ("tolerance", C.c_float), ("check_every", C.c_int),
                ("cheb_delay", C.c_int), ("ext_accel", C.c_float * 3),
                ("kernel_variant", C.c_int), ("use_graph", C.c_int), ("tile_verts", C.c_int),
                ("block_threads", C.c_int), ("persistent", C.c_int),
                ("iters_per_launch", C.c_int), ("element_regions", C.c_int),
                ("stall_checks", C.c_int), ("restart_rule", C.c_int),
                ("restart_factor", C.c_float), ("relax", C.c_float)]


class MeshDesc(C.Structure):
    _fields_ = [("num_verts", C.c_int), ("num_tris", C.c_int), ("num_bend", C.c_int),
                ("d_tris", _vp), ("d_bend_quads", _vp)]


STEP_FIELDS = ["d_pos", "d_vel", "d_mass", "d_tris", "d_Dm_inv", "d_rest_area", "d_stretch_k",
               "d_jacobi_diag", "d_bend_quads", "d_bend_w", "d_bend_n", "d_bend_k",
               "d_pin_idx", "d_pin_target"]


class StepArgs(C.Structure):
    _fields_ = ([(f, _vp) for f in STEP_FIELDS] +
                [("num_pinned", C.c_int), ("d_init_guess", _vp), ("d_trace", _vp),
                 ("d_cloth_accel", _vp)])


class StepStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("residual0", C.c_double), ("rhs_norm", C.c_double), ("frames", C.c_int),
                ("restarts", C.c_int), ("rho", C.c_float), ("relax", C.c_float),
                ("jacobi_bound", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [("variant", C.c_int), ("num_tiles", C.c_int), ("tile_verts", C.c_int),
                ("halo_tri_factor", C.c_double), ("halo_bend_factor", C.c_double),
                ("bytes_per_iter", C.c_double), ("lds_bytes", C.c_double),
                ("block_threads", C.c_int), ("num_cloths", C.c_int), ("padded_verts", C.c_int),
                ("persistent", C.c_int), ("tri_max", C.c_int), ("hinge_max", C.c_int),
                ("halo_max", C.c_int), ("pair", C.c_int), ("iters_per_launch", C.c_int),
                ("pair_halo_max", C.c_int), ("pair_tri_max", C.c_int),
                ("pair_hinge_max", C.c_int), ("pair_bytes_per_launch", C.c_double),
                ("pair_compute_max", C.c_int), ("pair_lds_bytes", C.c_double),
                ("element_regions", C.c_int), ("split_lds_bytes", C.c_double),
                ("lds_limit", C.c_int), ("gather_lanes", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [("num_candidates", C.c_int), ("tile_verts", C.c_int * TUNE_MAX),
                ("block_threads", C.c_int * TUNE_MAX), ("iters_per_launch", C.c_int * TUNE_MAX),
                ("element_regions", C.c_int * TUNE_MAX), ("us_per_iteration", C.c_float * TUNE_MAX),
                ("best", C.c_int), ("from_cache", C.c_int), ("key", C.c_ulonglong)]

    def candidates(self):
        """[{tile_verts, block_threads, iters_per_launch, element_regions,
        us_per_iteration}]"""
        return [{"tile_verts": self.tile_verts[i], "block_threads": self.block_threads[i],
                 "iters_per_launch": self.iters_per_launch[i],
                 "element_regions": self.element_regions[i],
                 "us_per_iteration": self.us_per_iteration[i]}
                for i in range(self.num_candidates)]


class SceneSpec(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("size", C.c_float), ("type", C.c_int),
                ("rest_jitter", C.c_float), ("seed", C.c_uint64), ("density", C.c_float),
                ("stretch_k", C.c_float), ("stretch_k_max", C.c_float), ("bend_k", C.c_float),
                ("pin_mode", C.c_int), ("pin_count", C.c_int), ("pos_jitter", C.c_float),
                ("dt", C.c_float)]


class PDConfig(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("tolerance", C.c_float),
                ("use_chebyshev", C.c_int), ("rho", C.c_float), ("gravity", C.c_float),
                ("dt", C.c_float), ("damping", C.c_float), ("mode", C.c_int),
                ("external_accel", C.c_float * 3), ("check_every", C.c_int),
                ("cheb_delay", C.c_int), ("kernel_variant", C.c_int), ("use_graph", C.c_int)]


# every symbol include/clothhip.h declares
CLH_SYMBOLS = [
    "clh_last_error", "clh_device_count", "clh_get_device_info", "clh_alloc", "clh_free",
    "clh_memcpy", "clh_memset", "clh_synchronize", "clh_apply_constraints",
    "clh_config_default", "clh_solver_create", "clh_solver_destroy", "clh_solver_reset",
    "clh_solver_step", "clh_solver_get_stats", "clh_solver_residual_history",
    "clh_solver_set_stream", "clh_solver_run_frames", "clh_solver_time_iteration_kernel",
    "clh_solver_plan_info", "clh_measure_copy", "clh_fused_plan_check",
    "clh_batch_create", "clh_batch_cloth_residuals", "clh_solver_autotune",
    "clh_ws_create", "clh_ws_destroy", "clh_ws_set_stream", "clh_ws_param_count",
    "clh_ws_get_params", "clh_ws_set_params", "clh_ws_predict", "clh_ws_features",
    "clh_ws_dataset_reserve", "clh_ws_dataset_add", "clh_ws_dataset_size", "clh_ws_dataset_clear",
    "clh_ws_dataset_read", "clh_ws_fit_output_scale", "clh_ws_loss_grad", "clh_ws_train",
    "clh_ws_adam_step", "clh_ws_set_smoothing", "clh_solver_start_residual",
    "clh_ws_dataset_append", "clh_step_noalloc", "clh_ws_set_step_filter",
    "clh_batch_cloth_status",
]


class WsDesc(C.Structure):
    _fields_ = [("num_verts", C.c_int), ("num_tris", C.c_int), ("d_tris", _vp),
                ("num_cloths", C.c_int), ("cloth_vert_offsets", _vp), ("d_pin_idx", _vp),
                ("num_pinned", C.c_int), ("dt", C.c_float), ("gravity", C.c_float),
                ("ext_accel", C.c_float * 3), ("seed", C.c_ulonglong),
                ("ext_feature", C.c_int)]

_hip = None
_core = None


def _load():
    global _hip, _core
    if _hip is not None:
        return _hip, _core
    for path in (LIB_HIP, LIB_CORE):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build() first")
    hip = C.CDLL(LIB_HIP, mode=C.RTLD_GLOBAL)
    core = C.CDLL(LIB_CORE)
    hip.clh_last_error.restype = C.c_char_p
    for name in CLH_SYMBOLS:
        if name != "clh_last_error":
            getattr(hip, name).restype = C.c_int
    hip.clh_alloc.argtypes = [C.POINTER(_vp), C.c_size_t]
    hip.clh_free.argtypes = [_vp]
    hip.clh_memcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
    hip.clh_memset.argtypes = [_vp, C.c_int, C.c_size_t]
    hip.clh_apply_constraints.argtypes = [_vp, _vp, _vp, _vp, C.c_int]
    hip.clh_config_default.argtypes = [C.POINTER(Config)]
    hip.clh_solver_create.argtypes = [C.POINTER(Config), C.POINTER(MeshDesc), C.POINTER(_vp)]
    hip.clh_solver_destroy.argtypes = [_vp]
    hip.clh_solver_reset.argtypes = [_vp]
    hip.clh_solver_step.argtypes = [_vp, C.POINTER(StepArgs)]
    hip.clh_solver_get_stats.argtypes = [_vp, C.POINTER(StepStats)]
    hip.clh_solver_residual_history.argtypes = [_vp, C.POINTER(C.c_double), C.c_int, _ip]
    hip.clh_solver_set_stream.argtypes = [_vp, _vp]
    hip.clh_solver_run_frames.argtypes = [_vp, C.POINTER(StepArgs), C.c_int, _fp]
    hip.clh_solver_time_iteration_kernel.argtypes = [_vp, C.POINTER(StepArgs), C.c_int, _fp]
    hip.clh_solver_plan_info.argtypes = [_vp, C.POINTER(PlanInfo)]
    hip.clh_measure_copy.argtypes = [_vp, _vp, C.c_size_t, C.c_int, _fp]
    hip.clh_fused_plan_check.argtypes = [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int,
                                         C.c_int, C.c_int, _vp, C.POINTER(PlanInfo)]
    hip.clh_batch_create.argtypes = [C.POINTER(Config), C.POINTER(MeshDesc), C.c_int, _vp,
                                     C.POINTER(_vp)]
    hip.clh_batch_cloth_residuals.argtypes = [_vp, C.POINTER(C.c_double), C.c_int, _ip]
    hip.clh_batch_cloth_status.argtypes = [_vp, _vp, _vp, C.c_int, _ip]
    hip.clh_solver_autotune.argtypes = [_vp, C.POINTER(StepArgs), C.c_int, C.POINTER(TuneReport)]
    _dp = C.POINTER(C.c_double)
    hip.clh_ws_create.argtypes = [C.POINTER(WsDesc), C.POINTER(_vp)]
    hip.clh_ws_destroy.argtypes = [_vp]
    hip.clh_ws_set_stream.argtypes = [_vp, _vp]
    hip.clh_ws_param_count.argtypes = []
    hip.clh_ws_get_params.argtypes = [_vp, _fp, C.c_int, _fp]
    hip.clh_ws_set_params.argtypes = [_vp, _fp, C.c_int, C.c_float]
    hip.clh_ws_predict.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int]
    hip.clh_ws_features.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    hip.clh_ws_dataset_reserve.argtypes = [_vp, C.c_longlong]
    hip.clh_ws_dataset_add.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    hip.clh_ws_dataset_size.argtypes = [_vp, C.POINTER(C.c_longlong)]
    hip.clh_ws_dataset_append.argtypes = [_vp, _vp, _vp, C.c_int]
    hip.clh_ws_dataset_clear.argtypes = [_vp]
    hip.clh_ws_dataset_read.argtypes = [_vp, C.c_longlong, C.c_int, _vp, _vp]
    hip.clh_ws_fit_output_scale.argtypes = [_vp, _fp]
    hip.clh_ws_loss_grad.argtypes = [_vp, C.c_longlong, C.c_int, _fp, _fp]
    hip.clh_ws_train.argtypes = [_vp, C.c_int, C.c_int, C.c_float, _fp]
    hip.clh_ws_adam_step.argtypes = [_vp, _fp, C.c_float]
    hip.clh_ws_set_smoothing.argtypes = [_vp, C.c_int]
    hip.clh_ws_set_step_filter.argtypes = [_vp, C.c_int, C.c_float]
    hip.clh_solver_start_residual.argtypes = [_vp, C.POINTER(StepArgs), C.POINTER(C.c_double)]
    hip.clh_get_device_info.argtypes = [C.c_int, C.POINTER(DeviceInfo)]
    hip.clh_device_count.argtypes = [_ip]
    hip.clh_step_noalloc.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int,
                                     C.POINTER(StepArgs), _vp]

    core.cc_scene_spec_default.argtypes = [C.POINTER(SceneSpec)]
    core.cc_scene_spec_default.restype = None
    core.cc_scene_build.argtypes = [C.POINTER(SceneSpec)]
    core.cc_scene_build.restype = _vp
    core.cc_scene_build_obj.argtypes = [C.c_char_p, C.POINTER(SceneSpec)]
    core.cc_scene_build_obj.restype = _vp
    core.cc_batch_build.argtypes = [C.POINTER(SceneSpec), C.c_int]
    core.cc_batch_build.restype = _vp
    core.cc_scene_num_cloths.argtypes = [_vp]
    core.cc_scene_num_cloths.restype = C.c_int
    core.cc_scene_cloth_offsets.argtypes = [_vp, _vp, _vp, _vp, _vp]
    core.cc_scene_cloth_offsets.restype = C.c_int
    core.cc_scene_destroy.argtypes = [_vp]
    core.cc_scene_destroy.restype = None
    core.cc_scene_counts.argtypes = [_vp, _ip, _ip, _ip, _ip]
    core.cc_scene_counts.restype = None
    core.cc_scene_get.argtypes = [_vp, C.c_char_p, _vp, C.c_size_t]
    core.cc_scene_get.restype = C.c_int
    core.cc_scene_num_edges.argtypes = [_vp]
    core.cc_scene_num_edges.restype = C.c_int
    core.cc_scene_upload.argtypes = [_vp]
    core.cc_scene_upload.restype = C.c_int
    core.cc_scene_device.argtypes = [_vp, C.c_char_p]
    core.cc_scene_device.restype = _vp
    core.cc_scene_set_state.argtypes = [_vp, _vp, _vp]
    core.cc_scene_set_state.restype = C.c_int
    core.cc_scene_get_state.argtypes = [_vp, _vp, _vp]
    core.cc_scene_get_state.restype = C.c_int
    core.cc_pd_config_default.argtypes = [C.POINTER(PDConfig)]
    core.cc_pd_config_default.restype = None
    core.cc_pdsolver_create.argtypes = [_vp, C.POINTER(PDConfig)]
    core.cc_pdsolver_create.restype = _vp
    core.cc_pdsolver_step.argtypes = [_vp, _vp]
    core.cc_pdsolver_step.restype = None
    core.cc_pdsolver_reset.argtypes = [_vp]
    core.cc_pdsolver_reset.restype = None
    core.cc_pdsolver_native.argtypes = [_vp]
    core.cc_pdsolver_native.restype = _vp
    core.cc_pdsolver_destroy.argtypes = [_vp]
    core.cc_pdsolver_destroy.restype = None
    core.cc_scene_energy.argtypes = [_vp, C.c_float, _vp, _vp, _vp]
    core.cc_scene_energy.restype = None
    core.cc_scene_apply_pins.argtypes = [_vp]
    core.cc_scene_apply_pins.restype = None
    core.cc_scene_get_cpu_state.argtypes = [_vp, _vp, _vp]
    core.cc_scene_get_cpu_state.restype = C.c_int
    core.cc_scene_set_cpu_state.argtypes = [_vp, _vp, _vp]
    core.cc_scene_set_cpu_state.restype = C.c_int
    core.cc_cpuref_create.argtypes = [_vp, C.c_float, C.c_float, C.c_int]
    core.cc_cpuref_create.restype = _vp
    core.cc_cpuref_step.argtypes = [_vp, _vp, C.c_int]
    core.cc_cpuref_step.restype = None
    core.cc_cpuref_factorizations.argtypes = [_vp]
    core.cc_cpuref_factorizations.restype = C.c_int
    core.cc_cpuref_destroy.argtypes = [_vp]
    core.cc_cpuref_destroy.restype = None
    _hip, _core = hip, core
    return hip, core


def hip():
    return _load()[0]


def core():
    return _load()[1]


class ClothHipError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status != CLH_OK:
        msg = hip().clh_last_error().decode(errors="replace")
        raise ClothHipError(f"{what}: status {status}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    check(hip().clh_device_count(C.byref(n)), "clh_device_count")
    return n.value


def device_info(dev: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    check(hip().clh_get_device_info(dev, C.byref(info)), "clh_get_device_info")
    return info


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


class DeviceArray:
    """Device buffer from clh_alloc with host copies in and out."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        check(hip().clh_alloc(C.byref(p), self.nbytes), "clh_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(hip().clh_memcpy(self.ptr, _ptr(a), a.nbytes, CLH_H2D), "clh_memcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        check(hip().clh_memcpy(_ptr(out), self.ptr, out.nbytes, CLH_D2H), "clh_memcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().clh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def scene_spec(**kw) -> SceneSpec:
    sp = SceneSpec()
    core().cc_scene_spec_default(C.byref(sp))
    for k, v in kw.items():
        if not hasattr(sp, k):
            raise KeyError(k)
        setattr(sp, k, v)
    return sp


ARRAY_SPECS = {
    # name: (dtype, components per element, count attribute)
    "rest_pos": (np.float32, 3, "N"), "pos0": (np.float32, 3, "N"), "tris": (np.int32, 3, "T"),
    "Dm_inv": (np.float32, 4, "T"), "deltaUV": (np.float32, 4, "T"),
    "dF_dx": (np.float32, 54, "T"), "rest_area": (np.float32, 1, "T"),
    "mass": (np.float32, 1, "N"), "stretch_k": (np.float32, 1, "T"),
    "jacobi_diag": (np.float32, 1, "N"), "bend_quads": (np.int32, 4, "E"),
    "bend_w": (np.float32, 4, "E"), "bend_n": (np.float32, 1, "E"),
    "bend_k": (np.float32, 1, "E"), "pin_idx": (np.int32, 1, "P"),
    "pin_target": (np.float32, 3, "P"), "edges": (np.int32, 4, "NE"),
    "edge_len": (np.float32, 1, "NE"), "vt_ptr": (np.int32, 1, "N1"),
    "vt_idx": (np.int32, 1, "T3"), "tri_to_tris": (np.int32, 3, "T"),
}


class Scene:
    """A cloth built through the C++ API (mesh, topology, constraints, pins);
    with `specs`, a batch of cloths concatenated into one scene."""

    def __init__(self, obj_path=None, specs=None, **kw):
        if specs is not None:
            arr = (SceneSpec * len(specs))(*[scene_spec(**sp) for sp in specs])
            self.spec = arr[0]
            self.h = core().cc_batch_build(arr, len(specs))
        elif obj_path is not None:
            self.spec = scene_spec(**kw)
            self.h = core().cc_scene_build_obj(str(obj_path).encode(), C.byref(self.spec))
        else:
            self.spec = scene_spec(**kw)
            self.h = core().cc_scene_build(C.byref(self.spec))
        if not self.h:
            raise ClothHipError("cc_scene_build failed")
        self.C = core().cc_scene_num_cloths(self.h)
        offs = [np.zeros(self.C + 1, np.int32) for _ in range(4)]
        core().cc_scene_cloth_offsets(self.h, *[_ptr(o) for o in offs])
        self.vert_off, self.tri_off, self.bend_off, self.pin_off = offs
        n, t, e, p = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        core().cc_scene_counts(self.h, C.byref(n), C.byref(t), C.byref(e), C.byref(p))
        self.N, self.T, self.E, self.P = n.value, t.value, e.value, p.value
        self.NE = core().cc_scene_num_edges(self.h)
        self.uploaded = False

    def get(self, name: str) -> np.ndarray:
        dtype, comp, cnt = ARRAY_SPECS[name]
        count = {"N1": self.N + 1, "T3": self.T * 3}.get(cnt) or getattr(self, cnt)
        shape = (count, comp) if comp > 1 else (count,)
        out = np.empty(shape, dtype=dtype)
        st = core().cc_scene_get(self.h, name.encode(), _ptr(out), out.nbytes)
        if st != 0:
            raise ClothHipError(f"cc_scene_get({name}) failed ({st})")
        return out

    def upload(self):
        if core().cc_scene_upload(self.h) != 0:
            raise ClothHipError("cc_scene_upload failed: " + hip().clh_last_error().decode())
        self.uploaded = True

    def device(self, name: str):
        return core().cc_scene_device(self.h, name.encode())

    def set_state(self, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32)
        st = core().cc_scene_set_state(self.h, None if pos is None else _ptr(pos),
                                       None if vel is None else _ptr(vel))
        if st != 0:
            raise ClothHipError("cc_scene_set_state failed")

    def get_state(self):
        pos = np.empty((self.N, 3), np.float32)
        vel = np.empty((self.N, 3), np.float32)
        if core().cc_scene_get_state(self.h, _ptr(pos), _ptr(vel)) != 0:
            raise ClothHipError("cc_scene_get_state failed")
        return pos, vel

    def energy(self, pos, vel, gravity=-9.81):
        """(kinetic, potential, stretch, bend) of a host state."""
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        out = np.zeros(4, np.float64)
        core().cc_scene_energy(self.h, gravity, _ptr(pos), _ptr(vel), _ptr(out))
        return out

    def get_cpu_state(self):
        """The host state pos_cpu, vel_cpu of the CPU reference backend."""
        pos = np.empty((self.N, 3), np.float32)
        vel = np.empty((self.N, 3), np.float32)
        if core().cc_scene_get_cpu_state(self.h, _ptr(pos), _ptr(vel)) != 0:
            raise ClothHipError("cc_scene_get_cpu_state failed")
        return pos, vel

    def set_cpu_state(self, pos=None, vel=None):
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float32)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32)
        core().cc_scene_set_cpu_state(self.h, None if pos is None else _ptr(pos),
                                      None if vel is None else _ptr(vel))

    def step_args(self, stretch=True, bend=True) -> StepArgs:
        a = StepArgs()
        for f in STEP_FIELDS:
            name = f[2:]
            name = {"stretch_k": "stretch_k", "pin_idx": "pin_idx"}.get(name, name)
            setattr(a, f, self.device(name))
        if not stretch:
            a.d_stretch_k = None
        if not bend or self.E == 0:
            a.d_bend_quads = a.d_bend_w = a.d_bend_n = a.d_bend_k = None
        a.num_pinned = self.P
        return a

    def close(self):
        if self.h:
            core().cc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def config(**kw) -> Config:
    c = Config()
    check(hip().clh_config_default(C.byref(c)), "clh_config_default")
    for k, v in kw.items():
        if k == "ext_accel":
            for i in range(3):
                c.ext_accel[i] = float(v[i])
        elif not hasattr(c, k):
            raise KeyError(k)
        else:
            setattr(c, k, v)
    return c


class Solver:
    """clh_solver through the C ABI; with cloth_offsets, a batched solver
    (clh_batch_create)."""

    def __init__(self, cfg: Config, N: int, T: int, E: int, d_tris=None, d_quads=None,
                 cloth_offsets=None):
        desc = MeshDesc(N, T, E, d_tris, d_quads)
        h = _vp()
        if cloth_offsets is not None:
            offs = np.ascontiguousarray(cloth_offsets, np.int32)
            check(hip().clh_batch_create(C.byref(cfg), C.byref(desc), len(offs) - 1, _ptr(offs),
                                         C.byref(h)), "clh_batch_create")
        else:
            check(hip().clh_solver_create(C.byref(cfg), C.byref(desc), C.byref(h)),
                  "clh_solver_create")
        self.h = h.value
        self.cfg = cfg

    def start_residual(self, args: StepArgs) -> float:
        r = C.c_double(0)
        check(hip().clh_solver_start_residual(self.h, C.byref(args), C.byref(r)),
              "clh_solver_start_residual")
        return r.value

    def cloth_residuals(self, max_count=1 << 16) -> np.ndarray:
        out = np.zeros(max_count, np.float64)
        n = C.c_int(0)
        check(hip().clh_batch_cloth_residuals(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                              max_count, C.byref(n)), "cloth_residuals")
        return out[:min(n.value, max_count)]

    def cloth_status(self, max_count=1 << 16):
        """(state CLH_CLOTH_* per cloth, iterations per cloth) of the last frame."""
        n = C.c_int()
        check(hip().clh_batch_cloth_status(self.h, None, None, 0, C.byref(n)), "status")
        st = np.zeros(n.value, np.int32)
        it = np.zeros(n.value, np.int32)
        check(hip().clh_batch_cloth_status(self.h, _ptr(st), _ptr(it), n.value, C.byref(n)),
              "clh_batch_cloth_status")
        return st, it

    def autotune(self, args: StepArgs, launches: int = 20) -> TuneReport:
        rep = TuneReport()
        check(hip().clh_solver_autotune(self.h, C.byref(args), launches, C.byref(rep)),
              "clh_solver_autotune")
        return rep

    def step(self, args: StepArgs):
        check(hip().clh_solver_step(self.h, C.byref(args)), "clh_solver_step")

    def stats(self) -> StepStats:
        st = StepStats()
        check(hip().clh_solver_get_stats(self.h, C.byref(st)), "clh_solver_get_stats")
        return st

    def residual_history(self, max_count=4096) -> np.ndarray:
        out = np.zeros(max_count, np.float64)
        n = C.c_int(0)
        check(hip().clh_solver_residual_history(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                max_count, C.byref(n)), "residual_history")
        return out[:min(n.value, max_count)]

    def run_frames(self, args: StepArgs, frames: int) -> float:
        ms = C.c_float(0)
        check(hip().clh_solver_run_frames(self.h, C.byref(args), frames, C.byref(ms)),
              "clh_solver_run_frames")
        return ms.value

    def time_iteration_kernel(self, args: StepArgs, launches: int) -> float:
        ms = C.c_float(0)
        check(hip().clh_solver_time_iteration_kernel(self.h, C.byref(args), launches,
                                                     C.byref(ms)), "time_iteration_kernel")
        return ms.value

    def plan_info(self) -> PlanInfo:
        pi = PlanInfo()
        check(hip().clh_solver_plan_info(self.h, C.byref(pi)), "clh_solver_plan_info")
        return pi

    def reset(self):
        check(hip().clh_solver_reset(self.h), "clh_solver_reset")

    def close(self):
        if self.h:
            hip().clh_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CpuRefSolver:
    """CpuStretchReferenceSolver (include/stretch_reference.h) on a scene: the
    CPU reference backend, stepping the scene's host state pos_cpu / vel_cpu."""

    def __init__(self, scene: "Scene", gravity=-9.81, damping=0.0, refactor=False):
        self.scene = scene
        self.h = core().cc_cpuref_create(scene.h, gravity, damping, 1 if refactor else 0)

    def step(self, steps=1):
        core().cc_cpuref_step(self.h, self.scene.h, steps)

    def factorizations(self) -> int:
        return core().cc_cpuref_factorizations(self.h)

    def close(self):
        if self.h:
            core().cc_cpuref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pd_config(**kw) -> PDConfig:
    c = PDConfig()
    core().cc_pd_config_default(C.byref(c))
    for k, v in kw.items():
        if k == "external_accel":
            for i in range(3):
                c.external_accel[i] = float(v[i])
        else:
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
    return c


class PDSolver:
    """The C++ PDSolver class (include/pd_solver.h) driven through cloth_scene.h."""

    def __init__(self, scene: Scene, cfg: PDConfig):
        self.scene = scene
        self.h = core().cc_pdsolver_create(scene.h, C.byref(cfg))

    def step(self):
        core().cc_pdsolver_step(self.h, self.scene.h)

    def reset(self):
        core().cc_pdsolver_reset(self.h)

    def stats(self) -> StepStats:
        st = StepStats()
        check(hip().clh_solver_get_stats(core().cc_pdsolver_native(self.h), C.byref(st)),
              "clh_solver_get_stats")
        return st

    def close(self):
        if self.h:
            core().cc_pdsolver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fused_plan_check(N, tris, quads, pos, tile_verts=128, mode=MODE_PD,
                     cloth_offsets=None) -> PlanInfo:
    """Builds and validates the fused tile plan on the host (no device)."""
    tris = np.ascontiguousarray(tris, np.int32)
    quads = None if quads is None or len(quads) == 0 else np.ascontiguousarray(quads, np.int32)
    pos = np.ascontiguousarray(pos, np.float32)
    offs = None if cloth_offsets is None else np.ascontiguousarray(cloth_offsets, np.int32)
    info = PlanInfo()
    check(hip().clh_fused_plan_check(N, len(tris), 0 if quads is None else len(quads),
                                     _ptr(tris), None if quads is None else _ptr(quads),
                                     _ptr(pos), tile_verts, mode,
                                     0 if offs is None else len(offs) - 1,
                                     None if offs is None else _ptr(offs), C.byref(info)),
          "clh_fused_plan_check")
    return info


def measure_copy(nbytes: int, launches: int = 20) -> float:
    """Average ms per launch of a device copy of nbytes (copy-peak probe)."""
    src, dst = DeviceArray(nbytes), DeviceArray(nbytes)
    ms = C.c_float(0)
    try:
        check(hip().clh_measure_copy(dst.ptr, src.ptr, nbytes, launches, C.byref(ms)),
              "clh_measure_copy")
    finally:
        src.free()
        dst.free()
    return ms.value


def _dptr(a):
    """Device pointer of a DeviceArray or a raw pointer value."""
    if a is None:
        return None
    return a.ptr if isinstance(a, DeviceArray) else a


class WarmStart:
    """The learned warm start (clh_ws_*): per-vertex BF16 MFMA MLP predicting
    the converged positions of the next frame."""

    def __init__(self, N, T, d_tris, dt, gravity=-9.81, ext_accel=(0.0, 0.0, 0.0),
                 d_pin_idx=None, num_pinned=0, cloth_offsets=None, seed=1, ext_feature=False):
        d = WsDesc()
        d.num_verts, d.num_tris, d.d_tris = N, T, d_tris
        self._offs = None
        if cloth_offsets is not None and len(cloth_offsets) > 2:
            self._offs = np.ascontiguousarray(cloth_offsets, np.int32)
            d.num_cloths, d.cloth_vert_offsets = len(self._offs) - 1, _ptr(self._offs)
        else:
            d.num_cloths, d.cloth_vert_offsets = 1, None
        d.d_pin_idx, d.num_pinned = d_pin_idx, num_pinned
        d.dt, d.gravity = dt, gravity
        for i in range(3):
            d.ext_accel[i] = float(ext_accel[i])
        d.seed = seed
        d.ext_feature = 1 if ext_feature else 0
        h = _vp()
        check(hip().clh_ws_create(C.byref(d), C.byref(h)), "clh_ws_create")
        self.h = h.value
        self.N = N
        self.P = hip().clh_ws_param_count()

    @classmethod
    def for_scene(cls, scene, dt, gravity=-9.81, ext_accel=(0.0, 0.0, 0.0), seed=1,
                  ext_feature=False):
        return cls(scene.N, scene.T, scene.device("tris"), dt, gravity, ext_accel,
                   scene.device("pin_idx"), scene.P, scene.vert_off, seed, ext_feature)

    def params(self):
        out = np.zeros(self.P, np.float32)
        sc = C.c_float(0)
        check(hip().clh_ws_get_params(self.h, out.ctypes.data_as(_fp), self.P, C.byref(sc)),
              "clh_ws_get_params")
        return out, sc.value

    def set_params(self, p, scale):
        p = np.ascontiguousarray(p, np.float32)
        check(hip().clh_ws_set_params(self.h, p.ctypes.data_as(_fp), len(p), scale),
              "clh_ws_set_params")

    def predict(self, xm2, xm1, x, guess, cloth_accel=None, use_network=True):
        check(hip().clh_ws_predict(self.h, _dptr(xm2), _dptr(xm1), _dptr(x), _dptr(cloth_accel),
                                   _dptr(guess), 1 if use_network else 0), "clh_ws_predict")

    def set_step_filter(self, passes, gain):
        check(hip().clh_ws_set_step_filter(self.h, passes, gain), "clh_ws_set_step_filter")

    def set_smoothing(self, passes):
        check(hip().clh_ws_set_smoothing(self.h, passes), "clh_ws_set_smoothing")

    def features(self, xm2, xm1, x, cloth_accel=None):
        out = DeviceArray(self.N * 16 * 2)
        check(hip().clh_ws_features(self.h, _dptr(xm2), _dptr(xm1), _dptr(x), _dptr(cloth_accel),
                                    out.ptr), "clh_ws_features")
        return out.download(np.uint16, (self.N, 16))

    def dataset_reserve(self, n):
        check(hip().clh_ws_dataset_reserve(self.h, n), "clh_ws_dataset_reserve")

    def dataset_add(self, xm2, xm1, x, xnext, cloth_accel=None, cloth_weight=None):
        check(hip().clh_ws_dataset_add(self.h, _dptr(xm2), _dptr(xm1), _dptr(x), _dptr(xnext),
                                       _dptr(cloth_accel), _dptr(cloth_weight)),
              "clh_ws_dataset_add")

    def dataset_append(self, features, targets):
        F = np.ascontiguousarray(features, np.uint16)
        T = np.ascontiguousarray(targets, np.float32)
        check(hip().clh_ws_dataset_append(self.h, _ptr(F), _ptr(T), len(F)),
              "clh_ws_dataset_append")

    def dataset_size(self):
        n = C.c_longlong(0)
        check(hip().clh_ws_dataset_size(self.h, C.byref(n)), "clh_ws_dataset_size")
        return n.value

    def dataset_clear(self):
        check(hip().clh_ws_dataset_clear(self.h), "clh_ws_dataset_clear")

    def dataset_read(self, first, count):
        F = np.zeros((count, 16), np.uint16)
        T = np.zeros((count, 4), np.float32)
        check(hip().clh_ws_dataset_read(self.h, first, count, _ptr(F), _ptr(T)),
              "clh_ws_dataset_read")
        return F, T

    def fit_output_scale(self):
        sc = C.c_float(0)
        check(hip().clh_ws_fit_output_scale(self.h, C.byref(sc)), "clh_ws_fit_output_scale")
        return sc.value

    def loss_grad(self, first, count):
        g = np.zeros(self.P, np.float32)
        loss = C.c_float(0)
        check(hip().clh_ws_loss_grad(self.h, first, count, C.byref(loss),
                                     g.ctypes.data_as(_fp)), "clh_ws_loss_grad")
        return loss.value, g

    def train(self, steps, batch, lr):
        losses = np.zeros(max(steps, 1), np.float32)
        check(hip().clh_ws_train(self.h, steps, batch, lr, losses.ctypes.data_as(_fp)),
              "clh_ws_train")
        return losses[:steps]

    def adam_step(self, grad, lr):
        g = np.ascontiguousarray(grad, np.float32)
        check(hip().clh_ws_adam_step(self.h, g.ctypes.data_as(_fp), lr), "clh_ws_adam_step")

    def close(self):
        if self.h:
            hip().clh_ws_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
