"""The C-ABI boundary: the libraries load, export every symbol their headers
declare, and reject invalid arguments without touching a device."""
import ctypes as C
import os
import re

import pytest

import clothhip as ch
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_clothhip_exports_every_declared_symbol():
    names = declared(os.path.join(ROOT, "include", "clothhip.h"), "clh_")
    assert len(names) >= 20
    assert sorted(names) == sorted(ch.CLH_SYMBOLS)
    lib = ch.hip()
    for n in names:
        assert hasattr(lib, n), n


def test_scene_api_exports():
    names = declared(os.path.join(ROOT, "cloth-cuda_amd", "host", "cloth_scene.h"), "cc_")
    lib = ch.core()
    for n in names:
        assert hasattr(lib, n), n


def test_oracle_exports():
    names = declared(os.path.join(ROOT, "oracle", "pd_oracle.h"), "or_")
    lib = po.lib()
    for n in names:
        assert hasattr(lib, n), n


def test_core_exports_reference_cpp_api():
    # the C++ API of the reference (mangled names) is in libcloth_core
    out = os.popen("nm -D --defined-only " + ch.LIB_CORE + " | c++filt").read()
    for sym in ["generate_square_cloth(int, int, float, int, ClothMesh&)",
                "MeshTopology::build(ClothMesh const&)",
                "ClothMesh::precompute_rest_state(float)", "ClothMesh::upload_to_gpu()",
                "ClothMesh::load_obj(", "SimConstraints::build_bend(",
                "SimConstraints::precompute_jacobi_diag(ClothMesh const&, float)",
                "Constraints::pin_top_row(ClothMesh const&, int)",
                "PDSolver::PDSolver(PDSolverConfig const&, ClothMesh const&, SimConstraints const&)",
                "PDSolver::step(ClothMesh&, SimConstraints const&, Constraints const&)",
                "PDSolver::reset()", "launch_apply_constraints_kernel("]:
        assert sym in out, sym


def test_invalid_arguments_rejected_without_device():
    lib = ch.hip()
    assert lib.clh_alloc(None, 16) == 1
    assert lib.clh_memcpy(None, None, 4, ch.CLH_H2D) == 1
    assert lib.clh_memcpy(C.c_void_p(1), C.c_void_p(1), 4, 9) == 1
    assert lib.clh_apply_constraints(None, None, None, None, 3) == 1
    assert lib.clh_config_default(None) == 1
    cfg = ch.config(mode=7)
    desc = ch.MeshDesc(4, 2, 0, None, None)
    h = C.c_void_p()
    assert lib.clh_solver_create(C.byref(cfg), C.byref(desc), C.byref(h)) == 1
    assert b"mode" in lib.clh_last_error()
    cfg = ch.config(dt=0.0)
    assert lib.clh_solver_create(C.byref(cfg), C.byref(desc), C.byref(h)) == 1
    cfg = ch.config(max_iterations=-1)
    assert lib.clh_solver_create(C.byref(cfg), C.byref(desc), C.byref(h)) == 1
    assert lib.clh_solver_step(None, None) == 1
    assert lib.clh_solver_destroy(None) == 0


def test_config_defaults_mirror_reference():
    # include/pd_solver.h:5-14 of the reference
    c = ch.config()
    assert c.max_iterations == 50 and c.use_chebyshev == 1
    assert abs(c.rho - 0.9) < 1e-7 and abs(c.gravity + 9.8) < 1e-6 and abs(c.dt - 0.01) < 1e-9
    assert c.damping == 0.0 and c.mode == ch.MODE_REF_CUDA
    p = ch.pd_config()
    assert p.max_iterations == 50 and abs(p.tolerance - 1e-4) < 1e-9 and p.mode == 0


def test_noalloc_refuses_without_device():
    # clh_step_noalloc: sizes beyond its static array, tolerance frames and
    # traces are refused before any device work
    lib = ch.hip()
    args = ch.StepArgs()
    cfg = ch.config(mode=ch.MODE_PD, max_iterations=5, dt=1 / 60)
    assert lib.clh_step_noalloc(C.byref(cfg), 2_000_000, 4_000_000, 6_000_000, C.byref(args),
                                None) == ch.ERR_UNSUPPORTED
    assert b"bytes of scratch" in lib.clh_last_error()
    cfg.tolerance = 1e-5
    assert lib.clh_step_noalloc(C.byref(cfg), 100, 162, 0, C.byref(args), None) == \
        ch.ERR_UNSUPPORTED
    cfg.tolerance = 0.0
    args.d_trace = C.c_void_p(16)
    assert lib.clh_step_noalloc(C.byref(cfg), 100, 162, 0, C.byref(args), None) == \
        ch.ERR_UNSUPPORTED
    assert lib.clh_step_noalloc(None, 100, 162, 0, C.byref(args), None) == ch.ERR_INVALID
    assert lib.clh_step_noalloc(C.byref(cfg), -1, 162, 0, C.byref(args), None) == ch.ERR_INVALID
    # a fixed frame of a fitting mesh with null state pointers: invalid, no launch
    args.d_trace = None
    assert lib.clh_step_noalloc(C.byref(cfg), 100, 162, 0, C.byref(args), None) == ch.ERR_INVALID
