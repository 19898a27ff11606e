"""CPU tests of the FEM generator's host side: descriptor validation of
ernn_fem_plan_create (rejected before any HIP call), and the host mirrors of
the reference's mesh, damage, CFL and load helpers against the FP64 oracle.

PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py.
"""
import ctypes as C

import numpy as np
import pytest

from ernn import _lib, fem
from oracle import fem_fp64 as F


def _fem_desc(loads=((5, 5, 1.0, 1),), **kw):
    L = (_lib.Source * max(1, len(loads)))()
    for k, (i, j, w, c) in enumerate(loads):
        L[k] = _lib.Source(i, j, w, c)
    d = _lib.FemDesc(nx=17, ny=13, nt=20, mode=0, dx=1e-3, dy=1.2e-3, dt=1e-8, rho=1610.0, rec_stride=5,
                     rec_scale=0.7, tile_v=0, tile_r=0, use_graph=0, blk_rw=0, blk_ks=0, loads=L,
                     n_loads=len(loads))
    for f, v in kw.items():
        setattr(d, f, v)
    return d, L


@pytest.mark.parametrize("kw,msg", [
    ({"nx": 2}, ">= 3"),
    ({"ny": 0}, ">= 3"),
    ({"nt": 1}, "nt must be"),
    ({"mode": 2}, "unknown mode"),
    ({"dx": 0.0}, "positive and finite"),
    ({"dt": float("nan")}, "positive and finite"),
    ({"rho": -1.0}, "positive and finite"),
    ({"rec_stride": 0}, "rec_stride"),
    ({"rec_scale": float("inf")}, "rec_scale"),
    ({"tile_v": 3, "tile_r": 2}, "tile (V, R)"),
    ({"tile_v": 2, "tile_r": 2}, "tile (V, R)"),
    ({"tile_v": 1, "tile_r": 3}, "tile (V, R)"),
    ({"n_loads": -1}, "load array"),
    ({"blk_rw": 4, "blk_ks": 3}, "time-blocked shape"),
    ({"blk_rw": 6, "blk_ks": 4}, "time-blocked shape"),
])
def test_fem_plan_create_rejects_bad_descriptors(kw, msg):
    d, keep = _fem_desc(**kw)
    h = C.c_void_p()
    assert _lib.lib().ernn_fem_plan_create(C.byref(d), C.byref(h)) == -1
    assert msg in _lib.lib().ernn_last_error().decode()
    assert not h.value


@pytest.mark.parametrize("load", [(17, 3, 1.0, 1), (3, -1, 1.0, 1), (3, 3, 1.0, 2), (3, 3, float("nan"), 1)])
def test_fem_plan_create_rejects_bad_loads(load):
    d, keep = _fem_desc(loads=((5, 5, 1.0, 1), load))
    h = C.c_void_p()
    assert _lib.lib().ernn_fem_plan_create(C.byref(d), C.byref(h)) == -1
    assert "load 1" in _lib.lib().ernn_last_error().decode()


def test_fem_null_arguments():
    lib = _lib.lib()
    assert lib.ernn_fem_plan_create(None, None) == -1
    info = _lib.FemInfo()
    assert lib.ernn_fem_plan_info(None, C.byref(info)) == -1
    assert lib.ernn_fem_run(None, None, None) == -1
    assert lib.ernn_fem_set_tiling(None, 1, 2) == -1
    assert lib.ernn_fem_plan_destroy(None) == -1


@pytest.mark.parametrize("nx,ny", [(9, 8), (12, 7)])
def test_host_mesh_matches_oracle(nx, ny):
    a = fem.create_structured_triangular_mesh(nx, ny, 0.04, 0.03)
    b = F.structured_mesh(nx, ny, 0.04, 0.03)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_host_damage_and_stiffness_match_oracle():
    x = np.linspace(-0.2, 0.2, 41)
    X, Y = np.meshgrid(x, x, indexing="ij")
    cfg = [{"type": "circle", "center": (0.1, 0.1), "size": 0.02},
           {"type": "square", "center": (-0.1, -0.06), "size": 0.02, "reduction_factor": 0.3}]
    m1, i1 = fem.create_damage_regions(X, Y, cfg)
    m2, i2 = F.damage_regions(X, Y, cfg)
    np.testing.assert_array_equal(m1, m2)
    assert i1["total_damaged_points"] == i2["total_damaged_points"] > 0
    assert [r["reduction_factor"] for r in i1["regions"]] == [0.5, 0.3]
    C6 = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)
    for u, v in zip(fem.apply_damage_to_stiffness(*C6, m1), F.damaged_stiffness(C6, m1)):
        np.testing.assert_array_equal(u, v)


def test_host_time_axis_and_helpers_match_oracle():
    for args in ((2001, 60e-6, 0.4 / 300, 0.4 / 300, 9011.5), (101, 60e-6, 0.4 / 300, 0.4 / 300, 9011.5)):
        dt1, nt1, t1 = fem.time_axis(*args)
        dt2, nt2, t2 = F.time_axis(*args)
        assert nt1 == nt2 and dt1 == dt2
        np.testing.assert_array_equal(t1, t2)
    t = np.linspace(0, 60e-6, 50)
    np.testing.assert_array_equal(fem.ricker_wavelet(t, 1e5, 1e-5, 1e10), F.ricker_wavelet(t, 1e5, 1e-5, 1e10))
    assert abs(fem.calculate_fwhm(5e6) - 2 * np.sqrt(np.log(2) / 5e6)) < 1e-18
    vp, vs, an = fem.calculate_wave_velocities_anisotropic(130.74e9, 11.5e9, 0, 0, 0, 14.77e9, 1610.0)
    assert (vp, vs, an) == F.wave_velocities(130.74e9, 11.5e9, 0, 0, 0, 14.77e9, 1610.0)


def test_nodal_loads_keep_fp32_nonzeros():
    x = np.linspace(-0.2, 0.2, 301)
    X, Y = np.meshgrid(x, x, indexing="ij")
    g = fem.gaussian_point_load(X, Y, 0.0, 0.0, 5e6)          # PF main's source (PF:839)
    loads = fem.nodal_loads(g)
    g32 = g.astype(np.float32)
    assert len(loads) == int((g32 != 0).sum()) > 0
    for i, j, w, c in loads:
        assert c == 1 and w == float(g32[i, j]) and w > 0
    # all of the load's mass is in the list: sum of kept weights = sum of the FP32 array
    assert abs(sum(w for _, _, w, _ in loads) - float(g32.astype(np.float64).sum())) < 1e-6 * float(g32.sum())


# ------------------------------------------------------------------ TV + Adam descriptor validation
def _tv_desc(**kw):
    d = _lib.TvAdamDesc()
    d.nx, d.ny, d.n_fields, d.log_params, d.step = 16, 12, 6, 1, 1
    d.lambda_tv, d.lr, d.beta1, d.beta2, d.eps, d.weight_decay = 1e-12, 1e-3, 0.9, 0.999, 1e-8, 0.0
    for f in range(8):
        d.lo[f], d.hi[f] = 0.0, 1e12
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,msg", [
    ({"nx": 0}, "nx and ny"),
    ({"n_fields": 0}, "n_fields"),
    ({"n_fields": 9}, "n_fields"),
    ({"step": 0}, "step"),
    ({"eps": 0.0}, "hyper-parameters"),
    ({"beta1": 1.0}, "hyper-parameters"),
    ({"lr": -1.0}, "hyper-parameters"),
    ({"lambda_tv": float("nan")}, "hyper-parameters"),
])
def test_tv_adam_rejects_bad_descriptors(kw, msg):
    n = C.c_int64(-1)
    assert _lib.lib().ernn_tv_adam_workspace_bytes(C.byref(_tv_desc(**kw)), C.byref(n)) == -1
    assert msg in _lib.lib().ernn_last_error().decode()


def test_tv_adam_workspace_size_and_bounds_check():
    n = C.c_int64(-1)
    assert _lib.lib().ernn_tv_adam_workspace_bytes(C.byref(_tv_desc()), C.byref(n)) == 0
    assert n.value == 6 * ((16 * 12 + 255) // 256) * 8          # one FP64 partial per block and field
    d = _tv_desc()
    d.lo[2], d.hi[2] = 5.0, 1.0
    assert _lib.lib().ernn_tv_adam_workspace_bytes(C.byref(d), C.byref(n)) == -1
    assert "bounds of field 2" in _lib.lib().ernn_last_error().decode()
    call = _lib.TvAdamCall()
    assert _lib.lib().ernn_tv_adam_step(C.byref(_tv_desc()), C.byref(call), None) == -1   # null workspace


def test_tv_adam_static_entry_point_rejects_fields_beyond_its_scratch():
    d = _tv_desc(nx=2048, ny=2048)                       # 6 x 16384 partials > 65536
    call = _lib.TvAdamCall()
    for f in range(6):                                     # never dereferenced: the size check comes first
        call.theta[f] = call.grad[f] = call.m[f] = call.v[f] = 256
    assert _lib.lib().ernn_tv_adam_step_static(C.byref(d), C.byref(call), None) == -3
    assert "static scratch" in _lib.lib().ernn_last_error().decode()
    call2 = _lib.TvAdamCall()
    assert _lib.lib().ernn_tv_adam_step_static(C.byref(_tv_desc()), C.byref(call2), None) == -1   # null fields
