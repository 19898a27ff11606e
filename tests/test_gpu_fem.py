"""GPU parity of the FEM generator (ernn_fem_* through ctypes) against the
FP64 restatement of the reference's solver (oracle/fem_fp64.py).

Bar: FP32 frames within 1e-5 relative L2 of the FP64 oracle (the traces bar
of the ERNN path) on small grids of a few hundred steps (measured 1.5e-6 to
2.4e-6) and within 2e-5 at the reference's full size (301 x 301 nodes, 2001
steps, PF:823-845); every tile shape gives bit-identical frames (each node's
arithmetic is fixed); graph replays equal the eager run bit for bit.

PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py.
"""
import os

import numpy as np
import pytest
import torch

from ernn import data, fem
from oracle import fem_fp64 as F

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
C_PF = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)   # PF:840-845
RHO = 1610.0


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small_case(nx=37, ny=29, Lx=0.04, Ly=0.03, nt=301, a=4e5, f0=4e5, amp=1e10, C=C_PF):
    """Grid with dx != dy, odd/even sizes, the PF source at the centre."""
    x = np.linspace(-Lx / 2, Lx / 2, nx)
    y = np.linspace(-Ly / 2, Ly / 2, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    dx, dy = Lx / (nx - 1), Ly / (ny - 1)
    vmax = F.wave_velocities(*C, RHO)[0]
    dt = 0.45 * min(dx, dy) / vmax
    g = F.gaussian_point_load(X, Y, 0.001, -0.0005, a).astype(np.float32)
    t = np.arange(nt) * dt
    w = F.ricker_wavelet(t, f0, 1.2 / f0, amp).astype(np.float32)
    return dict(nx=nx, ny=ny, Lx=Lx, Ly=Ly, dx=dx, dy=dy, dt=dt, nt=nt, g=g, w=w,
                C32=np.asarray(C, np.float32), X=X, Y=Y)


def gpu_run(c, stride=1, tile=(0, 0), D=None, use_graph=False, runs=1, blocking=(0, 0)):
    print(f"[fem] nx={c['nx']} ny={c['ny']} nt={c['nt']} stride={stride} tile={tile} blocking={blocking} "
          f"damage={D is not None} graph={use_graph}", flush=True)
    eng = fem.FEMEngine(c["nx"], c["ny"], c["dx"], c["dy"], c["nt"], c["dt"], RHO, fem.nodal_loads(c["g"]),
                        damage=D is not None, stride=stride, tile=tile, use_graph=use_graph, device=DEV,
                        blocking=blocking)
    assert eng.blocking == (tuple(blocking) if blocking[1] > 1 else (0, 0))
    outs = []
    for _ in range(runs):
        fx, fy = eng.run(c["C32"], c["w"], D)
        torch.cuda.synchronize()
        outs.append((fx.cpu().numpy(), fy.cpu().numpy()))
    eng.close()
    return outs if runs > 1 else outs[0]


def oracle_run(c, stride=1, D=None):
    fields = None
    if D is not None:
        fields = [np.asarray(D, np.float64) * np.float64(v) for v in c["C32"]]
    return F.solve(c["nx"], c["ny"], c["Lx"], c["Ly"], c["dt"], c["nt"], c["C32"].astype(np.float64), RHO,
                   c["g"], c["w"], C_fields=fields, stride=stride)


@pytest.mark.parametrize("nx,ny,Ly", [(37, 29, 0.03), (23, 131, 0.13)])   # one / three 63-column tiles
def test_fem_small_vs_fp64_oracle_all_tiles(nx, ny, Ly):
    c = small_case(nx=nx, ny=ny, Ly=Ly)
    ox, oy = oracle_run(c, stride=3)
    ref = None
    for tile in fem.FEM_SHAPES:
        gx, gy = gpu_run(c, stride=3, tile=tile)
        assert gx.shape == ox.shape == (101, nx, ny)
        assert np.abs(oy).max() > 0
        e = rel_l2(np.stack([gx, gy]), np.stack([ox, oy]))
        print(f"[fem] tile {tile}: rel L2 vs FP64 oracle {e:.3e}", flush=True)
        assert e < 1e-5, (tile, e)
        if ref is None:
            ref = (gx, gy)
        else:   # a node's arithmetic does not depend on the tiling
            np.testing.assert_array_equal(gx, ref[0])
            np.testing.assert_array_equal(gy, ref[1])


@pytest.mark.parametrize("nx,ny,Ly", [(37, 29, 0.03), (23, 131, 0.13), (70, 61, 0.05)])
def test_fem_time_blocked_equals_one_step(nx, ny, Ly):
    """KS steps per launch give the one-step kernel's frames bit for bit
    (83 samples = 81 steps: the last launch of every shape is partial)."""
    c = small_case(nx=nx, ny=ny, Ly=Ly, nt=83)
    ref = gpu_run(c, stride=3, tile=(1, 1))
    assert np.abs(ref[1]).max() > 0
    for blk in fem.FEM_BLOCK_SHAPES:
        got = gpu_run(c, stride=3, blocking=blk)
        np.testing.assert_array_equal(got[0], ref[0], err_msg=f"blocking {blk}")
        np.testing.assert_array_equal(got[1], ref[1], err_msg=f"blocking {blk}")


def test_fem_time_blocked_damage_and_graph():
    c = small_case(nx=41, ny=70, Ly=0.05, nt=64)
    mask, _ = F.damage_regions(c["X"], c["Y"], [{"type": "circle", "center": (0.006, 0.004), "size": 0.006}])
    D = torch.tensor(np.where(mask, 0.5, 1.0).astype(np.float32))
    ref = gpu_run(c, stride=2, tile=(1, 2), D=D)
    for blk in ((4, 4), (8, 8)):
        got = gpu_run(c, stride=2, D=D, blocking=blk)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
    for gx, gy in gpu_run(c, stride=2, D=D, blocking=(8, 4), use_graph=True, runs=3):
        np.testing.assert_array_equal(gx, ref[0])
        np.testing.assert_array_equal(gy, ref[1])


def test_fem_frames_boundary_and_first_steps():
    c = small_case(nt=40)
    gx, gy = gpu_run(c, stride=1)
    assert np.all(gx[0] == 0) and np.all(gy[0] == 0) and np.all(gx[1] == 0) and np.all(gy[1] == 0)
    for f in (gx, gy):   # boundary ring held at zero in every frame (PF:495-497)
        assert np.all(f[:, 0, :] == 0) and np.all(f[:, -1, :] == 0)
        assert np.all(f[:, :, 0] == 0) and np.all(f[:, :, -1] == 0)
    # u^1 = 0, so u^2 = dt^2 w[1] g / rho exactly up to FP32 rounding (PF:482-493)
    want = 0.7 * c["dt"] ** 2 * np.float64(c["w"][1]) * c["g"].astype(np.float64) / RHO
    np.testing.assert_allclose(gy[2][1:-1, 1:-1], want[1:-1, 1:-1], rtol=2e-6, atol=1e-30)
    assert np.all(gx[2] == 0)


def test_fem_damage_vs_fp64_oracle():
    c = small_case(nx=41, ny=35, nt=251)
    mask, _ = F.damage_regions(c["X"], c["Y"], [{"type": "circle", "center": (0.006, 0.004), "size": 0.005},
                                                {"type": "square", "center": (-0.008, -0.006), "size": 0.003}])
    D = np.where(mask, 0.5, 1.0).astype(np.float32)
    ox, oy = oracle_run(c, stride=5, D=D)
    gx, gy = gpu_run(c, stride=5, tile=(1, 2), D=torch.tensor(D))
    e = rel_l2(np.stack([gx, gy]), np.stack([ox, oy]))
    print(f"[fem] damage rel L2 vs FP64 oracle {e:.3e}", flush=True)
    assert e < 1e-5
    # the damage changes the wavefield (a factor of 1 everywhere would not)
    cx, cy = oracle_run(c, stride=5)
    assert rel_l2(np.stack([ox, oy]), np.stack([cx, cy])) > 1e-3
    gx1, gy1 = gpu_run(c, stride=5, tile=(1, 4), D=torch.tensor(D))
    np.testing.assert_array_equal(gx1, gx)
    np.testing.assert_array_equal(gy1, gy)


def test_fem_graph_replay_matches_eager():
    c = small_case(nt=120)
    eager = gpu_run(c, stride=2, tile=(1, 2))
    reps = gpu_run(c, stride=2, tile=(1, 2), use_graph=True, runs=3)
    for gx, gy in reps:
        np.testing.assert_array_equal(gx, eager[0])
        np.testing.assert_array_equal(gy, eager[1])


def test_fem_reference_size_vs_fp64_oracle_and_mirror_symmetry():
    """PF main (PF:823-845): 301 x 301 nodes, 0.4 m, 2001 steps of 30 ns,
    Ricker 100 kHz at 10 us, 1e10 N, Gaussian a = 5e6 at the centre, frames
    every 5 steps.  With C16 = C26 = 0 and nx odd the mesh and the load are
    mirror-symmetric in x, so ux is odd and uy even under i -> nx-1-i."""
    nx = ny = 301
    L = 0.4
    x = np.linspace(-L / 2, L / 2, nx)
    X, Y = np.meshgrid(x, x, indexing="ij")
    dt, nt, t = fem.time_axis(2001, 60e-6, L / 300, L / 300, np.sqrt(130.74e9 / RHO))
    g = fem.gaussian_point_load(X, Y, 0.0, 0.0, 5e6).astype(np.float32)
    w = fem.ricker_wavelet(t, 100e3, 10e-6, 1e10).astype(np.float32)
    c = dict(nx=nx, ny=ny, Lx=L, Ly=L, dx=L / 300, dy=L / 300, dt=dt, nt=nt, g=g, w=w,
             C32=np.asarray(C_PF, np.float32))
    gx, gy = gpu_run(c, stride=5, tile=(1, 2))
    assert gx.shape == (401, nx, ny)
    ox, oy = oracle_run(c, stride=5)
    e = rel_l2(np.stack([gx, gy]), np.stack([ox, oy]))
    print(f"[fem] 301^2 x 2001 rel L2 vs FP64 oracle {e:.3e}", flush=True)
    assert e < 2e-5
    bx, by = gpu_run(c, stride=5, blocking=(4, 4))   # time-blocked: the same frames bit for bit
    np.testing.assert_array_equal(bx, gx)
    np.testing.assert_array_equal(by, gy)
    c["C32"] = np.asarray((130.74e9, 11.50e9, 10.49e9, 0.0, 0.0, 14.77e9), np.float32)
    sx, sy = gpu_run(c, stride=5, tile=(1, 4))
    scale = np.abs(sy).max()
    assert np.abs(sx + sx[:, ::-1, :]).max() < 1e-4 * scale
    assert np.abs(sy - sy[:, ::-1, :]).max() < 1e-4 * scale


def test_fem_solver_api_and_reference_file_round_trip(tmp_path):
    """The reference's solver signature and return values (PF:196-590), and
    the observed-data file it writes (PF:900-945) read back by the loader of
    PI:516-591."""
    kw = dict(nx=41, ny=37, Lx=0.05, Ly=0.045, nt=200, t_total=4e-6, source_amplitude=1e10, source_f0=1e6,
              source_t0=1.2e-6, gaussian_param=4e6, C11=130.74e9, C22=11.5e9, C12=10.49e9, C16=30.67e9,
              C26=3.75e9, C66=14.77e9, rho=RHO, device=DEV)
    x, y, t, ux, uy, tot, smask, sidx, nodes, elements, info = fem.solve_2d_anisotropic_wave_fem_ricker_fixed(**kw)
    assert ux.shape == (41, 37, 200) and t.shape == (200,) and elements.shape == (2 * 40 * 36, 3)
    assert info["max_u_boundary"] == 0.0 and info["boundary_condition"] == "fixed"
    assert nodes.shape == (41 * 37, 2) and np.abs(uy).max() > 0
    assert tot.dtype == np.float32   # FP32 as PF:528 (squares of tiny values flush to zero there too)
    np.testing.assert_array_equal(tot, np.sqrt(ux ** 2 + uy ** 2))
    path = fem.generate_reference_data(str(tmp_path / "aniso_FEM_fixed_bc.npz"), stride=5, **kw)
    assert path.endswith("_ds5.npz")
    fx, fy, meta = data.load_reference_data(path, device=DEV)
    assert tuple(fx.shape) == (1, 40, 41, 37) and meta["primary_stride"] == 5
    np.testing.assert_array_equal(fx[0].cpu().numpy(), np.moveaxis(ux[:, :, ::5], -1, 0))
    assert float(meta["source_info"]["C16"]) == 30.67e9
    dpath = fem.generate_reference_data(str(tmp_path / "anisotropic_ricker_damage_data_ds5_FEM_fixed_bc.mat"),
                                        damage_configs=[{"type": "square", "center": (0.01, 0.0), "size": 0.005}],
                                        stride=5, **kw)
    assert dpath.endswith("_ds5_FEM_fixed_bc.mat")   # HDF5, as DF:1622-1684 writes it
    dx_, dy_, dmeta = data.load_fem_damage_data(dpath, device=DEV)
    assert dmeta["damage_mask"].sum() > 0 and int(dmeta["damage_info"]["total_damaged_points"]) > 0
    assert rel_l2(dy_.cpu().numpy(), fy.cpu().numpy()) > 1e-4


def test_fem_damage_reference_size_vs_fp64_oracle():
    """DF main (DF:1486-1545): 301 x 301 nodes, 0.4 m, 4001 samples of 30 ns,
    four damage regions with the stiffness halved (DF:877-880), frames every 5
    steps; FP32 vs the FP64 oracle, and time-blocked = one step per launch."""
    nx = ny = 301
    L = 0.4
    x = np.linspace(-L / 2, L / 2, nx)
    X, Y = np.meshgrid(x, x, indexing="ij")
    dt, nt, t = fem.time_axis(4001, 120e-6, L / 300, L / 300, np.sqrt(130.74e9 / RHO))
    g = fem.gaussian_point_load(X, Y, 0.0, 0.0, 5e6).astype(np.float32)
    w = fem.ricker_wavelet(t, 100e3, 10e-6, 1e10).astype(np.float32)
    mask, info = fem.create_damage_regions(X, Y, [
        {"type": "circle", "center": (0.1, 0.1), "size": 0.02},
        {"type": "square", "center": (-0.1, -0.1), "size": 0.02},
        {"type": "square", "center": (-0.1, -0.06), "size": 0.02},
        {"type": "square", "center": (-0.06, -0.1), "size": 0.02}])
    assert info["total_damaged_points"] > 0
    D = np.where(mask, 0.5, 1.0).astype(np.float32)
    c = dict(nx=nx, ny=ny, Lx=L, Ly=L, dx=L / 300, dy=L / 300, dt=dt, nt=nt, g=g, w=w,
             C32=np.asarray(C_PF, np.float32))
    gx, gy = gpu_run(c, stride=5, tile=(1, 1), D=torch.tensor(D))
    assert gx.shape == (801, nx, ny)
    ox, oy = oracle_run(c, stride=5, D=D)
    e = rel_l2(np.stack([gx, gy]), np.stack([ox, oy]))
    print(f"[fem] DF main 301^2 x 4001 rel L2 vs FP64 oracle {e:.3e}", flush=True)
    assert e < 2e-5
    bx, by = gpu_run(c, stride=5, D=torch.tensor(D), blocking=(2, 4))
    np.testing.assert_array_equal(bx, gx)
    np.testing.assert_array_equal(by, gy)


def _engine_run(c, loads, stride=1, blocking=(0, 0)):
    eng = fem.FEMEngine(c["nx"], c["ny"], c["dx"], c["dy"], c["nt"], c["dt"], RHO, loads, stride=stride,
                        blocking=blocking, device=DEV)
    fx, fy = eng.run(c["C32"], c["w"])
    torch.cuda.synchronize()
    out = (fx.cpu().numpy(), fy.cpu().numpy())
    eng.close()
    return out


def test_fem_edge_cases():
    """One interior node; no time step at all (nt = 2); x-component loads;
    a frame every step; duplicate loads merged into one weight."""
    print("[fem] edge cases", flush=True)
    # 3 x 3 nodes: the single interior node carries the load
    c = small_case(nx=3, ny=3, Lx=0.004, Ly=0.004, nt=30)
    g = np.zeros((3, 3), np.float32)
    g[1, 1] = 2.5e5
    c["g"] = g
    ox, oy = oracle_run(c, stride=1)
    for blk in ((0, 0), (2, 2)):
        gx, gy = _engine_run(c, fem.nodal_loads(g), stride=1, blocking=blk)
        # C16 != 0 couples the y load into ux as well
        assert np.abs(oy).max() > 0 and rel_l2(np.stack([gx, gy]), np.stack([ox, oy])) < 1e-5
    # nt = 2: frame 0 only, zero
    c2 = small_case(nt=2)
    fx, fy = _engine_run(c2, fem.nodal_loads(c2["g"]), stride=5)
    assert fx.shape == (1, c2["nx"], c2["ny"]) and np.all(fx == 0) and np.all(fy == 0)
    # loads on the x component (the kernel takes either; the reference uses y)
    c3 = small_case(nt=120)
    loads_x = fem.nodal_loads(c3["g"], comp=0)
    ox, oy = F.solve(c3["nx"], c3["ny"], c3["Lx"], c3["Ly"], c3["dt"], c3["nt"], c3["C32"].astype(np.float64), RHO,
                     np.zeros_like(c3["g"]), c3["w"], stride=1, load_x=c3["g"])
    for blk in ((0, 0), (4, 4)):
        gx, gy = _engine_run(c3, loads_x, stride=1, blocking=blk)
        assert np.abs(ox).max() > 0
        assert rel_l2(np.stack([gx, gy]), np.stack([ox, oy])) < 1e-5
    # duplicates: w1 + w2 on one node equals one load of the FP32 sum
    i, j = c3["nx"] // 2, c3["ny"] // 2
    w1, w2 = np.float32(1.5e5), np.float32(0.7e5)
    a = _engine_run(c3, [(i, j, float(w1), 1), (i, j, float(w2), 1)], stride=3)
    b = _engine_run(c3, [(i, j, float(np.float32(w1 + w2)), 1)], stride=3)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
