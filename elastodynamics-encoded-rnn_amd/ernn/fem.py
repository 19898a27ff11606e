"""FEM observed-data generator on the HIP path.

The reference generates its "real" observed data with an explicit P1 finite
element solver on a checkerboard triangle mesh:
  PF = Parameter_inversion/FEM_generate_data.py  (constant stiffness)
  DF = Damage_inversion/FEM_generate_data.py     (stiffness halved in damage regions)
This module mirrors its functions with the same names, arguments and return
values; the time loop (PF:479-513, ~100 tensor ops and six index_add_
scatters per step) runs as one gather kernel launch per step in
libernn_hip.so (ernn_fem_* in include/ernn.h).  There is no PyTorch fallback.

  create_structured_triangular_mesh     PF:117-189
  gaussian_point_load, ricker_wavelet   PF:51-88
  calculate_fwhm, wave velocities       PF:91-114
  create_damage_regions                 DF:121-180
  apply_damage_to_stiffness             DF:183-212
  solve_2d_anisotropic_wave_fem_ricker_fixed         PF:196-590
  solve_2d_anisotropic_wave_fem_ricker_fixed_damage  DF:770-1224
  generate_reference_data               PF:823-956, DF:1486-1700 (HDF5 as the reference writes it, or .npz)
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .data import save_reference_data

FEM_SHAPES = ((1, 1), (1, 2), (1, 4), (1, 8))   # (V, R): R node rows x 63 columns per wavefront
FEM_BLOCK_SHAPES = ((2, 2), (2, 4), (4, 2), (4, 4), (8, 2), (8, 4), (8, 8))   # (RW, KS): KS steps per launch
STIFFNESS_NAMES = ("C11", "C22", "C12", "C16", "C26", "C66")
DEFAULT_DAMAGE = [{"type": "circle", "center": (0.1, 0.1), "size": 0.02, "reduction_factor": 0.5}]   # DF:856-866


# --------------------------------------------------------------------------
# Reference helpers (host, NumPy)
# --------------------------------------------------------------------------
def gaussian_point_load(x, y, x0, y0, a=150):
    """(a / pi) exp(-a ((x - x0)^2 + (y - y0)^2)) (PF:51-66)."""
    return (a / np.pi) * np.exp(-a * ((x - x0) ** 2 + (y - y0) ** 2))


def ricker_wavelet(t, f0, t0, amplitude=1.0):
    """amplitude (1 - 2 (pi f0 (t - t0))^2) exp(-(pi f0 (t - t0))^2) (PF:69-88)."""
    s = (np.pi * f0 * (t - t0)) ** 2
    return amplitude * (1 - 2 * s) * np.exp(-s)


def calculate_fwhm(a):
    """2 sqrt(ln 2 / a) (PF:91-93)."""
    return 2 * np.sqrt(np.log(2) / a)


def calculate_wave_velocities_anisotropic(C11, C22, C12, C16, C26, C66, rho):
    """(sqrt(max(C11, C22) / rho), sqrt(C66 / rho), |C11 - C22| / (C11 + C22)) (PF:96-114)."""
    return np.sqrt(max(C11, C22) / rho), np.sqrt(C66 / rho), abs(C11 - C22) / (C11 + C22)


def create_structured_triangular_mesh(nx, ny, Lx, Ly):
    """(x, y, elements, boundary_nodes, dx, dy) of the checkerboard mesh
    (PF:117-189): node = j nx + i; cell (i, j) is split along n0-n2 when
    i + j is even, along n1-n3 when odd, lower triangle first."""
    dx = Lx / (nx - 1)
    dy = Ly / (ny - 1)
    x = np.linspace(-Lx / 2, Lx / 2, nx)
    y = np.linspace(-Ly / 2, Ly / 2, ny)
    j, i = (a.ravel() for a in np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij"))
    n0 = j * nx + i
    corners = np.stack([n0, n0 + 1, n0 + nx + 1, n0 + nx], 1)           # n0, n1, n2, n3
    even = ((i + j) % 2 == 0)[:, None]
    elements = np.empty((2 * n0.size, 3), np.int64)
    elements[0::2] = np.where(even, corners[:, [0, 1, 2]], corners[:, [0, 1, 3]])
    elements[1::2] = np.where(even, corners[:, [0, 2, 3]], corners[:, [1, 2, 3]])
    k = np.arange(nx * ny)
    on_edge = (k % nx == 0) | (k % nx == nx - 1) | (k // nx == 0) | (k // nx == ny - 1)
    return x, y, elements, k[on_edge].astype(np.int64), dx, dy


def create_damage_regions(X, Y, damage_configs):
    """(damage_mask, damage_info) for circles and squares (DF:121-180)."""
    mask = np.zeros_like(X, dtype=bool)
    regions = []
    for k, cfg in enumerate(damage_configs):
        cx, cy = cfg["center"]
        if cfg["type"] == "circle":
            region = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2) <= cfg["size"]
        elif cfg["type"] == "square":
            region = (np.abs(X - cx) <= cfg["size"]) & (np.abs(Y - cy) <= cfg["size"])
        else:
            raise ValueError(f"Unknown damage type: {cfg['type']}")
        mask = mask | region
        regions.append({"id": k, "type": cfg["type"], "center": cfg["center"], "size": cfg["size"],
                        "reduction_factor": cfg.get("reduction_factor", 0.5), "n_points": int(region.sum())})
    return mask, {"regions": regions, "total_damaged_points": int(mask.sum()),
                  "damage_percentage": 100 * mask.sum() / mask.size}


def apply_damage_to_stiffness(C11_base, C22_base, C12_base, C16_base, C26_base, C66_base, damage_mask,
                              reduction_factor=0.5):
    """Six (nx, ny) stiffness fields, multiplied by reduction_factor on the
    mask (DF:183-212)."""
    out = []
    for c in (C11_base, C22_base, C12_base, C16_base, C26_base, C66_base):
        f = np.full(damage_mask.shape, float(c))
        f[damage_mask] *= reduction_factor
        out.append(f)
    return tuple(out)


def time_axis(nt: int, t_total: float, dx: float, dy: float, vmax: float):
    """(dt, nt, t) with the CFL adjustment of PF:280-296: when vmax dt /
    min(dx, dy) > 0.5, nt = int(t_total / dt_stable) + 1 and dt = t_total /
    (nt - 1), dt_stable = 0.5 min(dx, dy) / vmax (the recomputed dt is
    >= dt_stable, so the number can end slightly above 0.5, as there)."""
    dt = t_total / (nt - 1)
    if vmax * dt / min(dx, dy) > 0.5:
        nt = int(t_total / (0.5 * min(dx, dy) / vmax)) + 1
        dt = t_total / (nt - 1)
    return dt, nt, np.linspace(0, t_total, nt)


def nodal_loads(load: np.ndarray, comp: int = 1) -> List[Tuple[int, int, float, int]]:
    """The nonzero FP32 entries of a nodal load array (nx, ny) as (i, j, w,
    comp) (PF:322 casts the Gaussian to FP32; zero entries add nothing)."""
    g = np.asarray(load, np.float32)
    ii, jj = np.nonzero(g)
    return [(int(a), int(b), float(g[a, b]), comp) for a, b in zip(ii, jj)]


# --------------------------------------------------------------------------
# Engine over the C ABI
# --------------------------------------------------------------------------
class FEMEngine:
    """Plan + workspace of one FEM run configuration on one device.

    loads: (i, j, w, comp) nodal loads (w: the Gaussian at the node, comp 1 =
    y as in PF:482-484).  run(C, wavelet[, damage]) returns (frames_x,
    frames_y) of shape (ceil(nt / stride), nx, ny): frame k = scale * u^{k stride}."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, nt: int, dt: float, rho: float,
                 loads: Sequence[Tuple[int, int, float, int]], damage: bool = False, stride: int = 1,
                 scale: float = 0.7, tile: Tuple[int, int] = (0, 0), use_graph: bool = False,
                 device: Optional[torch.device] = None, blocking: Tuple[int, int] = (0, 0)):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise RuntimeError("FEMEngine runs on a ROCm device only (there is no CPU path)")
        self.nx, self.ny, self.nt, self.stride = int(nx), int(ny), int(nt), int(stride)
        self.damage = bool(damage)
        arr = (_lib.Source * max(1, len(loads)))()
        for k, (i, j, w, c) in enumerate(loads):
            arr[k] = _lib.Source(int(i), int(j), float(w), int(c))
        d = _lib.FemDesc(nx=self.nx, ny=self.ny, nt=self.nt, mode=_lib.FEM_DAMAGE if damage else _lib.FEM_CONST,
                         dx=float(dx), dy=float(dy), dt=float(dt), rho=float(rho), rec_stride=self.stride,
                         rec_scale=float(scale), tile_v=int(tile[0]), tile_r=int(tile[1]),
                         use_graph=int(bool(use_graph)), blk_rw=int(blocking[0]), blk_ks=int(blocking[1]),
                         loads=arr, n_loads=len(loads))
        lib = _lib.lib()
        self._plan = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.ernn_fem_plan_create(C.byref(d), C.byref(self._plan)), "ernn_fem_plan_create")
        self._refresh()
        self._ws = torch.empty(int(self.info.workspace_bytes), dtype=torch.uint8, device=self.device)
        if self._ws.data_ptr() % 256:
            raise RuntimeError("workspace allocation is not 256-byte aligned")

    def _refresh(self) -> None:
        self.info = _lib.FemInfo()
        _lib.check(_lib.lib().ernn_fem_plan_info(self._plan, C.byref(self.info)), "ernn_fem_plan_info")

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib().ernn_fem_plan_destroy(self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_frames(self) -> int:
        return int(self.info.n_frames)

    @property
    def tile(self) -> Tuple[int, int]:
        return int(self.info.tile_v), int(self.info.tile_r)

    @property
    def blocking(self) -> Tuple[int, int]:
        """(RW, KS) of the time-blocked launches; (0, 0) = one step per launch."""
        return int(self.info.blk_rw), int(self.info.blk_ks)

    def _inputs(self, C6, wavelet, damage_field):
        f32 = dict(dtype=torch.float32, device=self.device)
        C6 = torch.as_tensor(C6, **f32).reshape(-1).contiguous()
        w = torch.as_tensor(wavelet, **f32).reshape(-1).contiguous()
        if C6.numel() != 6:
            raise ValueError("C must hold the six constants C11, C22, C12, C16, C26, C66")
        if w.numel() != self.nt:
            raise ValueError(f"wavelet must hold nt = {self.nt} time factors, got {w.numel()}")
        D = None
        if self.damage:
            if damage_field is None:
                raise ValueError("a damage engine needs the per-node factor field")
            D = torch.as_tensor(damage_field, **f32).contiguous()
            if tuple(D.shape) != (self.nx, self.ny):
                raise ValueError(f"damage field must be ({self.nx}, {self.ny}), got {tuple(D.shape)}")
        return C6, w, D

    def _call(self, C6, w, D, fx, fy) -> _lib.FemCall:
        return _lib.FemCall(workspace=self._ws.data_ptr(), params=C6.data_ptr(),
                            field=D.data_ptr() if D is not None else None, wavelet=w.data_ptr(),
                            frames_x=fx.data_ptr(), frames_y=fy.data_ptr())

    def run(self, C6, wavelet, damage_field=None, out=None) -> Tuple[torch.Tensor, torch.Tensor]:
        """Enqueue the whole run on the current stream; returns the frames."""
        C6, w, D = self._inputs(C6, wavelet, damage_field)
        shape = (self.n_frames, self.nx, self.ny)
        fx, fy = out if out is not None else (torch.empty(shape, dtype=torch.float32, device=self.device),
                                              torch.empty(shape, dtype=torch.float32, device=self.device))
        if tuple(fx.shape) != shape or tuple(fy.shape) != shape or not (fx.is_contiguous() and fy.is_contiguous()):
            raise ValueError(f"frame arrays must be contiguous {shape}")
        c = self._call(C6, w, D, fx, fy)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_fem_run(self._plan, C.byref(c), C.c_void_p(_lib.stream_handle(self.device))),
                       "ernn_fem_run")
        self._keep = (C6, w, D)   # inputs stay alive until the stream has used them
        return fx, fy

    def time_steps(self, steps: int, C6, wavelet, damage_field=None, out=None) -> float:
        """Milliseconds of steps n = 1 .. min(steps, nt - 2) (hipEvents; synchronizes)."""
        C6, w, D = self._inputs(C6, wavelet, damage_field)
        shape = (self.n_frames, self.nx, self.ny)
        fx, fy = out if out is not None else (torch.empty(shape, dtype=torch.float32, device=self.device),
                                              torch.empty(shape, dtype=torch.float32, device=self.device))
        c = self._call(C6, w, D, fx, fy)
        ms = C.c_float()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_fem_time_steps(self._plan, int(steps), C.byref(c),
                                                      C.c_void_p(_lib.stream_handle(self.device)), C.byref(ms)),
                       "ernn_fem_time_steps")
        return float(ms.value)

    def set_tiling(self, v: int, r: int) -> None:
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_fem_set_tiling(self._plan, int(v), int(r)), "ernn_fem_set_tiling")
        self._refresh()

    def set_blocking(self, rw: int, ks: int) -> None:
        """KS steps per launch on blocks of 4 x RW x 64 nodes (ks <= 1: off)."""
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().ernn_fem_set_blocking(self._plan, int(rw), int(ks)), "ernn_fem_set_blocking")
        self._refresh()

    def tune(self, C6, wavelet, damage_field=None, steps: int = 100, rounds: int = 2,
             shapes: Sequence[Tuple[int, int]] = FEM_SHAPES,
             block_shapes: Sequence[Tuple[int, int]] = FEM_BLOCK_SHAPES) -> Dict[Tuple, float]:
        """Time each one-step tile shape and each time-blocked shape over
        `steps` steps (best of `rounds`) and keep the fastest; returns
        {("tile", V, R) or ("block", RW, KS): ms}."""
        shape = (self.n_frames, self.nx, self.ny)
        out = (torch.empty(shape, dtype=torch.float32, device=self.device),
               torch.empty(shape, dtype=torch.float32, device=self.device))
        times = {}
        self.set_blocking(0, 0)
        for v, r in shapes:
            self.set_tiling(v, r)
            times[("tile", v, r)] = min(self.time_steps(steps, C6, wavelet, damage_field, out) for _ in range(rounds))
        best_tile = min((k for k in times if k[0] == "tile"), key=times.get)
        self.set_tiling(*best_tile[1:])
        for rw, ks in block_shapes:
            self.set_blocking(rw, ks)
            times[("block", rw, ks)] = min(self.time_steps(steps, C6, wavelet, damage_field, out)
                                           for _ in range(rounds))
        best = min(times, key=times.get)
        self.set_blocking(*(best[1:] if best[0] == "block" else (0, 0)))
        return times


# --------------------------------------------------------------------------
# The reference's solver functions
# --------------------------------------------------------------------------
def _solve(nx, ny, Lx, Ly, nt, t_total, source_amplitude, source_f0, source_t0, gaussian_param, C11, C22, C12,
           C16, C26, C66, rho, source_position, damage_configs, device, record_stride, tile, use_graph):
    vp, vs, aniso = calculate_wave_velocities_anisotropic(C11, C22, C12, C16, C26, C66, rho)
    x, y, elements, boundary_nodes, dx, dy = create_structured_triangular_mesh(nx, ny, Lx, Ly)
    X, Y = np.meshgrid(x, y, indexing="ij")
    dt, nt, t = time_axis(nt, t_total, dx, dy, max(vp, vs))                        # PF:280-296
    x0, y0 = source_position
    gauss = gaussian_point_load(X, Y, x0, y0, gaussian_param)                       # PF:298-301
    wavelet = np.asarray([ricker_wavelet(tn, source_f0, source_t0, source_amplitude) for tn in t], np.float32)
    damage_mask = damage_info = None
    D = None
    if damage_configs is not None:
        damage_mask, damage_info = create_damage_regions(X, Y, damage_configs)     # DF:871-880
        D = np.where(damage_mask, 0.5, 1.0).astype(np.float32)                     # reduction 0.5 (DF:877-880)
    eng = FEMEngine(nx, ny, dx, dy, nt, dt, rho, nodal_loads(gauss), damage=D is not None,
                    stride=record_stride, scale=0.7, tile=tile, use_graph=use_graph, device=device)
    fx, fy = eng.run(np.array([C11, C22, C12, C16, C26, C66], np.float32), wavelet, D)
    # (nx, ny, frames) like the reference's (nx, ny, nt) arrays (PF:517-526)
    u_x = np.moveaxis(fx.cpu().numpy(), 0, -1)
    u_y = np.moveaxis(fy.cpu().numpy(), 0, -1)
    eng.close()
    total = np.sqrt(u_x ** 2 + u_y ** 2)
    bi, bj = boundary_nodes % nx, boundary_nodes // nx
    max_u_boundary = float(max(np.abs(u_x[bi, bj, -1]).max(), np.abs(u_y[bi, bj, -1]).max()))   # PF:531-537
    nodes = np.stack([x[np.arange(nx * ny) % nx], y[np.arange(nx * ny) // nx]], 1)                # PF:560-565
    source_mask = gauss > 0.01 * np.max(gauss)
    source_indices = [int(np.argmin(np.abs(y - y0))) * nx + int(np.argmin(np.abs(x - x0)))]
    source_info = {"source_amplitude": source_amplitude, "source_f0": source_f0, "source_t0": source_t0,
                   "gaussian_param": gaussian_param, "fwhm": calculate_fwhm(gaussian_param),
                   "C11": C11, "C22": C22, "C12": C12, "C16": C16, "C26": C26, "C66": C66, "rho": rho,
                   "vp_approx": vp, "vs_approx": vs, "anisotropy_factor": aniso, "boundary_condition": "fixed",
                   "max_u_boundary": max_u_boundary}
    if damage_configs is not None:
        source_info.update({"damage_mask": damage_mask, "damage_info": damage_info, "damage_configs": damage_configs})
    t_rec = t[::record_stride]
    return x, y, t_rec, u_x, u_y, total, source_mask, source_indices, nodes, elements, source_info


def solve_2d_anisotropic_wave_fem_ricker_fixed(nx=301, ny=301, Lx=0.4, Ly=0.4, nt=1001, t_total=30e-6,
                                               source_amplitude=1e13, source_f0=170e3, source_t0=6e-6,
                                               gaussian_param=5e8, C11=16.5e10, C22=6.2e10, C12=5.0e10,
                                               C16=0.0, C26=0.0, C66=3.96e10, rho=7100.0,
                                               source_position=(0.0, 0.0), device=None, record_stride=1,
                                               tile=(0, 0), use_graph=False):
    """PF:196-590 on the HIP path: (x, y, t, u_x, u_y, total_displacement,
    source_mask, source_indices, nodes, elements, source_info) with u_x, u_y
    of shape (nx, ny, ceil(nt / record_stride)) holding 0.7 u at t[::record_stride]
    (record_stride = 1: every step, the reference's (nx, ny, nt))."""
    return _solve(nx, ny, Lx, Ly, nt, t_total, source_amplitude, source_f0, source_t0, gaussian_param, C11, C22,
                  C12, C16, C26, C66, rho, source_position, None, device, record_stride, tile, use_graph)


def solve_2d_anisotropic_wave_fem_ricker_fixed_damage(nx=301, ny=301, Lx=0.4, Ly=0.4, nt=1001, t_total=30e-6,
                                                      source_amplitude=1e13, source_f0=170e3, source_t0=6e-6,
                                                      gaussian_param=5e8, C11=16.5e10, C22=6.2e10, C12=5.0e10,
                                                      C16=0.0, C26=0.0, C66=3.96e10, rho=7100.0,
                                                      source_position=(0.0, 0.0), damage_configs=None,
                                                      device=None, record_stride=1, tile=(0, 0),
                                                      use_graph=False):
    """DF:770-1224 on the HIP path (the damage variant of the same solver;
    damage_configs None = one circle of radius 2 cm at (10, 10) cm, DF:856-866;
    every damaged node's stiffness is halved whatever reduction_factor the
    configuration names, as DF:877-880 does)."""
    return _solve(nx, ny, Lx, Ly, nt, t_total, source_amplitude, source_f0, source_t0, gaussian_param, C11, C22,
                  C12, C16, C26, C66, rho, source_position,
                  DEFAULT_DAMAGE if damage_configs is None else damage_configs, device, record_stride, tile,
                  use_graph)


def generate_reference_data(path: str, damage_configs=None, stride: int = 5, **solver_kwargs) -> str:
    """Run the solver and write the observed-data file of PF:900-945 /
    DF:1619-1684 (ux, uy as (nt_ds, nx, ny) float32, x, y, t[::stride],
    source_info, and for damage runs damage_mask and damage_info with its
    region_<i> subgroups): an HDF5 file in the reference's layout for a .mat /
    .h5 name, an .npz archive for an .npz name; data.load_reference_data /
    load_fem_damage_data read both.  The stride goes into the name as
    "_ds<stride>" when the name lacks it."""
    damaged = damage_configs is not None
    solver = solve_2d_anisotropic_wave_fem_ricker_fixed_damage if damaged else solve_2d_anisotropic_wave_fem_ricker_fixed
    kw = dict(solver_kwargs, record_stride=stride)
    if damaged:
        kw["damage_configs"] = damage_configs
    x, y, t, u_x, u_y, _, _, _, _, _, info = solver(**kw)
    if f"_ds{stride}" not in os.path.basename(path):
        root, ext = os.path.splitext(path)
        path = f"{root}_ds{stride}{ext}"
    src = {k: v for k, v in info.items() if k not in ("damage_mask", "damage_info", "damage_configs")}
    dmg = None
    if damaged:
        di = info["damage_info"]
        dmg = {"total_damaged_points": di["total_damaged_points"], "damage_percentage": di["damage_percentage"]}
        for r in di["regions"]:
            for k in ("type", "center", "size", "reduction_factor", "n_points"):
                dmg[f"region_{r['id']}/{k}"] = np.asarray(r[k])
    return save_reference_data(path, np.moveaxis(u_x, -1, 0), np.moveaxis(u_y, -1, 0), x.astype(np.float32),
                               y.astype(np.float32), np.asarray(t, np.float32), source_info=src,
                               damage_mask=info["damage_mask"] if damaged else None, damage_info=dmg)
