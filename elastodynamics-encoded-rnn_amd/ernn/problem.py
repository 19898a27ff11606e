"""Problem description: grids, shots, sources, receivers, stiffness.

Reference (paths relative to the reference repository root):
  PI = Parameter_inversion/Inversion_adam.py, DI = Damage_inversion/Inversion_adam.py,
  PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py

Stiffness is always handled in the reference's order and names
``C11, C22, C12, C16, C26, C66`` (PI:86-98); in x-z notation these are
C11, C33, C13, C15, C35, C55.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

PARAM_NAMES = ("C11", "C22", "C12", "C16", "C26", "C66")
XZ_NAMES = ("C11", "C33", "C13", "C15", "C35", "C55")
MODES = {"const6": 0, "dfield": 1, "field6": 2}
FUSE_MAX_CELLS = 1 << 20


# --------------------------------------------------------------------------
# Material helpers
# --------------------------------------------------------------------------
def cfl_dt(C: Sequence[float], rho: float, h: float, factor: float = 0.5) -> float:
    """dt = factor * h / c_max with c_max = sqrt(max(C11, C22) / rho).

    The reference's convention (PF:96-110 velocity estimate, PF:279-292 CFL
    0.5).  Exact for orthotropic stiffness aligned with the grid; use
    ``max_phase_velocity`` for rotated stiffness.
    """
    return factor * h / math.sqrt(max(C[0], C[1]) / rho)


def max_phase_velocity(C: Sequence[float], rho: float, n_dir: int = 3600) -> float:
    """Largest plane-wave speed over all directions (Christoffel matrix)."""
    C11, C22, C12, C16, C26, C66 = C
    th = np.linspace(0.0, np.pi, n_dir, endpoint=False)
    c, s = np.cos(th), np.sin(th)
    g11 = C11 * c * c + 2 * C16 * c * s + C66 * s * s
    g22 = C66 * c * c + 2 * C26 * c * s + C22 * s * s
    g12 = C16 * c * c + (C12 + C66) * c * s + C26 * s * s
    lam = 0.5 * (g11 + g22) + np.sqrt(0.25 * (g11 - g22) ** 2 + g12 ** 2)
    return float(np.sqrt(lam.max() / rho))


def bond_rotate(C: Sequence[float], theta_deg: float) -> Tuple[float, ...]:
    """In-plane rotation of a 2-D stiffness (Voigt 1, 2, 6) by theta degrees.

    C'_ijkl = R_ip R_jq R_kr R_ls C_pqrs with R the rotation by theta, i.e. a
    material whose principal (fibre) axis lies at +theta from x.  The
    invariants C11 + C22 + 2 C12 and C11 + C22 + 2 C66 are preserved.
    """
    C11, C22, C12, C16, C26, C66 = C
    voigt = {(0, 0): 0, (1, 1): 1, (0, 1): 2, (1, 0): 2}
    M = np.array([[C11, C12, C16], [C12, C22, C26], [C16, C26, C66]], dtype=np.float64)
    T = np.zeros((2, 2, 2, 2))
    for (i, j), a in voigt.items():
        for (k, l), b in voigt.items():
            T[i, j, k, l] = M[a, b]
    t = math.radians(theta_deg)
    R = np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])
    Tr = np.einsum("ip,jq,kr,ls,pqrs->ijkl", R, R, R, R, T)
    return (Tr[0, 0, 0, 0], Tr[1, 1, 1, 1], Tr[0, 0, 1, 1], Tr[0, 0, 0, 1], Tr[1, 1, 0, 1],
            Tr[0, 1, 0, 1])


# --------------------------------------------------------------------------
# Sources and receivers
# --------------------------------------------------------------------------
def ricker(nt: int, dt: float, f0: float, t0: float, amplitude: float = 1.0) -> np.ndarray:
    """Source time function of PI:28-35, sampled at t = n dt in FP64 and stored
    in FP32 as PI:210-217 does."""
    t = np.arange(nt) * dt
    q = (np.pi * f0 * (t - t0)) ** 2
    return (amplitude * (1 - 2 * q) * np.exp(-q)).astype(np.float32)


def point_source(i: int, j: int, h: float, comp: int = 1) -> List[Tuple[int, int, float, int]]:
    """A point force of unit strength on cell (i, j): force density 1/h^2."""
    return [(int(i), int(j), float(np.float32(1.0 / (h * h))), int(comp))]


def gaussian_sources(nx: int, ny: int, h: float, x0: float = 0.0, y0: float = 0.0,
                     a: float = 5e6, comp: int = 1, rel_cut: float = 1e-12):
    """The reference's spatial load (a/pi) exp(-a r^2) (PI:38-41) on the grid
    of PI:186-195, evaluated in FP32 like the reference, as a sparse list.

    Cells whose weight is below rel_cut times the peak are dropped; their
    contribution to u is below rel_cut relative to the peak force.
    """
    Lx, Ly = h * (nx - 1), h * (ny - 1)
    x = torch.linspace(-Lx / 2, Lx / 2, nx)
    y = torch.linspace(-Ly / 2, Ly / 2, ny)
    X, Y = torch.meshgrid(x, y, indexing="ij")
    g = ((a / np.pi) * torch.exp(-a * ((X - x0) ** 2 + (Y - y0) ** 2))).numpy()
    keep = np.argwhere(g > rel_cut * g.max())
    return [(int(i), int(j), float(g[i, j]), int(comp)) for i, j in keep], g


def row_receivers(nx: int, j: int, n: int) -> np.ndarray:
    """n receivers evenly spaced along x on row j (i = 0, s, 2s, ... with s = nx // n)."""
    step = max(nx // n, 1)
    i = np.arange(n) * step
    if i[-1] >= nx:
        raise ValueError(f"{n} receivers do not fit on {nx} cells")
    return np.stack([i, np.full(n, j)], axis=1).astype(np.int32)


def ellipse_damage(nx: int, ny: int, n: int, rng: np.random.Generator,
                   semi=(20, 80), level=(0.3, 0.7), taper: float = 5.0,
                   margin: int = 10) -> np.ndarray:
    """Stiffness-degradation factor D in [0, 1]: 1 outside n elliptical regions,
    D = level inside, with a smooth cosine taper of `taper` cells at the edge.
    Multiplier semantics follow the reference's damaged-element stiffness
    scaling (DF:145, DF:183-212)."""
    D = np.ones((nx, ny), dtype=np.float64)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for _ in range(n):
        a = rng.uniform(*semi)
        b = rng.uniform(*semi)
        ci = rng.uniform(margin + a, max(nx - margin - a, margin + a + 1))
        cj = rng.uniform(margin + b, max(ny - margin - b, margin + b + 1))
        ang = rng.uniform(0, np.pi)
        lv = rng.uniform(*level)
        ca, sa = np.cos(ang), np.sin(ang)
        di, dj = I - ci, J - cj
        u = (ca * di + sa * dj) / a
        v = (-sa * di + ca * dj) / b
        r = np.sqrt(u * u + v * v)                      # 1 on the boundary
        dist = (r - 1.0) * min(a, b)                    # ~ cells outside the edge
        w = np.clip(0.5 - dist / taper, 0.0, 1.0)       # 1 inside, 0 beyond the taper
        w = 0.5 - 0.5 * np.cos(np.pi * w)               # smooth step
        D = np.minimum(D, 1.0 - (1.0 - lv) * w)
    return D.astype(np.float32)


# --------------------------------------------------------------------------
# Shots and surveys
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Shot:
    """One instance: one source time function on one grid with one parameter set."""
    nx: int
    ny: int
    nt: int
    h: float
    dt: float
    rho: float
    sources: List[Tuple[int, int, float, int]]
    wavelet: np.ndarray
    receivers: Optional[np.ndarray] = None   # (n_rec, 2) int; None in snapshot mode
    set: int = 0

    @property
    def cells(self) -> int:
        return self.nx * self.ny


@dataclasses.dataclass
class Survey:
    """A batch of shots sharing a parameter mode and recording convention."""
    shots: List[Shot]
    mode: str = "const6"
    n_sets: int = 1
    rec_stride: int = 1       # frame k records u^{kP+1} (PI:294-301)
    snapshot: bool = False    # record whole fields (the reference's output)
    ckpt_every: int = 0       # 0 = choose by memory (autotune.choose_ckpt)
    # HBM bytes the checkpointed states may take when ckpt_every is 0: the
    # interval is the one within them that recomputes the fewest steps (the
    # forward pass keeps the last segment's states, so an interval of nt
    # recomputes nothing).  None = 80 % of the device's free memory when the
    # engine is built; 0 = the interval of least memory.
    ckpt_budget: Optional[int] = None
    # forward V, R, adjoint V, R (0 = default), then grid caps in blocks per CU
    # for the forward and adjoint (0 = one tile per wavefront)
    tiles: Tuple[int, ...] = (0, 0, 0, 0, 0, 0)
    use_graph: bool = False
    # Share each adjoint launch with recompute work of the previous checkpoint
    # segment: one step each, or KS each with a time-blocked adj_blocking.
    # None = automatic: on while a step has at most FUSE_MAX_CELLS cells
    # (latency bound), off above (no gain at 8 x 1024^2, and the fused
    # single-step kernel inherits the adjoint's register footprint).
    # autotune.tune_backward picks it by timing on the real problem.
    fuse: Optional[bool] = None
    # Time-blocked forward (V, RW, KS): KS steps per launch on blocks of four
    # wavefronts x RW rows x 64 V columns; (0, 0, 0) = one step per launch.
    blocking: Tuple[int, int, int] = (0, 0, 0)
    # Time-blocked adjoint (V, RW, KS), const6 / dfield; with fuse, the shape of
    # the fused time-blocked backward ((1, 4, 2), (1, 4, 4) or (1, 8, 4))
    adj_blocking: Tuple[int, int, int] = (0, 0, 0)
    # Fused time-blocked backward: rows per wavefront of the recompute half
    # (0 = the adjoint's RW; 8 beside RW = 4 with KS in {2, 4})
    adj_ff_rw: int = 0
    # dfield only: the six background constants are known (stiffness-degradation
    # imaging) and the adjoint returns dJ/dD alone.  The checkpoint bank then
    # holds (dt^2/rho) A(C0) u^m, which the forward forms anyway, instead of
    # u^m, and the adjoint reads it pointwise instead of taking a stencil of
    # u (no constant-gradient sums).  Needs the unfused backward.
    fixed_c0: bool = False
    # dfield (joint, or with fixed_c0): no bank and no recompute.  The adjoint
    # runs the forward recurrence backward in time from each segment's top
    # checkpoint (the leapfrog step is time-symmetric) and forms
    # (dt^2/rho) A(C0) u^m on the way.  Needs the unfused backward and a
    # reversed time-blocked adjoint shape (REV_BLOCK_SHAPES; (0, 0, 0) =
    # (1, 4, 4)); the checkpoint interval bounds how far each reconstruction
    # runs from its checkpoint (ckpt_every 0: REV_SEGMENT).
    reverse_u: bool = False
    # with fixed_c0: the known constants have C16 = C26 = 0 in every set (a
    # background orthotropic along the grid axes), so the damage kernels drop
    # those terms of the operator sums (the same bits, fewer instructions).
    # Constants that break the promise give wrong results; ERNN checks C_init.
    orthotropic: bool = False

    def __post_init__(self):
        if self.mode not in MODES:
            raise ValueError(f"mode must be one of {sorted(MODES)}")
        if self.fixed_c0 and self.mode != "dfield":
            raise ValueError("fixed_c0 applies to the dfield mode only")
        if self.fixed_c0 and self.fuse:
            raise ValueError("fixed_c0 needs the unfused backward (fuse False or None)")
        if self.reverse_u and self.mode != "dfield":
            raise ValueError("reverse_u applies to the dfield mode only")
        if self.reverse_u and self.fuse:
            raise ValueError("reverse_u needs the unfused backward (fuse False or None)")
        if self.orthotropic and not self.fixed_c0:
            raise ValueError("orthotropic applies to a dfield survey over known constants (fixed_c0) only")

    @property
    def fused(self) -> bool:
        if self.fixed_c0 or self.reverse_u:
            return False
        if self.fuse is not None:
            return bool(self.fuse)
        return sum(s.nx * s.ny for s in self.shots) <= FUSE_MAX_CELLS

    @property
    def nt_max(self) -> int:
        return max(s.nt for s in self.shots)

    def cell_steps(self) -> int:
        return sum(s.nx * s.ny * s.nt for s in self.shots)

    def n_frames(self, shot: Shot) -> int:
        return (shot.nt + self.rec_stride - 1) // self.rec_stride

    def wavelets(self) -> np.ndarray:
        return np.concatenate([np.asarray(s.wavelet, np.float32)[: s.nt] for s in self.shots])

    def set_grid(self, s: int) -> Tuple[int, int]:
        for sh in self.shots:
            if sh.set == s:
                return sh.nx, sh.ny
        raise ValueError(f"set {s} has no shot")
