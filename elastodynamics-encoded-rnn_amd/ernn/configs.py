"""The five benchmark workloads as concrete synthetic surveys.

All FP32, density 1600 kg/m^3, Ricker f0 = 1 MHz with t0 = 1 us (the
reference's f0 * t0 = 1, PI:692-693), amplitude 1e10 (PI:694), y-directed
point forces.  "Top" is row j = ny - 1 (the reference plots field.T with
origin='lower', DI:713).  dt = 0.5 h / c_max (PF:279-292) with c_max from
max(C11, C22) for grid-aligned stiffness and from the Christoffel matrix for
rotated stiffness.  Stiffness order C11, C22, C12, C16, C26, C66 (= C11, C33,
C13, C15, C35, C55).

  plain-128       128^2, h 0.2 mm, 600 steps, 1 shot at top-centre, 64 top-row receivers
  single-512      512^2, h 0.05 mm, 2000 steps, 1 shot, 256 top-row receivers, K = 50
  damage-1024     1024^2, h 0.05 mm, 3000 steps, 8 top-edge shots, 512 receivers on the top
                  and bottom rows, damage field D with 3 elliptical regions over the known
                  orthotropic background (dJ/dD only: fixed_c0), checkpoint interval chosen
                  under a 32 GB cap on the checkpointed states
  batched-uneven  24 instances, sizes {256,512,768,1024}^2, h 0.05 mm, 2000 steps,
                  constants +/-30 %, own top-edge source, 256 top-row receivers each
  aniso-2048      2048^2, h 0.05 mm, 4000 steps, 16 shots, 1024 receivers, stiffness
                  rotated 30 deg (C16, C26 != 0), damage field with 5 regions
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .problem import (Shot, Survey, bond_rotate, cfl_dt, ellipse_damage, max_phase_velocity,
                      point_source, ricker, row_receivers)

RHO = 1600.0
F0 = 1e6
T0 = 1e-6
AMP = 1e10
C_ORTHO = (130e9, 12e9, 6e9, 0.0, 0.0, 5e9)   # C11, C22(C33), C12(C13), C16, C26, C66(C55)


@dataclasses.dataclass
class Workload:
    name: str
    survey: Survey
    C_true: np.ndarray                 # (n_sets, 6) Pa
    D_true: Optional[np.ndarray] = None  # (nx, ny) damage factor of set 0 (field workloads)
    description: str = ""


def _shot(n, h, nt, dt, src_i, rec, set_=0, src_j=None):
    src_j = n - 1 if src_j is None else src_j
    return Shot(nx=n, ny=n, nt=nt, h=h, dt=dt, rho=RHO, sources=point_source(src_i, src_j, h, comp=1),
                wavelet=ricker(nt, dt, F0, T0, AMP), receivers=rec, set=set_)


def plain_128(nt: int = 600, n: int = 128) -> Workload:
    h = 0.2e-3
    dt = cfl_dt(C_ORTHO, RHO, h)
    rec = row_receivers(n, n - 1, n // 2)
    sv = Survey([_shot(n, h, nt, dt, n // 2, rec)], mode="const6")
    return Workload("plain-128", sv, np.array([C_ORTHO]), description="128^2 x 600, 1 shot, 64 receivers")


def single_512(nt: int = 2000, n: int = 512) -> Workload:
    h = 0.05e-3
    dt = cfl_dt(C_ORTHO, RHO, h)
    rec = row_receivers(n, n - 1, n // 2)
    sv = Survey([_shot(n, h, nt, dt, n // 2, rec)], mode="const6", ckpt_every=50)
    return Workload("single-512", sv, np.array([C_ORTHO]), description="512^2 x 2000, 1 shot, 256 receivers")


DAMAGE_MEM_CAP = 32 * 10**9   # bytes of checkpointed state the damage-imaging workload may take


def damage_1024(nt: int = 3000, n: int = 1024, shots: int = 8, seed: int = 2,
                mem_cap: Optional[int] = DAMAGE_MEM_CAP, fixed_c0: bool = True,
                orthotropic: Optional[bool] = None) -> Workload:
    """mem_cap: HBM bytes for the checkpointed states (Survey.ckpt_budget);
    None = the engine's default, 80 % of the free device memory.  fixed_c0:
    stiffness-degradation imaging inverts D over the known background
    constants (dJ/dD only); False also returns dJ/dC0 (the joint gradient).
    orthotropic (None: as fixed_c0): the plan takes C16 = C26 = 0 of the
    orthotropic background as given (Survey.orthotropic)."""
    h = 0.05e-3
    dt = cfl_dt(C_ORTHO, RHO, h)
    rec = np.concatenate([row_receivers(n, n - 1, n // 4), row_receivers(n, 0, n // 4)])
    sh = [_shot(n, h, nt, dt, int((k + 0.5) * n / shots), rec) for k in range(shots)]
    ortho = fixed_c0 if orthotropic is None else orthotropic
    sv = Survey(sh, mode="dfield", ckpt_budget=mem_cap, fixed_c0=fixed_c0, orthotropic=ortho)
    D = ellipse_damage(n, n, 3, np.random.default_rng(seed), semi=(20, 80), level=(0.3, 0.7), taper=5)
    return Workload("damage-1024", sv, np.array([C_ORTHO]), D,
                    description=f"{n}^2 x {nt}, {shots} shots, 512 receivers, D field 3 ellipses"
                                + (f", checkpoints under {mem_cap / 1e9:g} GB" if mem_cap else ""))


def batched_uneven(nt: int = 2000, count: int = 24, seed: int = 3,
                   sizes=(256, 512, 768, 1024)) -> Workload:
    h = 0.05e-3
    rng = np.random.default_rng(seed)
    ns = rng.choice(sizes, size=count)
    Cs, shots = [], []
    for b, n in enumerate(ns):
        n = int(n)
        C = np.array(C_ORTHO) * (1.0 + rng.uniform(-0.3, 0.3, size=6))
        C[3] = C[4] = 0.0
        dt = cfl_dt(C, RHO, h)
        src_i = int(rng.integers(n // 8, n - n // 8))
        shots.append(_shot(n, h, nt, dt, src_i, row_receivers(n, n - 1, 256), set_=b))
        Cs.append(C)
    sv = Survey(shots, mode="const6", n_sets=count)
    return Workload("batched-uneven", sv, np.array(Cs),
                    description=f"{count} instances of {{256,512,768,1024}}^2 x {nt}")


def aniso_2048(nt: int = 4000, n: int = 2048, shots: int = 16, seed: int = 4) -> Workload:
    h = 0.05e-3
    C = np.array(bond_rotate(C_ORTHO, 30.0))
    dt = 0.5 * h / max_phase_velocity(C, RHO)
    rec = row_receivers(n, n - 1, min(1024, n))   # 1024 at full size; every cell of the row on crops
    sh = [_shot(n, h, nt, dt, int((k + 0.5) * n / shots), rec) for k in range(shots)]
    sv = Survey(sh, mode="dfield")
    D = ellipse_damage(n, n, 5, np.random.default_rng(seed), semi=(20, 80), level=(0.2, 0.8), taper=5)
    return Workload("aniso-2048", sv, np.array([C]), D,
                    description=f"{n}^2 x {nt}, {shots} shots, 1024 receivers, rotated 30 deg, D 5 ellipses")


WORKLOADS = {
    "plain-128": plain_128,
    "single-512": single_512,
    "damage-1024": damage_1024,
    "batched-uneven": batched_uneven,
    "aniso-2048": aniso_2048,
}
