"""Shared builders for the GPU parity tests (not collected by pytest)."""
import math

import numpy as np
import torch

import ernn
from ernn.function import ernn_traces
from oracle import ernn_fp64 as orc

RHO = 1600.0
C_ANISO = np.array([130e9, 12e9, 6e9, 3e9, 1.5e9, 5e9])
DEV = torch.device("cuda:0")


def rel_l2(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_shot(nx, ny, nt, h=0.2e-3, C=C_ANISO, sources=None, receivers=None, set_=0, f0=1e6, amp=1e10):
    dt = 0.5 * h / ernn.max_phase_velocity(C, RHO)
    if sources is None:
        sources = ernn.point_source(nx // 2, ny - 1, h, comp=1)
    wav = ernn.ricker(nt, dt, f0, 1.0 / f0, amp)
    return ernn.Shot(nx=nx, ny=ny, nt=nt, h=h, dt=dt, rho=RHO, sources=sources, wavelet=wav,
                     receivers=receivers, set=set_)


def oracle_run(shot, C_cells, stride, snapshot, G=None):
    """FP64 oracle traces (2, nf, X) and, if G is given, the per-cell gradient."""
    rec = None if snapshot else shot.receivers
    tr, st = orc.forward(C_cells, shot.h, shot.dt, RHO, shot.nt, shot.nx, shot.ny, shot.sources,
                         shot.wavelet.astype(np.float64), stride, rec, keep_states=G is not None)
    if G is None:
        return tr, None
    g = orc.adjoint_gradient(C_cells, shot.h, shot.dt, RHO, shot.nt, shot.nx, shot.ny, st,
                             G.reshape(tr.shape), stride, rec)
    return tr, g


def run_hip(survey, params=None, field=None, G=None, engine=None):
    """HIP traces (flat), and (grad_params, grad_field) fp64 if G (flat) is given."""
    eng = engine or ernn.Engine(survey, device=DEV)
    wav = torch.tensor(survey.wavelets(), device=DEV)
    p = None if params is None else torch.tensor(np.asarray(params, np.float32).reshape(-1), device=DEV)
    f = None if field is None else torch.tensor(np.asarray(field, np.float32).reshape(-1), device=DEV)
    ws = eng.acquire_workspace()
    tr = eng.forward(p, f, wav, ws)
    grads = None
    if G is not None:
        grads = eng.adjoint(p, f, wav, torch.tensor(np.asarray(G, np.float32).reshape(-1), device=DEV), ws)
        grads = tuple(None if x is None else x.cpu().numpy() for x in grads)
    torch.cuda.synchronize()
    eng.release_workspace(ws)
    return eng, tr.cpu().numpy(), grads
