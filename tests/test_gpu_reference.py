"""HIP path vs the reference's own method: the op-for-op PyTorch restatement
(oracle/plain_torch.py) with autograd, run on the same GPU.  Bars: FP32 traces
within 1e-5 relative L2 of the plain path, gradients within 1e-4 relative of
its autograd gradient; the FP64 oracle error is printed per case."""
import json
import math
import os

import numpy as np
import pytest
import torch

import ernn
from ernn import configs
from ernn.function import ernn_traces
from oracle import ernn_fp64 as orc
from oracle import plain_torch as pt
from gpu_helpers import DEV, rel_l2

pytestmark = pytest.mark.gpu
FD_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fd_single512_crop.json")


def _rel(a, b):
    return float((a - b).double().norm() / b.double().norm())


def _grad_rel(a, b):
    a = torch.as_tensor(a).double().cpu(); b = torch.as_tensor(b).double().cpu()
    return float((a - b).abs().max() / b.abs().max())


# ------------------------------------------------------------------ PI mirror
def test_cfrp_simulator_matches_plain_path_forward_and_gradient():
    """CFRPWaveSimulatorRicker (PI:132-307) + CFRPParameterInverterRicker loss
    (PI:358-391) vs autograd through the plain path."""
    n, h, nt, P = 64, 1.333e-3, 400, 5
    rho = 1610.0
    dt = 0.5 * h / math.sqrt(160.06e9 / rho)
    kw = dict(domain_shape=(n, n), h=h, dt=dt, nt=nt, primary_time_stride=P, source_component='y',
              source_f0=100e3, source_t0=10e-6, source_amplitude=1e10, rho_known=rho,
              gaussian_param=5e6, source_position=(0.0, 0.0))
    init = dict(C11_init=90e9, C22_init=20e9, C12_init=30e9, C16_init=20e9, C26_init=10e9, C66_init=10e9)
    sim = ernn.CFRPWaveSimulatorRicker(**kw, **init).to(DEV)
    # observed: plain path at the true constants (PI:699-706)
    true = [160.06e9, 70.86e9, 48.04e9, 9.48e9, 16.27e9, 37.64e9]
    sig = sim.source_signal
    g = sim.gaussian_dist
    rho_t = torch.tensor(rho, device=DEV)
    with torch.no_grad():
        ox, oy = pt.forward([torch.tensor(c, device=DEV) for c in true], h, dt, nt, P, sig, g, rho_t,
                            hoist_weights=True)
    # HIP module
    ux, uy = sim()
    assert ux.shape == (1, (nt + P - 1) // P, n, n)
    loss = pt.loss_subsampled(ux, uy, ox, oy, 2, 8)
    loss.backward()
    g_hip = torch.stack([p.grad for p in (sim.log_C11, sim.log_C22, sim.log_C12, sim.log_C16,
                                          sim.log_C26, sim.log_C66)])
    # plain path with the same parameters
    logs = [torch.log(torch.tensor(v, dtype=torch.float32)).to(DEV).requires_grad_() for v in init.values()]
    bounds = [(1e9, 10000e9)] * 6
    C = pt.material_params_log(logs, bounds)
    px, py = pt.forward(C, h, dt, nt, P, sig, g, rho_t)
    lp = pt.loss_subsampled(px, py, ox, oy, 2, 8)
    lp.backward()
    g_ref = torch.stack([l.grad for l in logs])
    e_tr = _rel(torch.stack([ux, uy]), torch.stack([px, py]).detach())
    e_g = _grad_rel(g_hip, g_ref)
    print(f"PI mirror: traces rel-L2 vs plain {e_tr:.2e}, loss {loss.item():.6e} vs {lp.item():.6e}, "
          f"grad rel {e_g:.2e}")
    assert e_tr < 1e-5
    assert abs(loss.item() - lp.item()) / lp.item() < 1e-5
    assert e_g < 1e-4


# ------------------------------------------------------------------ DI mirror
def test_spatially_varying_simulator_matches_plain_path():
    """SpatiallyVaryingAnisotropicSimulator (DI:198-393) + the damage loss with
    TV (DI:441-463) vs autograd through the per-cell plain path (DI:100-176,
    force transposed as the reference does; centred source on a square grid)."""
    n, h, nt, P = 48, 1.333e-3, 300, 5
    rho = 1610.0
    dt = 0.5 * h / math.sqrt(250e9 / rho)
    sim = ernn.SpatiallyVaryingAnisotropicSimulator((n, n), h, dt, nt, P, rho_known=rho,
                                                    source_f0=100e3, source_t0=10e-6).to(DEV)
    with torch.no_grad():   # perturb the initial fields so every term of the gradient is exercised
        for k, p in enumerate((sim.log_C11, sim.log_C22, sim.log_C12, sim.log_C16, sim.log_C26, sim.log_C66)):
            p.add_(0.05 * torch.randn(p.shape, generator=torch.Generator().manual_seed(k)).to(DEV))
    sig, g = sim.source_signal, sim.gaussian_dist
    rho_t = torch.tensor(rho, device=DEV)
    # observed: damaged true fields (half stiffness in a block)
    with torch.no_grad():
        Ct = [f.detach().clone() for f in sim.get_material_params()[:6]]
        for f in Ct:
            f[:, 10:20, 25:35] *= 0.5
        ox, oy = pt.forward(Ct, h, dt, nt, P, sig, g, rho_t, per_cell=True, transpose_force=True,
                            hoist_weights=True)
    ux, uy = sim()
    fields = sim.get_material_params()[:6]
    loss = pt.loss_damage(ux, uy, ox, oy, 2, fields, 1e-21)
    loss.backward()
    params = (sim.log_C11, sim.log_C22, sim.log_C12, sim.log_C16, sim.log_C26, sim.log_C66)
    g_hip = [p.grad.clone() for p in params]
    logs = [p.detach().clone().requires_grad_() for p in params]
    C = pt.material_params_log(logs, [sim.C11_bounds, sim.C22_bounds, sim.C12_bounds, sim.C16_bounds,
                                      sim.C26_bounds, sim.C66_bounds])
    px, py = pt.forward(C, h, dt, nt, P, sig, g, rho_t, per_cell=True, transpose_force=True)
    lp = pt.loss_damage(px, py, ox, oy, 2, C, 1e-21)
    lp.backward()
    e_tr = _rel(torch.stack([ux, uy]), torch.stack([px, py]).detach())
    e_g = [rel_l2(a.cpu().numpy(), l.grad.cpu().numpy()) for a, l in zip(g_hip, logs)]
    print(f"DI mirror: traces rel-L2 {e_tr:.2e}, grad rel-L2 per field {['%.1e' % e for e in e_g]}")
    assert e_tr < 1e-5
    assert abs(loss.item() - lp.item()) / lp.item() < 1e-5
    assert max(e_g) < 1e-4


# ------------------------------------------------------------------ full sizes
def _plain_receiver_run(wl, C, G=None):
    """Plain path (reference op sequence) for a one-shot receiver survey."""
    s = wl.survey.shots[0]
    sig = torch.tensor(s.wavelet, device=DEV).unsqueeze(0)
    spatial = torch.zeros(s.nx, s.ny, device=DEV)
    for (i, j, w, comp) in s.sources:
        spatial[i, j] += w
    Ct = [torch.tensor(float(c), device=DEV, requires_grad=G is not None) for c in C]
    ux, uy = pt.forward(Ct, s.h, s.dt, s.nt, 1, sig, spatial, torch.tensor(s.rho, device=DEV),
                        hoist_weights=True)
    r = torch.as_tensor(s.receivers, device=DEV).long()
    tr = torch.stack([ux[0][:, r[:, 0], r[:, 1]], uy[0][:, r[:, 0], r[:, 1]]])
    grad = None
    if G is not None:
        (tr * G).sum().backward()
        grad = torch.stack([c.grad for c in Ct])
    return tr.detach(), grad


@pytest.mark.parametrize("name", ["plain-128"])
def test_full_size_traces_and_gradient_vs_plain_autograd(name):
    """(single-512 and damage-1024 at full size, with every kernel shape, are in
    test_gpu_fullsize.py.)"""
    wl = configs.WORKLOADS[name]()
    C = wl.C_true[0] * np.array([1.05, 0.97, 1.1, 1.0, 1.0, 0.93]) + np.array([0, 0, 0, 0.4e9, -0.3e9, 0])
    eng = ernn.Engine(wl.survey, device=DEV)
    wav = torch.tensor(wl.survey.wavelets(), device=DEV)
    p = torch.tensor(C, dtype=torch.float32, device=DEV, requires_grad=True)
    tr = ernn_traces(eng, params=p, wavelets=wav)
    G = torch.randn(tr.shape, generator=torch.Generator().manual_seed(7)).to(DEV)
    (tr * G).sum().backward()
    ref, gref = _plain_receiver_run(wl, C, G.view(eng.shot_traces(tr, 0).shape))
    e_tr = _rel(eng.shot_traces(tr.detach(), 0), ref)
    e_g = _grad_rel(p.grad, gref)
    print(f"{name}: traces rel-L2 vs plain FP32 {e_tr:.2e}, grad rel vs autograd {e_g:.2e}")
    assert e_tr < 1e-5
    assert e_g < 1e-4


def test_single_512_crop_vs_fp64_oracle_and_finite_differences():
    """single-512 physics (h 0.05 mm, dt, C, 2000 steps) on a 128^2 crop problem:
    HIP traces vs the FP64 oracle; the HIP gradient vs the FP64 discrete adjoint
    and vs FP64 central differences at step 1e-4 relative (both from the
    committed fixture tests/golden/fd_single512_crop.json); and vs central
    differences of the HIP forward itself at step 1e-3 C (J sums 2000 FP32
    steps, so a smaller step lets the FP32 noise of J(C +- d) dominate the
    difference, while the O(d^2) truncation of a 1e-3 step stays below 1e-5 of
    the gradient)."""
    fx = json.load(open(FD_FIXTURE))
    wl = configs.single_512(nt=2000, n=128)
    s = wl.survey.shots[0]
    C = wl.C_true[0] * np.array([1.05, 0.97, 1.1, 1.0, 1.0, 0.93]) + np.array([0, 0, 0, 0.4e9, -0.3e9, 0])
    C32 = np.asarray(C, np.float32).astype(np.float64)
    assert np.array_equal(C32, np.array(fx["C"])) and fx["nt"] == s.nt and fx["grid"] == [s.nx, s.ny]
    rng = np.random.default_rng(fx["g_seed"])
    G = rng.standard_normal((2, s.nt, len(s.receivers)))
    ref, _ = orc.forward(C32, s.h, s.dt, s.rho, s.nt, s.nx, s.ny, s.sources, s.wavelet.astype(np.float64),
                         1, s.receivers)
    gref, fd64 = np.array(fx["adjoint"]), np.array(fx["fd"])
    eng = ernn.Engine(wl.survey, device=DEV)
    wav = torch.tensor(wl.survey.wavelets(), device=DEV)
    Gt = torch.tensor(G.ravel(), dtype=torch.float32, device=DEV)
    ws = eng.acquire_workspace()
    p = torch.tensor(C, dtype=torch.float32, device=DEV)
    tr = eng.forward(p, None, wav, ws)
    gp, _ = eng.adjoint(p, None, wav, Gt, ws)
    gp = gp.cpu().numpy()
    e_tr = rel_l2(tr.cpu().numpy().reshape(ref.shape), ref)
    e_g = np.abs(gp - gref).max() / np.abs(gref).max()
    e_fd64 = np.abs(gp - fd64).max() / np.abs(fd64).max()
    # central differences of the HIP forward, J = <traces, G>
    fd = []
    for k in range(6):
        d = 1e-3 * max(abs(C[k]), 5e9)
        cp = C.copy(); cp[k] += d
        cm = C.copy(); cm[k] -= d
        jp = float((eng.forward(torch.tensor(cp, dtype=torch.float32, device=DEV), None, wav, ws).double() * Gt.double()).sum())
        jm = float((eng.forward(torch.tensor(cm, dtype=torch.float32, device=DEV), None, wav, ws).double() * Gt.double()).sum())
        fd.append((jp - jm) / (2 * d))
    fd = np.array(fd)
    e_fd = np.abs(gp - fd).max() / np.abs(gp).max()
    print(f"single-512 crop 128^2 x 2000: traces rel-L2 vs FP64 {e_tr:.2e}, grad vs FP64 adjoint {e_g:.2e}, "
          f"vs FP64 central FD (1e-4) {e_fd64:.2e}, vs FP32 central FD (1e-3) {e_fd:.2e}")
    assert e_tr < 1e-5
    assert e_g < 1e-4
    assert e_fd64 < 1e-4
    assert e_fd < 2e-2      # FP32 forward noise / step bounds the FD accuracy


def test_aniso_2048_crop_vs_fp64_oracle():
    """aniso-2048 physics (30-degree fibre frame, so C15 and C35 are nonzero;
    damage field of five ellipses) on a 256^2 crop at 500 steps, two shots:
    HIP traces, dJ/dC0 and dJ/dD vs the FP64 oracle's forward and discrete
    adjoint on the fields C_p = D C0_p."""
    from gpu_helpers import run_hip
    wl = configs.aniso_2048(nt=500, n=256, shots=2)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    nx, ny = sv.shots[0].nx, sv.shots[0].ny
    D = np.asarray(wl.D_true, np.float32).reshape(nx, ny)
    rng = np.random.default_rng(12)
    G = rng.standard_normal(sum(2 * s.nt * len(s.receivers) for s in sv.shots))
    eng, tr, (gC, gD) = run_hip(sv, params=C0, field=D, G=G)
    D64, C64 = D.astype(np.float64), C0.astype(np.float64)
    cells = [D64 * c for c in C64]
    e_tr, gcell = 0.0, 0.0
    for b, s in enumerate(sv.shots):
        lo, hi = eng.trace_off[b], eng.trace_off[b + 1]
        ref, st = orc.forward(cells, s.h, s.dt, s.rho, s.nt, nx, ny, s.sources, s.wavelet.astype(np.float64),
                              1, s.receivers, keep_states=True)
        e_tr = max(e_tr, rel_l2(tr[lo:hi].reshape(ref.shape), ref))
        gcell = gcell + orc.adjoint_gradient(cells, s.h, s.dt, s.rho, s.nt, nx, ny, st,
                                             G[lo:hi].reshape(ref.shape), 1, s.receivers)
        del st
    gC_ref, gD_ref = orc.dfield_gradient(gcell, C64, D64)
    e_gc = np.abs(gC - gC_ref).max() / np.abs(gC_ref).max()
    e_gd = rel_l2(gD.reshape(nx, ny), gD_ref)
    print(f"aniso-2048 crop 256^2 x 500, 2 shots: traces rel-L2 vs FP64 {e_tr:.2e}, dJ/dC0 max rel {e_gc:.2e}, "
          f"dJ/dD rel-L2 {e_gd:.2e}")
    assert e_tr < 1e-5
    assert e_gc < 1e-4
    assert e_gd < 1e-4
    eng.close()
