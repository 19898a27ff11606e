"""GPU parity of the fused TV + Adam update (ernn_tv_adam_step through
FusedTVAdam) against the reference's own path: autograd through
clamp(exp(theta)) and total_variation_2d (DI:179-195, DI:304-322, DI:453-460)
and torch.optim.Adam (DI:478-483), on the same inputs, over several steps.

Bars: theta within 2 FP32 ulps of torch's after every step (|theta| ~ 25:
4e-6 absolute), Adam moments within 1e-5 relative L2, the total gradient left
in theta.grad within 1e-5 relative L2 of autograd's, TV within 1e-5 of torch's
FP32 sum; and the FP64 oracle's closed-form first step.
"""
import numpy as np
import pytest
import torch

import ernn
from ernn import FusedTVAdam
from oracle import tv_adam_fp64 as T

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BOUNDS = [(50e9, 250e9), (5e9, 50e9), (5e9, 50e9), (0.0, 60e9), (0.0, 20e9), (5e9, 30e9)]   # DI:265-270
C0 = [130.74e9, 11.5e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9]


def rel_l2(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm())


def make_fields(nx, ny, seed=0, log_params=True):
    g = torch.Generator().manual_seed(seed)
    th = []
    for k, (c, (lo, hi)) in enumerate(zip(C0, BOUNDS)):
        f = torch.full((1, nx, ny), float(c), dtype=torch.float64)
        f *= 1 + 0.3 * torch.randn(f.shape, generator=g, dtype=torch.float64)
        f[0, 2:6, 3:9] = float(c)                         # a uniform patch: ties, sign(0) = 0
        f[0, 0, :4] = lo * 0.5 if lo > 0 else 1e6          # below the lower bound (clamped, no gradient)
        f[0, -1, -3:] = hi * 2.0                            # above the upper bound
        f = f.clamp_min(1e6)
        th.append((f.log() if log_params else f).float().to(DEV).contiguous())
    return th


def torch_path(th0, gdata, lam, steps, log_params=True, **adam):
    params = [t.clone().requires_grad_() for t in th0]
    opt = torch.optim.Adam(params, **adam)
    out = []
    for s in range(steps):
        opt.zero_grad()
        C = [torch.clamp(torch.exp(p) if log_params else p, *b) for p, b in zip(params, BOUNDS)]
        data = sum((gd * p).sum() for gd, p in zip(gdata[s], params))   # d data / d theta = gdata
        tvs = [ernn.total_variation_2d(c) for c in C]
        loss = data + lam * sum(tvs)
        loss.backward()
        grads = [p.grad.detach().clone() for p in params]
        opt.step()
        out.append(dict(theta=[p.detach().clone() for p in params], grad=grads,
                        m=[opt.state[p]["exp_avg"].clone() for p in params],
                        v=[opt.state[p]["exp_avg_sq"].clone() for p in params],
                        tv=[float(t) for t in tvs]))
    return out


def fused_path(th0, gdata, lam, steps, log_params=True, **adam):
    params = [t.clone() for t in th0]
    opt = FusedTVAdam(params, lambda_tv=lam, bounds=BOUNDS, log_params=log_params, **adam)
    out = []
    for s in range(steps):
        for p, gd in zip(params, gdata[s]):
            p.grad = gd.clone()
        opt.step()
        torch.cuda.synchronize()
        out.append(dict(theta=[p.clone() for p in params], grad=[p.grad.clone() for p in params],
                        m=[t.clone() for t in opt.state["exp_avg"]], v=[t.clone() for t in opt.state["exp_avg_sq"]],
                        tv=opt.last_tv.cpu().tolist()))
    return out


@pytest.mark.parametrize("nx,ny,log_params,adam", [
    (64, 48, True, dict(lr=1e-2, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0)),
    (33, 70, True, dict(lr=3e-3, betas=(0.8, 0.99), eps=1e-6, weight_decay=1e-3)),
    (40, 40, False, dict(lr=1e8, betas=(0.3, 0.9), eps=1e-8, weight_decay=0.0)),   # lerp weight >= 0.5 branch
])
def test_fused_tv_adam_matches_autograd_and_torch_adam(nx, ny, log_params, adam):
    print(f"[tv-adam] {nx}x{ny} log={log_params} {adam}", flush=True)
    th0 = make_fields(nx, ny, log_params=log_params)
    lam = 1e-12 if log_params else 1e-3
    gs = torch.Generator().manual_seed(7)
    scale = 0.1 if log_params else 1e-3
    gdata = [[(scale * torch.randn((1, nx, ny), generator=gs)).to(DEV) for _ in range(6)] for _ in range(4)]
    ref = torch_path(th0, gdata, lam, 4, log_params, **adam)
    got = fused_path(th0, gdata, lam, 4, log_params, **adam)
    for s, (r, g) in enumerate(zip(ref, got)):
        for f in range(6):
            tol = 4e-6 if log_params else 2.5e-7 * float(r["theta"][f].abs().max())
            dth = float((g["theta"][f] - r["theta"][f]).abs().max())
            assert dth <= tol, (s, f, dth)
            assert rel_l2(g["grad"][f], r["grad"][f]) < 1e-5, (s, f)
            assert rel_l2(g["m"][f], r["m"][f]) < 1e-5, (s, f)
            assert rel_l2(g["v"][f], r["v"][f]) < 1e-5, (s, f)
            assert abs(g["tv"][f] - r["tv"][f]) <= 1e-5 * abs(r["tv"][f]), (s, f, g["tv"][f], r["tv"][f])
    # the fields moved and the clamped cells got no TV gradient beyond the data term
    assert float((got[-1]["theta"][0] - th0[0]).abs().max()) > 0


def test_fused_first_step_vs_fp64_oracle():
    """Step 1 against the FP64 restatement: TV subgradient through clamp/exp,
    then Adam's closed-form first move lr g / (|g| + eps)."""
    nx, ny = 30, 26
    th0 = make_fields(nx, ny, seed=3)
    gd = [(0.05 * torch.randn((1, nx, ny), generator=torch.Generator().manual_seed(11 + f))).to(DEV) for f in range(6)]
    lam, lr = 1e-12, 1e-2
    got = fused_path(th0, [gd], lam, 1, lr=lr, betas=(0.9, 0.999), eps=1e-8)[0]
    for f in range(6):
        t = th0[f][0].double().cpu().numpy()
        C, chain = T.fields(t, *BOUNDS[f])
        # the kernel forms the fields in FP32: signs and TV are taken on the FP32-rounded C
        lo, hi = BOUNDS[f]
        C32 = np.clip(np.exp(t).astype(np.float32), np.float32(lo), np.float32(hi)).astype(np.float64)
        g_tot = gd[f][0].double().cpu().numpy() + lam * T.tv_subgradient(C32) * chain
        np.testing.assert_allclose(got["grad"][f][0].cpu().numpy(), g_tot, rtol=2e-6, atol=1e-7 * np.abs(g_tot).max())
        th1, _, _ = T.adam_step(t, g_tot, 0 * t, 0 * t, 1, lr, 0.9, 0.999, 1e-8)
        np.testing.assert_allclose(got["theta"][f][0].double().cpu().numpy(), th1, rtol=0, atol=4e-6)
        assert abs(got["tv"][f] - T.tv(C32)) <= 1e-6 * T.tv(C32)   # the kernel rounds each difference to FP32


def test_di_inverter_fused_update_matches_unfused():
    """SpatiallyVaryingParameterInverter.invert(fused_update=True) follows the
    same trajectory as the autograd + torch.optim.Adam loop (DI:465-560)."""
    import math
    n, h, nt, P = 40, 1.333e-3, 160, 5
    rho = 1610.0
    dt = 0.5 * h / math.sqrt(250e9 / rho)
    res = {}
    for fused in (False, True):
        sim = ernn.SpatiallyVaryingAnisotropicSimulator((n, n), h, dt, nt, P, rho_known=rho, source_f0=100e3,
                                                        source_t0=10e-6).to(DEV)
        with torch.no_grad():
            for k, p in enumerate((sim.log_C11, sim.log_C22, sim.log_C12, sim.log_C16, sim.log_C26, sim.log_C66)):
                p.add_(0.05 * torch.randn(p.shape, generator=torch.Generator().manual_seed(k)).to(DEV))
            ox, oy = sim()
            for p in (sim.log_C11, sim.log_C66):
                p.add_(0.02)
        inv = ernn.SpatiallyVaryingParameterInverter(sim, ox.detach(), oy.detach(), secondary_time_stride=2,
                                                     lambda_tv=1e-12, verbose=False)
        inv.invert(max_epochs=6, lr=1e-3, checkpoint_interval=0, fused_update=fused)
        res[fused] = (inv.history, [p.detach().clone() for p in sim.parameters()])
    h0, p0 = res[False]
    h1, p1 = res[True]
    np.testing.assert_allclose(h1["loss"], h0["loss"], rtol=1e-4)
    np.testing.assert_allclose(h1["loss_tv"], h0["loss_tv"], rtol=1e-5)
    for a, b in zip(p1, p0):
        assert float((a - b).abs().max()) < 2e-5


def test_vector_and_scalar_paths_agree():
    """16-byte-aligned fields with ny % 4 == 0 take the 4-cells-per-thread
    kernels; fields at an odd float offset take the one-cell kernels.  The
    per-cell arithmetic is the same, so theta, m, v and grad agree bit for
    bit over several steps (the TV sums differ only in summation order)."""
    nx, ny = 36, 44
    th0 = make_fields(nx, ny, seed=5)
    gs = torch.Generator().manual_seed(9)
    gdata = [[(0.1 * torch.randn((1, nx, ny), generator=gs)).to(DEV) for _ in range(6)] for _ in range(3)]
    adam = dict(lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    aligned = fused_path(th0, gdata, 1e-12, 3, **adam)
    # the same values one float into larger buffers: every pointer is 4 bytes off 16-byte alignment
    bufs = [torch.empty(nx * ny + 1, device=DEV) for _ in range(6)]
    shifted = []
    for b, t in zip(bufs, th0):
        v = b[1:].view(1, nx, ny)
        v.copy_(t)
        shifted.append(v)
    assert all(v.data_ptr() % 16 for v in shifted)
    params = shifted
    opt = FusedTVAdam(params, lambda_tv=1e-12, bounds=BOUNDS, **adam)
    gbufs = [torch.empty(nx * ny + 1, device=DEV) for _ in range(6)]
    for s in range(3):
        for p, gb, gd in zip(params, gbufs, gdata[s]):
            gv = gb[1:].view(1, nx, ny)
            gv.copy_(gd)
            p.grad = gv
        # Adam state at an odd offset too
        if s == 0:
            mb = [torch.zeros(nx * ny + 1, device=DEV) for _ in range(12)]
            opt.state["exp_avg"] = [b[1:].view(1, nx, ny) for b in mb[:6]]
            opt.state["exp_avg_sq"] = [b[1:].view(1, nx, ny) for b in mb[6:]]
        opt.step()
        torch.cuda.synchronize()
        for f in range(6):
            assert torch.equal(params[f], aligned[s]["theta"][f]), (s, f)
            assert torch.equal(params[f].grad, aligned[s]["grad"][f]), (s, f)
            assert torch.equal(opt.state["exp_avg"][f], aligned[s]["m"][f]), (s, f)
            assert torch.equal(opt.state["exp_avg_sq"][f], aligned[s]["v"][f]), (s, f)
            assert abs(opt.last_tv[f].item() - aligned[s]["tv"][f]) <= 1e-12 * aligned[s]["tv"][f]
