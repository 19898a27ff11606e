"""Benchmark: fwd+adjoint Gcell-steps/s and s/Adam-iter, HIP adjoint vs autograd, 1 GPU.

One step = one Adam iteration of the waveform-misfit inversion
(Parameter_inversion/Inversion_adam.py:436-450): forward recurrence over all
shots, MSE on receiver traces, gradient (checkpointed recompute + reverse-time
adjoint), Adam update.  Default workload: damage-1024 (1024^2 grid, 3000
steps, 8 top-edge shots, 512 receivers on the top and bottom rows, damage
field D(x) of three ellipses scaling all six constants; the gradient covers
the six constants and the 1 Mi-cell D field; checkpoint interval chosen under
a 32 GB cap on the checkpointed states), FP32, synthetic data (observed
traces = the same solver at the true constants and D + 1 % RMS seeded noise).

Also reported on the same line:
  autograd_baseline  the same iteration through the reference's own method
                     (op-for-op F.pad/F.conv2d restatement + autograd) on the
                     same GPU, in the reference-faithful form (per-step stencil
                     weight tensors, snapshot history) and with the weights
                     hoisted; the faster one is the denominator of `speedup`.
  roofline           dominant kernel's algorithmic bytes per launch / its average
                     launch duration (hipEvents on the launch stream, back-to-back
                     launches), against 8 TB/s HBM.  Algorithmic bytes per cell
                     and step: each state read once and written once, per-cell
                     fields read once (forward 24 / 28 / 48 B for six constants /
                     damage field / six fields, adjoint 32 / 36 / 56 B); a launch
                     of KS steps is credited KS times that, the recompute half of
                     a fused launch nothing (reported beside it).  Time-blocked
                     launches move fewer bytes than that (states stay in
                     registers between their steps), so the fraction can exceed
                     1; `traffic` is the measured HBM bytes per launch.
  cpu_baseline       the op-for-op CPU port (oracle/plain_torch.py, autograd) on
                     the host cores over a bounded window of steps.
  hbm_budget_run     workloads with a checkpoint cap: the same iteration with the
                     interval that 80 % of the free HBM gives (fewer recomputed
                     steps); `value` stays the capped figure.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "elastodynamics-encoded-rnn_amd"))

METRIC = "fwd+adjoint Gcell-steps/s & s/Adam-iter: HIP adjoint vs autograd, 1 GPU"
HBM_PEAK_GBS = 8000.0
# Vector-ALU issue peak in wavefront instructions per second: 256 CUs x 2.4
# GHz x one wave64 instruction per CU per cycle (four SIMD16 units, four
# cycles each).  The kernels are built without packed FP32, so this is also
# their FP32 FMA peak (78.6 TFLOP/s, half the packed 157.3).
VALU_PEAK_WIPS = 256 * 2.4e9
# Algorithmic FP32 bytes per cell and step (the minimum of one pass over the
# grid per step): the forward reads u^n, u^{n-1} and writes u^{n+1} (2
# components each); the adjoint reads lam^{m+1}, lam^{m+2}, u^m and writes
# lam^m; per-cell fields add one read each (damage field: D; six fields: the
# six C_p).  Gradient accumulators are amortised and not counted.
BYTES_PER_CELL = {
    ("const6", 0): 24, ("const6", 1): 32,
    ("dfield", 0): 28, ("dfield", 1): 36,
    ("field6", 0): 48, ("field6", 1): 56,
}
# The reversed-u adjoint (reverse_u plans) also carries u backward: per step it
# reads u^m, u^{m+1} and writes u^{m-1} on top of the damage adjoint's 36 B.
REV_BYTES_PER_CELL = 36 + 24


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _ranked(times, n=4):
    """The n fastest time-blocked shapes of a tuner's {shape: ms per step}."""
    best = sorted((t, s) for s, t in times.items() if s != (0, 0, 0))[:n]
    return "ranked " + ", ".join(f"{s} {t * 1e3:.2f}" for t, s in best)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="damage-1024")
    ap.add_argument("--nt", type=int, default=0, help="override the step count (profiling runs)")
    ap.add_argument("--joint-c0", type=int, default=0,
                    help="damage-1024: also invert the six background constants (joint gradient; the "
                         "workload's own setting inverts D over known constants)")
    ap.add_argument("--reverse-u", type=int, default=1,
                    help="damage-field workloads: 1 = the adjoint reconstructs u backward from the checkpoints "
                         "(no bank, no recompute); 0 = banked states with segment recompute")
    ap.add_argument("--ckpt-every", type=int, default=0,
                    help="fix the checkpoint interval (profiling runs; 0 = the workload's memory budget)")
    ap.add_argument("--graph", type=int, default=1, help="replay passes as captured hipGraphs")
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--tiles", default="", help="fixed tiles 'Vf,Rf,Va,Ra,Bf,Ba' (skips autotuning)")
    ap.add_argument("--blocking", default="auto",
                    help="time-blocked forward 'V,RW,KS', '0' for one step per launch, or 'auto' (autotuned)")
    ap.add_argument("--adj-blocking", default="auto",
                    help="time-blocked unfused adjoint 'V,RW,KS', '0', or 'auto' (autotuned)")
    ap.add_argument("--ff-rw", default="auto",
                    help="fused backward: rows per wavefront of the recompute half (the adjoint's RW, or 8 beside "
                         "RW = 4), or 'auto' (autotuned)")
    ap.add_argument("--fuse", default="auto",
                    help="backward schedule: '1' fused single-step, '0' unfused (time-blocked when tuned), "
                         "'auto' (small grids: timed both ways)")
    ap.add_argument("--baseline", type=int, default=1, help="time the autograd plain path on the GPU")
    ap.add_argument("--cpu", type=int, default=1, help="time the CPU port")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--baseline-window", type=int, default=2000,
                    help="steps of one shot timed for the autograd baseline (scaled to the survey)")
    ap.add_argument("--hbm-run", type=int, default=1,
                    help="workloads with a checkpoint memory cap: also time the interval that 80 %% of the free "
                         "HBM gives (reported beside the capped value)")
    ap.add_argument("--per-instance", type=int, default=1,
                    help="batched surveys: time every instance alone for p50 / max")
    ap.add_argument("--fem-n", type=int, default=301,
                    help="FEM workloads: nodes per side (the reference's 301; larger grids keep the same square and "
                         "physics, dt from the same CFL rule)")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# The reference's own method on the GPU (autograd plain path), self-contained
# restatement of PI:44-106 and PI:240-307 used only as the timing baseline.
# ---------------------------------------------------------------------------
def _weights(torch, h, device, dtype):
    s = h ** (-2)
    a = torch.tensor([[[[0.0, 0.0, 0.0], [1.0, -2.0, 1.0], [0.0, 0.0, 0.0]]]], device=device, dtype=dtype)
    b = torch.tensor([[[[0.0, 1.0, 0.0], [0.0, -2.0, 0.0], [0.0, 1.0, 0.0]]]], device=device, dtype=dtype)
    c = torch.tensor([[[[0.25, 0.0, -0.25], [0.0, 0.0, 0.0], [-0.25, 0.0, 0.25]]]], device=device, dtype=dtype)
    return s * a, s * b, s * c


def autograd_traces(torch, shot, C, faithful, D=None, nsteps=None, fields=None, stride=0):
    """Receiver traces (2, nsteps, n_rec) through the reference's op sequence.

    C: six 0-dim tensors; D: optional (nx, ny) damage field, in which case the
    per-cell operator of DI:100-176 runs on the fields D * C_p; fields: six
    (nx, ny) stiffness fields for that operator instead (C unused).  stride > 0:
    whole-field frames u^(kP+1) every stride steps (DI:367-393), returned as
    (X, Y), each (1, n_frames, nx, ny)."""
    import torch.nn.functional as F
    dev = (fields if fields is not None else C)[0].device
    nx, ny = shot.nx, shot.ny
    nt = shot.nt if nsteps is None else nsteps
    sig = torch.tensor(shot.wavelet, device=dev)
    g = torch.zeros(1, nx, ny, device=dev)
    for (i, j, w, comp) in shot.sources:
        g[0, i, j] += w
    rho = torch.tensor(shot.rho, device=dev)
    if not stride:
        ri = torch.as_tensor(shot.receivers[:, 0], device=dev).long()
        rj = torch.as_tensor(shot.receivers[:, 1], device=dev).long()
    if fields is not None:   # per-cell fields on the transposed view (DI:143-148)
        C11, C22, C12, C16, C26, C66 = (f.unsqueeze(0).transpose(1, 2) for f in fields)
    elif D is None:
        C11, C22, C12, C16, C26, C66 = C
    else:
        C11, C22, C12, C16, C26, C66 = ((D * c).unsqueeze(0).transpose(1, 2) for c in C)
    W = None if faithful else _weights(torch, shot.h, dev, torch.float32)
    ux1 = torch.zeros(1, nx, ny, device=dev); uy1 = torch.zeros_like(ux1)
    ux2 = torch.zeros_like(ux1); uy2 = torch.zeros_like(ux1)
    hx, hy = [], []
    conv = lambda f, k: F.conv2d(F.pad(f, (1, 1, 1, 1)), k)
    for n in range(nt):
        fy = sig[n] * g
        d2dx2, d2dy2, d2dxdy = W if W is not None else _weights(torch, shot.h, dev, torch.float32)
        ux = ux1.transpose(1, 2).unsqueeze(1); uy = uy1.transpose(1, 2).unsqueeze(1)
        uxx = conv(ux, d2dx2).squeeze(1); uxy_ = conv(ux, d2dy2).squeeze(1)
        uyx = conv(uy, d2dx2).squeeze(1); uyy = conv(uy, d2dy2).squeeze(1)
        uxxy = conv(ux, d2dxdy).squeeze(1); uyxy = conv(uy, d2dxdy).squeeze(1)
        Lx = (C11 * uxx + C66 * uxy_ + (C12 + C66) * uyxy + 2 * C16 * uxxy + C16 * uyx + C26 * uyy).transpose(1, 2)
        Ly = (C66 * uyx + C22 * uyy + (C12 + C66) * uxxy + C16 * uxx + C26 * uxy_ + 2 * C26 * uyxy).transpose(1, 2)
        Lx = (Lx + torch.zeros_like(fy)) / rho
        Ly = (Ly + fy) / rho
        uxn = 2 * ux1 - ux2 + shot.dt ** 2 * Lx
        uyn = 2 * uy1 - uy2 + shot.dt ** 2 * Ly
        ux2, uy2, ux1, uy1 = ux1, uy1, uxn, uyn
        if stride:
            if n % stride == 0:
                hx.append(ux1); hy.append(uy1)
        elif faithful:
            hx.append(ux1); hy.append(uy1)
        else:
            hx.append(ux1[0, ri, rj]); hy.append(uy1[0, ri, rj])
    if stride:
        return torch.stack(hx, dim=1), torch.stack(hy, dim=1)
    if faithful:
        X = torch.stack(hx, dim=1); Y = torch.stack(hy, dim=1)
        return torch.stack([X[0][:, ri, rj], Y[0][:, ri, rj]])
    return torch.stack([torch.stack(hx), torch.stack(hy)])


def time_autograd(torch, wl, theta0, scale, D0, observed, window, learn_c=True):
    """s/Adam-iter of the plain path (shots one at a time, as the reference).
    learn_c False: the constants are known (no gradient), only D is inverted.

    If the survey is longer than `window` steps or has several shots, one shot
    is timed over `window` steps and the per-step cost is scaled to
    n_shots * nt steps (per-step cost is constant in n).  The window is also
    capped so that the autograd tape (about 48 saved fields per step) fits in
    half of the device memory."""
    res = {}
    shots = wl.survey.shots
    s0 = shots[0]
    tape_step = 48 * 4 * s0.nx * s0.ny
    mem_cap = min(torch.cuda.get_device_properties(0).total_memory // 2, int(0.8 * torch.cuda.mem_get_info()[0]))
    W = max(10, min(window, s0.nt, mem_cap // tape_step))
    scale_f = sum(s.nt for s in shots) / W
    for faithful in (True, False):
        theta = torch.nn.Parameter(theta0.clone(), requires_grad=learn_c)
        params = [theta] if learn_c else []
        D = None
        if D0 is not None:
            D = torch.nn.Parameter(D0.clone())
            params.append(D)
        opt = torch.optim.Adam(params, lr=1e-2)

        def it():
            opt.zero_grad()
            C = list(theta[0] * scale[0])
            tr = autograd_traces(torch, s0, C, faithful, D=D, nsteps=W)
            loss = torch.mean((tr - observed[0][:, :W]) ** 2)
            loss.backward()
            opt.step()
        it()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it()
        torch.cuda.synchronize()
        key = "faithful" if faithful else "hoisted"
        res[key] = (time.perf_counter() - t0) * scale_f
        log(f"autograd plain path ({key}): {res[key]:.3f} s/iter"
            + (f" (extrapolated from {W} steps of one shot)" if scale_f != 1 else ""))
    res["window_steps"] = W
    res["extrapolated"] = scale_f != 1
    return res


# ---------------------------------------------------------------------------
# CPU baseline: the op-for-op port (oracle/plain_torch.py) on the host cores
# ---------------------------------------------------------------------------
def cpu_baseline(torch, wl, budget_s):
    sys.path.insert(0, REPO)
    from oracle import plain_torch as pt
    s = wl.survey.shots[0]
    threads = torch.get_num_threads()

    import numpy as np
    dfield = wl.survey.mode == "dfield"

    def run(nsteps):
        sig = torch.tensor(s.wavelet[:nsteps]).unsqueeze(0)
        spatial = torch.zeros(s.nx, s.ny)
        for (i, j, w, comp) in s.sources:
            spatial[i, j] += w
        t0 = time.perf_counter()
        if dfield:   # the per-cell operator of DI:100-176 on C_p = D C0_p, gradient w.r.t. D
            D = torch.tensor(np.asarray(wl.D_true, np.float32), requires_grad=True)
            C = [(D * float(c)).unsqueeze(0) for c in wl.C_true[0]]
        else:
            C = [torch.tensor(float(c), requires_grad=True) for c in wl.C_true[0]]
        ux, uy = pt.forward(C, s.h, s.dt, nsteps, 1, sig, spatial, torch.tensor(s.rho), hoist_weights=True,
                            per_cell=dfield)
        (ux.square().sum() + uy.square().sum()).backward()
        return time.perf_counter() - t0

    # the source fires near t0 = 1 us, so short windows still carry the same op count
    t_probe = run(10)
    n = int(max(10, min(s.nt, 10 * budget_s / max(t_probe, 1e-3))))
    t = run(n)
    rate = s.nx * s.ny * n / t / 1e9
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next(ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": round(rate, 5), "unit": "Gcell-steps/s (fwd+adjoint)", "cores": threads,
            "kind": "port", "cpu": cpu,
            "sample": f"oracle/plain_torch.py forward + autograd backward"
                      + (" (per-cell operator, dJ/dD)" if dfield else "") + f", {s.nx}x{s.ny} grid, "
                      f"{n} of {s.nt} steps, torch {threads} threads, {t:.1f} s"}


def cpu_oracle(wl, budget_s):
    """The FP64 NumPy oracle (forward with kept states + discrete adjoint) on
    one host core over a bounded window of the first shot's steps."""
    import numpy as np
    sys.path.insert(0, REPO)
    from oracle import ernn_fp64 as orc
    s = wl.survey.shots[0]
    C0 = np.asarray(wl.C_true[0], np.float64)
    C = C0 if wl.survey.mode == "const6" else [np.asarray(wl.D_true, np.float64) * c for c in C0]

    def run(n):
        t0 = time.perf_counter()
        tr, st = orc.forward(C, s.h, s.dt, s.rho, n, s.nx, s.ny, s.sources, s.wavelet[:n].astype(np.float64), 1,
                             s.receivers, keep_states=True)
        orc.adjoint_gradient(C, s.h, s.dt, s.rho, n, s.nx, s.ny, st, np.ones_like(tr), 1, s.receivers)
        return time.perf_counter() - t0

    t_probe = run(4)
    n = int(max(4, min(s.nt, 4 * budget_s / max(t_probe, 1e-3))))
    t = run(n)
    return {"value": round(s.nx * s.ny * n / t / 1e9, 6), "unit": "Gcell-steps/s (fwd+adjoint)", "cores": 1,
            "kind": "oracle",
            "sample": f"oracle/ernn_fp64.py FP64 forward + discrete adjoint, {s.nx}x{s.ny} grid, {n} of {s.nt} "
                      f"steps, NumPy on one core, {t:.1f} s"}


def time_instances(torch, ernn, wl, C_init, D_init, dev, iters):
    """Per-instance fwd+adjoint time (each instance alone on the GPU)."""
    import numpy as np
    times = []
    for b, s in enumerate(wl.survey.shots):
        one = ernn.Shot(**{**s.__dict__, "set": 0})
        sv1 = ernn.Survey([one], mode=wl.survey.mode, n_sets=1, ckpt_every=wl.survey.ckpt_every,
                          use_graph=wl.survey.use_graph, fixed_c0=wl.survey.fixed_c0,
                          reverse_u=wl.survey.reverse_u, orthotropic=wl.survey.orthotropic)
        m = ernn.ERNN(sv1, C_init[s.set:s.set + 1], D_init=D_init, device=dev, learn_constants=not sv1.fixed_c0)
        for k in range(iters + 1):
            if k == 1:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            m.zero_grad()
            m().square().mean().backward()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) / iters)
        m.engine.close()
    t = np.array(times)
    return {"p50_s": round(float(np.median(t)), 5), "max_s": round(float(t.max()), 5),
            "sum_s": round(float(t.sum()), 5), "n": len(t)}


# FEM observed-data generator runs (the reference scripts' main()):
# fem-pf  Parameter_inversion/FEM_generate_data.py:823-898, fem-df  Damage_inversion/FEM_generate_data.py:1486-1560
FEM_WORKLOADS = {
    "fem-pf": dict(t_total=60e-6, nt=2001, damage=None),
    "fem-df": dict(t_total=120e-6, nt=4001, damage=[
        {"type": "circle", "center": (0.1, 0.1), "size": 0.02, "reduction_factor": 0.5},
        {"type": "square", "center": (-0.1, -0.1), "size": 0.02, "reduction_factor": 0.5},
        {"type": "square", "center": (-0.1, -0.06), "size": 0.02, "reduction_factor": 0.5},
        {"type": "square", "center": (-0.06, -0.1), "size": 0.02, "reduction_factor": 0.5}]),
}
FEM_C = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)   # PF:840-845
FEM_RHO = 1610.0


def fem_torch_program(torch, fem, nx, L, C6, rho, dt, g32, w32, steps, dev, D=None):
    """Seconds per step of the reference's own FEM tensor program on the same
    GPU, restated op for op (PF:322-339 mass, PF:363-397 Jacobians, PF:417-457
    operator with three Gauss points and six index_add_ scatters, PF:481-497
    update on the (nodes, nt) history array), over `steps` steps."""
    import numpy as np
    x, y, el_np, bnd_np, dx, dy = fem.create_structured_triangular_mesh(nx, nx, L, L)
    N, E = nx * nx, el_np.shape[0]
    el = torch.from_numpy(el_np).to(dev)
    bnd = torch.from_numpy(bnd_np).to(dev)
    M = torch.zeros(N, dtype=torch.float32, device=dev)
    for i in range(3):
        M.index_add_(0, el[:, i], torch.full((E,), rho * (dx * dy / 2.0) / 3.0, dtype=torch.float32, device=dev))
    xc, yc = x[el_np % nx], y[el_np // nx]
    J = np.stack([np.stack([xc[:, 1] - xc[:, 0], xc[:, 2] - xc[:, 0]], 1),
                  np.stack([yc[:, 1] - yc[:, 0], yc[:, 2] - yc[:, 0]], 1)], 1)
    detJ = torch.from_numpy(np.abs(np.linalg.det(J))).float().to(dev)
    Jinv = torch.from_numpy(np.linalg.inv(J)).float().to(dev)
    dxi = torch.tensor([-1.0, 1.0, 0.0], device=dev)
    deta = torch.tensor([-1.0, 0.0, 1.0], device=dev)
    wg = torch.tensor([1 / 6] * 3, device=dev)
    if D is None:
        Ce = [torch.tensor(c, dtype=torch.float32, device=dev) for c in C6]
    else:   # per-element C at the centroid node (DF:1003-1023)
        ci = np.abs(x[None, :] - xc.mean(1)[:, None]).argmin(1)
        cj = np.abs(y[None, :] - yc.mean(1)[:, None]).argmin(1)
        Ce = [torch.from_numpy(c * D.astype(np.float64)[ci, cj]).float().to(dev) for c in C6]
    C11, C22, C12, C16, C26, C66 = Ce
    gn = torch.from_numpy(g32.T.flatten()).float().to(dev)
    tf = torch.from_numpy(w32).to(dev)
    ux = torch.zeros((N, steps + 2), dtype=torch.float32, device=dev)
    uy = torch.zeros((N, steps + 2), dtype=torch.float32, device=dev)

    def K(u_x, u_y):
        f_x = torch.zeros(N, dtype=torch.float32, device=dev)
        f_y = torch.zeros(N, dtype=torch.float32, device=dev)
        for gp in range(3):
            uex, uey = u_x[el], u_y[el]
            dNdx = Jinv[:, 0, 0].unsqueeze(1) * dxi.unsqueeze(0) + Jinv[:, 1, 0].unsqueeze(1) * deta.unsqueeze(0)
            dNdy = Jinv[:, 0, 1].unsqueeze(1) * dxi.unsqueeze(0) + Jinv[:, 1, 1].unsqueeze(1) * deta.unsqueeze(0)
            exx = torch.sum(dNdx * uex, dim=1)
            eyy = torch.sum(dNdy * uey, dim=1)
            gxy = torch.sum(dNdy * uex + dNdx * uey, dim=1)
            sxx = C11 * exx + C12 * eyy + C16 * gxy
            syy = C12 * exx + C22 * eyy + C26 * gxy
            txy = C16 * exx + C26 * eyy + C66 * gxy
            fac = detJ * wg[gp]
            fex = fac.unsqueeze(1) * (dNdx * sxx.unsqueeze(1) + dNdy * txy.unsqueeze(1))
            fey = fac.unsqueeze(1) * (dNdy * syy.unsqueeze(1) + dNdx * txy.unsqueeze(1))
            for i in range(3):
                f_x.index_add_(0, el[:, i], fex[:, i])
                f_y.index_add_(0, el[:, i], fey[:, i])
        return f_x, f_y

    def run(n_steps):
        for n in range(1, n_steps + 1):
            fy_ext = tf[n] * gn * (1.0 / rho) * M
            fix, fiy = K(ux[:, n], uy[:, n])
            ux[:, n + 1] = 2 * ux[:, n] - ux[:, n - 1] - (dt * dt) * fix / M
            uy[:, n + 1] = 2 * uy[:, n] - uy[:, n - 1] + (dt * dt) * (fy_ext - fiy) / M
            ux[bnd, n + 1] = 0
            uy[bnd, n + 1] = 0

    run(3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def fem_main(args):
    """One FEM generator run of the reference's main() on the HIP path: tile
    autotune, `warmup` + `steps` timed whole runs (graph replay), the step
    kernel's roofline, the reference's tensor program on the same GPU and the
    FP64 oracle on one CPU core over bounded windows."""
    import numpy as np
    import torch
    from ernn import fem
    dev = torch.device("cuda:0")
    spec = FEM_WORKLOADS[args.workload]
    nx, L = args.fem_n, 0.4
    h = L / (nx - 1)
    x = np.linspace(-L / 2, L / 2, nx)
    X, Y = np.meshgrid(x, x, indexing="ij")
    vp, vs, _ = fem.calculate_wave_velocities_anisotropic(*FEM_C, FEM_RHO)
    nt_req = args.nt or spec["nt"]
    dt, nt, t = fem.time_axis(nt_req, spec["t_total"] * (nt_req - 1) / (spec["nt"] - 1), h, h, max(vp, vs))
    g32 = fem.gaussian_point_load(X, Y, 0.0, 0.0, 5e6).astype(np.float32)
    w32 = np.asarray(fem.ricker_wavelet(t, 100e3, 10e-6, 1e10), np.float32)
    D = None
    if spec["damage"]:
        mask, _ = fem.create_damage_regions(X, Y, spec["damage"])
        D = np.where(mask, 0.5, 1.0).astype(np.float32)
    Dt = torch.tensor(D, device=dev) if D is not None else None
    stride = 5
    eng = fem.FEMEngine(nx, nx, h, h, nt, dt, FEM_RHO, fem.nodal_loads(g32), damage=D is not None,
                        stride=stride, use_graph=bool(args.graph), device=dev)
    C32 = np.asarray(FEM_C, np.float32)
    tune = None
    if args.tiles:   # fixed tile "V,R"
        eng.set_tiling(*(int(s) for s in args.tiles.split(",")))
    elif args.autotune:
        tune = eng.tune(C32, w32, Dt, steps=min(200, nt - 2))
    if args.blocking not in ("auto", "0"):   # fixed time-blocked shape "RW,KS"
        eng.set_blocking(*(int(s) for s in args.blocking.split(",")))
    if tune:
        log("fem shapes (us per step): " + ", ".join(f"{k}: {v / min(200, nt - 2) * 1e3:.2f}"
                                                    for k, v in sorted(tune.items())))
    out = (torch.empty((eng.n_frames, nx, nx), dtype=torch.float32, device=dev),
           torch.empty((eng.n_frames, nx, nx), dtype=torch.float32, device=dev))
    for _ in range(args.warmup):
        eng.run(C32, w32, Dt, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run(C32, w32, Dt, out=out)
    torch.cuda.synchronize()
    run_s = (time.perf_counter() - t0) / args.steps
    nodes = nx * nx
    node_steps = nodes * (nt - 2)
    value = node_steps / run_s / 1e9
    step_ms = eng.time_steps(nt - 2, C32, w32, Dt, out=out) / (nt - 2)
    RW, KS = eng.blocking
    if KS > 1:
        # time-blocked launch of KS steps: reads u^{n0-1}, u^{n0} (+ the damage factor) and writes its last
        # two states (2 components each), plus the frames of its steps
        kname = "fem_block_kernel"
        bpl = 32 + (4 if D is not None else 0) + 8.0 * KS / stride
        launches = -(-(nt - 2) // KS)
        launch_ms = step_ms * (nt - 2) / launches
    else:
        # one step: read u^n, u^{n-1} (2 components), write u^{n+1} (+ the damage factor), frame every `stride`
        kname = "fem_step_kernel"
        bpl = 24 + (4 if D is not None else 0) + 8.0 / stride
        launches, launch_ms = nt - 2, step_ms
    achieved = nodes * bpl / (launch_ms * 1e-3) / 1e9
    traffic = valu = None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        wname = args.workload if nx == 301 else f"{args.workload}-{nx}"
        entry = json.load(open(tpath)).get(f"{wname}/{kname}")
        if entry and entry.get("shape") in (None, [RW, KS] if KS > 1 else list(eng.tile)):
            traffic = entry["bytes"]
            if entry.get("valu_insts"):
                rate = entry["valu_insts"] / (launch_ms * 1e-3)
                valu = {"insts_per_launch": entry["valu_insts"], "issue_rate": round(rate / 1e9, 1),
                        "peak": VALU_PEAK_WIPS / 1e9, "unit": "G wave-instructions/s",
                        "frac": round(rate / VALU_PEAK_WIPS, 4)}
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "valu": valu,
                "bytes_per_node": bpl, "nodes_per_launch": nodes, "avg_launch_us": round(launch_ms * 1e3, 3),
                "launches_per_run": launches, "us_per_step": round(step_ms * 1e3, 3)}
    log(f"fem run {run_s * 1e3:.2f} ms ({value:.2f} Gnode-steps/s), step kernel {step_ms * 1e3:.2f} us")
    baseline = None
    if args.baseline:
        win = min(args.baseline_window, 200, nt - 2)
        s = fem_torch_program(torch, fem, nx, L, FEM_C, FEM_RHO, dt, g32, w32, win, dev, D)
        baseline = {"s_per_step": round(s, 6), "s_per_run": round(s * (nt - 2), 3), "window_steps": win,
                    "value": round(nodes / s / 1e9, 5), "unit": "Gnode-steps/s"}
        log(f"reference tensor program on the GPU: {s * 1e3:.3f} ms per step")
    cpu = None
    if args.cpu:
        from oracle import fem_fp64
        n_win = max(10, min(nt - 2, int(args.cpu_seconds * 20)))
        t0 = time.perf_counter()
        fem_fp64.solve(nx, nx, L, L, dt, n_win + 2, np.asarray(C32, np.float64), FEM_RHO, g32, w32[:n_win + 2],
                       C_fields=None if D is None else [D.astype(np.float64) * float(c) for c in C32], stride=stride)
        s = (time.perf_counter() - t0) / n_win
        cpu = {"value": round(nodes / s / 1e9, 5), "unit": "Gnode-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle (sparse element assembly + {n_win} steps) of the same run"}
        log(f"cpu FP64 oracle: {cpu['value']} Gnode-steps/s on one core")
    line = {
        "metric": "FEM generator Gnode-steps/s (1 GPU)", "value": round(value, 3), "unit": "Gnode-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(run_s * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload if nx == 301 else f"{args.workload}-{nx}", "grid": f"{nx}x{nx}",
                   "nt": nt, "frames_every": stride, "blocking": list(eng.blocking),
                   "damage": D is not None, "tile": list(eng.tile), "graph": bool(args.graph),
                   "parallelism": "none"},
        "s_per_run": round(run_s, 5),
        "reference_program_on_gpu": baseline,
        "speedup_vs_reference_program": round(baseline["s_per_run"] / run_s, 1) if baseline else None,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    eng.close()
    print(json.dumps(line), flush=True)


def tv_adam_main(args):
    """The damage inversion's regularised update (DI:179-195, DI:304-322,
    DI:453-460, DI:478-483) on six 2048^2 log-stiffness fields: the fused HIP
    step (FusedTVAdam) against autograd through clamp(exp) + TV and
    torch.optim.Adam on the same GPU (the data gradient is given), and the FP64
    oracle on one CPU core."""
    import numpy as np
    import torch
    import ernn
    from ernn import FusedTVAdam
    dev = torch.device("cuda:0")
    n = 2048
    bounds = [(50e9, 250e9), (5e9, 50e9), (5e9, 50e9), (0.0, 60e9), (0.0, 20e9), (5e9, 30e9)]
    C0 = [130.74e9, 11.5e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9]
    g = torch.Generator(device=dev).manual_seed(0)
    th0 = [(torch.full((1, n, n), math.log(c), device=dev) + 0.05 * torch.randn((1, n, n), generator=g, device=dev))
           .contiguous() for c in C0]
    gd = [1e-3 * torch.randn((1, n, n), generator=g, device=dev) for _ in C0]
    lam, lr = 1e-12, 1e-3
    cells = 6 * n * n

    fused_p = [t.clone() for t in th0]
    opt = FusedTVAdam(fused_p, lr=lr, lambda_tv=lam, bounds=bounds)

    def fused_step():
        for p, d in zip(fused_p, gd):
            p.grad = d.clone()
        opt.step()

    ref_p = [t.clone().requires_grad_() for t in th0]
    ropt = torch.optim.Adam(ref_p, lr=lr)

    def torch_step():
        ropt.zero_grad()
        for p, d in zip(ref_p, gd):
            p.grad = d.clone()
        C = [torch.clamp(torch.exp(p), *b) for p, b in zip(ref_p, bounds)]
        (lam * sum(ernn.total_variation_2d(c) for c in C)).backward()
        ropt.step()

    def timed(fn, k):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    k = max(args.steps, 5)
    s_fused = timed(fused_step, k)
    s_torch = timed(torch_step, k)
    # kernel time alone (hipEvents on the stream), without the gradient copy of fused_step
    for p, d in zip(fused_p, gd):
        p.grad = d.clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(k):
        opt.step()
    e1.record()
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / k
    # bytes per cell: TV pass reads theta (+ neighbours, cached) and grad, writes grad (12);
    # Adam pass reads theta, grad, m, v and writes theta, m, v (28)
    bpc = 40
    achieved = cells * bpc / (step_ms * 1e-3) / 1e9
    traffic = valu = None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        entry = json.load(open(tpath)).get("tv-adam-2048/tv_grad_kernel+adam_kernel")
        if entry:
            traffic = entry["bytes"]
    roofline = {"bound": "hbm", "kernel": "tv_grad_kernel+adam_kernel", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "valu": valu, "bytes_per_cell": bpc, "cells_per_launch": cells, "avg_launch_us": round(step_ms * 1e3, 2)}
    log(f"tv+adam: fused {s_fused * 1e3:.3f} ms per update ({step_ms * 1e3:.1f} us kernels), "
        f"autograd TV + torch Adam {s_torch * 1e3:.3f} ms")
    cpu = None
    if args.cpu:
        from oracle import tv_adam_fp64 as T
        th = th0[0][0].double().cpu().numpy()
        gg = gd[0][0].double().cpu().numpy()
        t0 = time.perf_counter()
        Cf, chain = T.fields(th, *bounds[0])
        gt = gg + lam * T.tv_subgradient(Cf) * chain
        T.tv(Cf)
        T.adam_step(th, gt, 0 * th, 0 * th, 1, lr, 0.9, 0.999, 1e-8)
        s = time.perf_counter() - t0
        cpu = {"value": round(n * n / s / 1e9, 5), "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
               "sample": "FP64 oracle: one 2048^2 field, TV subgradient + chain + one Adam step"}
    line = {
        "metric": "TV + Adam update Gcell-updates/s (1 GPU)", "value": round(cells / s_fused / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": k, "warmup": args.warmup,
        "ms_per_step": round(s_fused * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "tv-adam-2048", "fields": 6, "grid": f"{n}x{n}", "parallelism": "none"},
        "reference_path_on_gpu": {"ms_per_step": round(s_torch * 1e3, 4), "what": "autograd TV + torch.optim.Adam"},
        "speedup_vs_reference_path": round(s_torch / s_fused, 2),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Six-field damage inversion (Damage_inversion/Inversion_adam.py, DI) at the
# reference's own size: 301^2 grid over 0.4 m, 4001 steps over 120 us, one
# Gaussian y-force at the centre (f0 100 kHz, t0 10 us, amplitude 1e10,
# a = 5e6, DI:885-895), rho 1610, whole-field frames every P = 5 steps,
# six per-cell log-stiffness fields.  One step = one epoch of DI:465-560:
# forward, 5 MSE_x + MSE_y on every second frame, backward, TV (1e-21) +
# Adam (lr 0.02), here with the fused TV + Adam update.
# ---------------------------------------------------------------------------
DI_TRUE = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)   # DI:838-843


def di_fields(n, seed=5):
    """True fields: the healthy constants scaled by three elliptical damage
    regions (D = 0.5-0.8), as DF's damaged elements (DF:145, DF:183-212)."""
    import numpy as np
    from ernn import ellipse_damage
    D = ellipse_damage(n, n, 3, np.random.default_rng(seed), semi=(12, 30), level=(0.5, 0.8), taper=3)
    return [(D * c).astype(np.float32) for c in DI_TRUE], D


def di_autograd(torch, shot, fields0, obs, window, lam, stride):
    """s/epoch of the reference's own DI epoch on the GPU: clamp(exp(theta)) per
    cell (DI:304-322), the per-cell operator + frames (DI:100-176, DI:331-393),
    5 MSE_x + MSE_y on every second frame, lam TV (DI:441-463), torch Adam.
    Timed over `window` steps (memory-capped) and scaled to nt."""
    bounds = [(50e9, 250e9), (5e9, 50e9), (5e9, 50e9), (0.0, 60e9), (0.0, 20e9), (5e9, 30e9)]
    tape_step = 48 * 4 * shot.nx * shot.ny
    mem_cap = min(torch.cuda.get_device_properties(0).total_memory // 2, int(0.8 * torch.cuda.mem_get_info()[0]))
    W = max(10, min(window, shot.nt, mem_cap // tape_step))
    thetas = [torch.nn.Parameter(f.clamp_min(1e6).log()) for f in fields0]
    opt = torch.optim.Adam(thetas, lr=0.02)
    nf = (W + stride - 1) // stride

    def tv(c):
        return (c[1:, :] - c[:-1, :]).abs().sum() + (c[:, 1:] - c[:, :-1]).abs().sum()

    def it():
        opt.zero_grad()
        C = [torch.clamp(torch.exp(t), *b) for t, b in zip(thetas, bounds)]
        X, Y = autograd_traces(torch, shot, None, True, nsteps=W, fields=C, stride=stride)
        loss = 5 * torch.mean((X[:, ::2] - obs[0][:, :nf:2]) ** 2) + torch.mean((Y[:, ::2] - obs[1][:, :nf:2]) ** 2)
        loss = loss + lam * sum(tv(c) for c in C)
        loss.backward()
        opt.step()
    it()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) * shot.nt / W
    log(f"autograd DI epoch on the GPU: {t:.3f} s" + (f" (extrapolated from {W} of {shot.nt} steps)" if W < shot.nt else ""))
    return t, W


def di_cpu_baseline(torch, shot, fields, budget_s, stride):
    """The op-for-op CPU port (oracle/plain_torch.py, per-cell operator) with
    autograd on the host cores over a bounded window of steps."""
    sys.path.insert(0, REPO)
    from oracle import plain_torch as pt
    threads = torch.get_num_threads()

    def run(nsteps):
        C = [torch.tensor(f).unsqueeze(0).requires_grad_() for f in fields]
        sig = torch.tensor(shot.wavelet[:nsteps]).unsqueeze(0)
        spatial = torch.zeros(shot.nx, shot.ny)
        for (i, j, w, comp) in shot.sources:
            spatial[i, j] += w
        t0 = time.perf_counter()
        ux, uy = pt.forward(C, shot.h, shot.dt, nsteps, stride, sig, spatial, torch.tensor(shot.rho),
                            per_cell=True, hoist_weights=True)
        (ux.square().sum() + uy.square().sum()).backward()
        return time.perf_counter() - t0

    t_probe = run(10)
    n = int(max(10, min(shot.nt, 10 * budget_s / max(t_probe, 1e-3))))
    t = run(n)
    return {"value": round(shot.nx * shot.ny * n / t / 1e9, 5), "unit": "Gcell-steps/s (fwd+adjoint)",
            "cores": threads, "kind": "port",
            "sample": f"oracle/plain_torch.py per-cell forward + autograd backward, {shot.nx}x{shot.ny} grid, "
                      f"{n} of {shot.nt} steps, torch {threads} threads, {t:.1f} s"}


def di_main(args):
    import numpy as np
    import torch
    import ernn
    from ernn.autotune import timing_buffers, tune_blocking, tune_tiles
    from ernn.inversion import add_noise_rms

    dev = torch.device("cuda:0")
    n, P, lam = 301, 5, 1e-21
    nt = args.nt or 4001
    h, dt = 0.4 / (n - 1), 120e-6 / 4000
    sim = ernn.SpatiallyVaryingAnisotropicSimulator((n, n), h, dt, nt, P, total_time=(nt - 1) * dt,
                                                    rho_known=1610.0).to(dev)
    names = ("C11", "C22", "C12", "C16", "C26", "C66")
    thetas = [getattr(sim, "log_" + k) for k in names]
    init = [t.detach().clone() for t in thetas]
    true, D = di_fields(n)
    eng = sim.engine(False)
    shot = eng.survey.shots[0]
    log(f"workload di-301: {n}^2 x {nt}, six per-cell fields, frames every {P} steps, "
        f"{len(shot.sources)} source cells; {n * n * nt / 1e9:.3f} Gcell-steps per pass")
    with torch.no_grad():
        for t, f in zip(thetas, true):
            t.copy_(torch.tensor(f, device=dev).clamp_min(1e6).log().unsqueeze(0))
        ux, uy = sim()
        gen = torch.Generator().manual_seed(1)
        obs_x = add_noise_rms(ux, 0.01, gen)
        obs_y = add_noise_rms(uy, 0.01, gen)
        field_true = torch.cat([torch.tensor(f, device=dev).reshape(-1) for f in true])
        for t, t0 in zip(thetas, init):
            t.copy_(t0)
    wav = sim.source_signal[0].contiguous()
    steps = max(16, min(96, 40000 // max(1, n * n // 1000)))
    if args.autotune:
        tune = tune_tiles(eng, None, field_true, wav, steps=steps, which=(0, 2) if eng.info.fuse else (0, 1))
        log("autotune: " + ", ".join(f"{('fwd', 'adj', 'fused')[w]} V={v} R={r} B={b} {ms * 1e3:.1f} us/step"
                                     for w, (v, r, b, ms, _) in tune.items()))
        bv, brw, bks, bms, btimes = tune_blocking(eng, None, field_true, wav, steps=steps)
        log(f"autotune: forward blocking (V, RW, KS) = ({bv}, {brw}, {bks}) {bms * 1e3:.2f} us/step "
            f"(one step per launch {btimes[(0, 0, 0)] * 1e3:.2f}; {_ranked(btimes)})")
        bv, brw, bks, bms, btimes = tune_blocking(eng, None, field_true, wav, steps=steps, which=1)
        log(f"autotune: adjoint blocking (V, RW, KS) = ({bv}, {brw}, {bks}) {bms * 1e3:.2f} us/step "
            f"(one step per launch {btimes[(0, 0, 0)] * 1e3:.2f}; {_ranked(btimes)})")
    inv = ernn.SpatiallyVaryingParameterInverter(sim, obs_x, obs_y, secondary_time_stride=2, lambda_tv=lam,
                                                 verbose=False)
    opt = inv.fused_optimizer(lr=0.02)

    def epoch():
        opt.zero_grad()
        px, py = sim()
        data = inv.data_loss(px, py)
        data.backward()
        opt.step()
        return data

    for _ in range(args.warmup):
        epoch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = epoch()
    torch.cuda.synchronize()
    dt_iter = (time.perf_counter() - t0) / args.steps
    cells = n * n
    value = cells * nt / dt_iter / 1e9
    log(f"HIP: {dt_iter * 1e3:.2f} ms/epoch, {value:.2f} Gcell-steps/s fwd+adjoint, data loss {loss.item():.4e}")

    # per-kernel launch durations (hipEvents on the launch stream)
    ws = eng.acquire_workspace()
    tr0, gr0 = timing_buffers(eng)
    field = torch.cat([c.reshape(-1) for c in sim.get_material_params()[:6]]).detach().contiguous()
    nk = min(200, nt)
    eng.time_steps(0, 10, None, field, wav, ws, tr0, gr0)
    fwd_ms = eng.time_steps(0, nk, None, field, wav, ws, tr0, gr0) / nk
    adj_ms = eng.time_steps(1, nk, None, field, wav, ws, tr0, gr0) / nk
    fused_ms = eng.time_steps(2, nk, None, field, wav, ws, tr0, gr0) / nk if eng.info.fuse else None
    KS, AKS = eng.blocking[2], eng.adj_blocking[2]
    blk_ms = ablk_ms = None
    if KS > 1:
        L = max(2, nk // KS)
        eng.time_steps(3, 2, None, field, wav, ws, tr0, gr0)
        blk_ms = eng.time_steps(3, L, None, field, wav, ws, tr0, gr0) / L
    if AKS > 1:
        L = max(2, nk // AKS)
        eng.time_steps(4, 2, None, field, wav, ws, tr0, gr0)
        ablk_ms = eng.time_steps(4, L, None, field, wav, ws, tr0, gr0) / L
    eng.release_workspace(ws)
    del ws, tr0, gr0
    bf, ba = BYTES_PER_CELL[("field6", 0)], BYTES_PER_CELL[("field6", 1)]
    K = int(eng.info.ckpt_every)
    S = math.ceil(nt / K)
    if AKS > 1:
        kern = {"adj_block_kernel": (ablk_ms, sum(math.ceil(min(K, nt - s * K) / AKS) for s in range(S)), ba * AKS)}
    else:
        kern = {"adj_step_kernel": (adj_ms, nt, ba)}
    if KS > 1:
        kern["fwd_block_kernel"] = (blk_ms, math.ceil(nt / KS) + (S - 1) * math.ceil(K / KS), bf * KS)
    else:
        kern["fwd_step_kernel"] = (fwd_ms, nt + (S - 1) * K, bf)
    name = max(kern, key=lambda k: kern[k][0] * kern[k][1])
    ms, launches, bpc = kern[name]
    achieved = cells * bpc / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_cell": bpc,
                "cells_per_launch": cells, "avg_launch_us": round(ms * 1e3, 3), "launches_per_iter": launches,
                "fwd_step_us": round(fwd_ms * 1e3, 3), "adj_step_us": round(adj_ms * 1e3, 3),
                "fused_step_us": round(fused_ms * 1e3, 3) if fused_ms else None,
                "fwd_block_us_per_step": round(blk_ms / KS * 1e3, 3) if blk_ms else None,
                "adj_block_us_per_step": round(ablk_ms / AKS * 1e3, 3) if ablk_ms else None}
    log(f"kernels: fwd {fwd_ms * 1e3:.2f} us, adj {adj_ms * 1e3:.2f} us, fused "
        f"{fused_ms * 1e3 if fused_ms else float('nan'):.2f} us per launch, time-blocked fwd "
        f"{blk_ms / KS * 1e3 if blk_ms else float('nan'):.2f} us, adj "
        f"{ablk_ms / AKS * 1e3 if ablk_ms else float('nan'):.2f} us per step; {name}: {achieved:.0f} GB/s algorithmic")
    eng.trim()
    torch.cuda.empty_cache()

    baseline = None
    if args.baseline:
        fields0 = [torch.tensor(np.full((n, n), c, np.float32), device=dev) for c in DI_TRUE]
        t_ag, W = di_autograd(torch, shot, fields0, (obs_x, obs_y), args.baseline_window, lam, P)
        baseline = {"s_per_epoch": round(t_ag, 4), "value": round(cells * nt / t_ag / 1e9, 4),
                    "unit": "Gcell-steps/s", "window_steps": W, "extrapolated": W < nt}
    cpu = None
    if args.cpu:
        torch.set_num_threads(min(os.cpu_count() or 1, 16))
        cpu = di_cpu_baseline(torch, shot, true, args.cpu_seconds, P)
        log(f"cpu baseline: {cpu['value']} Gcell-steps/s on {cpu['cores']} threads")
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt_iter * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "di-301", "grid": f"{n}x{n}", "nt": nt, "fields": 6, "frame_stride": P,
                   "secondary_stride": 2, "lambda_tv": lam, "mode": "field6", "ckpt_every": K,
                   "fused_backward": bool(eng.info.fuse), "tiles": list(eng.tiles), "blocking": list(eng.blocking),
                   "adj_blocking": list(eng.adj_blocking), "parallelism": "none"},
        "s_per_epoch": round(dt_iter, 5),
        "autograd_baseline": baseline,
        "speedup_vs_autograd": round(baseline["s_per_epoch"] / dt_iter, 2) if baseline else None,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Observed-data file (HDF5) of the damage inversion at the reference's size:
# DF main (301^2 nodes over 0.4 m, 4001 samples over 120 us, four damage
# regions, frames every 5 steps) written by the FEM generator through
# ernn.hdf5 and read back by the loader; beside it the same arrays written and
# read through the HDF5 C library (the library h5py calls), where installed.
# ---------------------------------------------------------------------------
DF_REGIONS = [{"type": "circle", "center": (0.1, 0.1), "size": 0.02},
              {"type": "square", "center": (-0.1, -0.1), "size": 0.02},
              {"type": "square", "center": (-0.1, -0.06), "size": 0.02},
              {"type": "square", "center": (-0.06, -0.1), "size": 0.02}]   # DF:1510-1545


def _libhdf5():
    import ctypes as C
    for name in ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"):
        try:
            lib = C.CDLL(name)
        except OSError:
            continue
        hid = C.c_int64
        for fn, res, argt in [("H5Fopen", hid, [C.c_char_p, C.c_uint, hid]), ("H5Fclose", C.c_int, [hid]),
                              ("H5Fcreate", hid, [C.c_char_p, C.c_uint, hid, hid]),
                              ("H5Dopen2", hid, [hid, C.c_char_p, hid]), ("H5Dclose", C.c_int, [hid]),
                              ("H5Dread", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                              ("H5Dwrite", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
                              ("H5Dcreate2", hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
                              ("H5Pcreate", hid, [hid]), ("H5Pclose", C.c_int, [hid]),
                              ("H5Pset_chunk", C.c_int, [hid, C.c_int, C.POINTER(C.c_ulonglong)]),
                              ("H5Pset_deflate", C.c_int, [hid, C.c_uint]),
                              ("H5Screate_simple", hid, [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]),
                              ("H5Sclose", C.c_int, [hid])]:
            f = getattr(lib, fn)
            f.restype, f.argtypes = res, argt
        lib.H5open()
        return lib
    return None


def _lib_read(lib, path, shape):
    import ctypes as C
    import numpy as np
    flt = C.c_int64.in_dll(lib, "H5T_NATIVE_FLOAT_g").value
    f = lib.H5Fopen(path.encode(), 0, 0)
    out = []
    for k in ("ux", "uy"):
        d = lib.H5Dopen2(f, k.encode(), 0)
        a = np.empty(shape, np.float32)
        if lib.H5Dread(d, flt, 0, 0, 0, a.ctypes.data) < 0:
            raise RuntimeError(f"H5Dread {k}")
        lib.H5Dclose(d)
        out.append(a)
    lib.H5Fclose(f)
    return out


def _lib_write(lib, path, arrays):
    import ctypes as C
    flt = C.c_int64.in_dll(lib, "H5T_NATIVE_FLOAT_g").value
    dcpl_cls = C.c_int64.in_dll(lib, "H5P_CLS_DATASET_CREATE_ID_g").value
    f = lib.H5Fcreate(path.encode(), 2, 0, 0)
    for k, a in arrays.items():
        dims = (C.c_ulonglong * 3)(*a.shape)
        sp = lib.H5Screate_simple(3, dims, None)
        dcpl = lib.H5Pcreate(dcpl_cls)
        lib.H5Pset_chunk(dcpl, 3, (C.c_ulonglong * 3)(1, a.shape[1], a.shape[2]))
        lib.H5Pset_deflate(dcpl, 4)
        d = lib.H5Dcreate2(f, k.encode(), flt, sp, 0, dcpl, 0)
        if lib.H5Dwrite(d, flt, 0, 0, 0, a.ctypes.data) < 0:
            raise RuntimeError(f"H5Dwrite {k}")
        lib.H5Dclose(d); lib.H5Pclose(dcpl); lib.H5Sclose(sp)
    lib.H5Fclose(f)


def hdf5_main(args):
    import tempfile
    import numpy as np
    import torch
    from ernn import data, fem, hdf5
    dev = torch.device("cuda:0")
    threads = min(os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "anisotropic_ricker_damage_data_ds5_FEM_fixed_bc.mat")
        t0 = time.perf_counter()
        path = fem.generate_reference_data(path, damage_configs=DF_REGIONS, stride=5, nx=301, ny=301, Lx=0.4, Ly=0.4,
                                           nt=args.nt or 4001, t_total=120e-6, source_amplitude=1e10, source_f0=100e3,
                                           source_t0=10e-6, gaussian_param=5e6, rho=1610.0, device=dev,
                                           C11=130.74e9, C22=11.50e9, C12=10.49e9, C16=30.67e9, C26=3.75e9,
                                           C66=14.77e9)
        t_gen = time.perf_counter() - t0
        size = os.path.getsize(path)
        with hdf5.File(path) as f:
            shape = f["ux"].shape
        payload = 2 * 4 * int(np.prod(shape))
        log(f"hdf5: FEM run + file ({shape[0]} frames of {shape[1]}x{shape[2]}, {size / 1e6:.1f} MB on disk, "
            f"{payload / 1e6:.1f} MB of float32) in {t_gen:.2f} s")

        def ours_read():
            with hdf5.File(path) as f:
                return f["ux"][:], f["uy"][:]

        def timed(fn, k):
            fn()
            t = time.perf_counter()
            for _ in range(k):
                r = fn()
            return (time.perf_counter() - t) / k, r

        k = max(1, args.steps)
        t_read, (ux, uy) = timed(ours_read, k)
        t_load, _ = timed(lambda: data.load_fem_damage_data(path, device=dev), k)
        wpath = os.path.join(d, "w.mat")
        t_write, _ = timed(lambda: hdf5.write(wpath, {"ux": ux, "uy": uy}, chunked={"ux": (1,) + ux.shape[1:],
                                                                                    "uy": (1,) + uy.shape[1:]}), k)
        log(f"hdf5: read {t_read * 1e3:.1f} ms ({payload / t_read / 1e9:.2f} GB/s, {threads} threads), loader to "
            f"the GPU {t_load * 1e3:.1f} ms, write {t_write * 1e3:.1f} ms ({payload / t_write / 1e9:.2f} GB/s)")
        lib = _libhdf5()
        library = None
        if lib is not None:
            t_lread, (lx, ly) = timed(lambda: _lib_read(lib, path, shape), k)
            assert np.array_equal(lx, ux) and np.array_equal(ly, uy)
            lpath = os.path.join(d, "lib.mat")
            t_lwrite, _ = timed(lambda: _lib_write(lib, lpath, {"ux": ux, "uy": uy}), k)
            library = {"read_ms": round(t_lread * 1e3, 2), "read_GBps": round(payload / t_lread / 1e9, 3),
                       "write_ms": round(t_lwrite * 1e3, 2), "write_GBps": round(payload / t_lwrite / 1e9, 3),
                       "cores": 1, "what": "HDF5 C library (the calls h5py makes), same file / arrays"}
            log(f"hdf5: HDF5 C library read {t_lread * 1e3:.1f} ms ({payload / t_lread / 1e9:.2f} GB/s), write "
                f"{t_lwrite * 1e3:.1f} ms")
    line = {
        "metric": "observed-data HDF5 read, GB/s of float32 payload", "value": round(payload / t_read / 1e9, 3),
        "unit": "GB/s", "n_gpus": 1, "steps": k, "warmup": 1, "ms_per_step": round(t_read * 1e3, 2),
        "higher_is_better": True, "scaling": "none", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "hdf5-df", "frames": shape[0], "grid": f"{shape[1]}x{shape[2]}", "gzip": 4,
                   "chunk": [1, shape[1], shape[2]], "file_MB": round(size / 1e6, 1), "threads": threads},
        "write_ms": round(t_write * 1e3, 2), "loader_to_gpu_ms": round(t_load * 1e3, 2),
        "roofline": None, "cpu_baseline": library,
    }
    print(json.dumps(line), flush=True)


def time_hbm_budget(torch, ernn, sv, eng, C_init, D_init, obs, dev, args):
    """The same Adam iteration with the checkpoint interval that 80 % of the
    free HBM gives instead of the workload's cap (same tiles and blocking).
    The capped engine's workspaces are released first."""
    import dataclasses
    from ernn.autotune import state_bytes
    eng.trim()
    torch.cuda.empty_cache()
    sv2 = dataclasses.replace(sv, ckpt_budget=None, fuse=bool(eng.info.fuse))
    model = ernn.ERNN(sv2, C_init, D_init=D_init, device=dev, learn_constants=not sv2.fixed_c0)
    e2 = model.engine
    t = eng.tiles
    e2.set_tiling(0, t[0], t[1], t[4])
    e2.set_tiling(1, t[2], t[3], t[5])
    e2.set_blocking(*eng.blocking)
    e2.set_blocking(*eng.adj_blocking, which=1)
    if e2.info.fuse and eng.ff_rw:
        v, _, ks = eng.adj_blocking
        e2.set_blocking(v, eng.ff_rw, ks, which=2)
    opt = torch.optim.Adam(model.parameters(), lr=1e-2)

    def it():
        opt.zero_grad()
        loss = torch.mean((model() - obs) ** 2)
        loss.backward()
        opt.step()

    for _ in range(max(1, args.warmup)):
        it()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    K = int(e2.info.ckpt_every)
    out = {"budget": "80 % of free HBM", "ckpt_every": K, "ckpt_state_bytes": state_bytes(sv2, K),
           "fused_backward": bool(e2.info.fuse), "s_per_adam_iter": round(dt, 5),
           "value": round(sv.cell_steps() / dt / 1e9, 3), "unit": "Gcell-steps/s"}
    log(f"whole-HBM checkpoint budget: K = {K}, {dt * 1e3:.2f} ms/Adam-iter, {out['value']} Gcell-steps/s")
    e2.close()
    del model, opt
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Shots would shard across GPUs with one gradient all-reduce per Adam iteration.", file=sys.stderr)
        sys.exit(2)
    if args.workload in FEM_WORKLOADS:
        return fem_main(args)
    if args.workload == "tv-adam-2048":
        return tv_adam_main(args)
    if args.workload == "di-301":
        return di_main(args)
    if args.workload == "hdf5-df":
        return hdf5_main(args)
    import numpy as np
    import torch
    import ernn
    from ernn import configs
    from ernn.autotune import (state_bytes, timing_buffers, tune_backward, tune_blocking, tune_fused_recompute,
                               tune_tiles)

    dev = torch.device("cuda:0")
    kw = {"nt": args.nt} if args.nt else {}
    wl = configs.WORKLOADS[args.workload](**kw)
    sv = wl.survey
    sv.use_graph = bool(args.graph)
    if args.ckpt_every:
        sv.ckpt_every = args.ckpt_every
    variant = ""
    if args.joint_c0 and sv.fixed_c0:
        sv.fixed_c0, sv.orthotropic, variant = False, False, "-joint"   # the profiles of this variant carry the suffix
    if sv.mode == "dfield":   # damage field: reversed u by default, the banked schedule as a variant
        if args.reverse_u:
            sv.reverse_u = True
        else:
            variant += "-banked"
    rng = np.random.default_rng(0)
    C_true = wl.C_true
    pert = 1.0 + rng.uniform(-0.2, 0.2, size=C_true.shape)
    C_init = C_true * pert
    C_init[:, 3:5] = C_true[:, 3:5] + rng.uniform(-0.2, 0.2, size=(C_true.shape[0], 2)) * C_true[:, 5:6]
    if sv.fixed_c0:   # stiffness-degradation imaging: the background constants are known, D is inverted
        C_init = C_true.copy()
    D_init = None if sv.mode == "const6" else np.ones_like(wl.D_true)
    log(f"workload {wl.name}: {wl.description}; {sv.cell_steps() / 1e9:.3f} Gcell-steps per pass")

    params_true = torch.tensor(C_true.reshape(-1), dtype=torch.float32, device=dev)
    field_true = None if sv.mode == "const6" else torch.tensor(wl.D_true.reshape(-1), device=dev)
    cells_k = max(1, sum(s.nx * s.ny for s in sv.shots) // 1000)   # thousands of cells per step
    if args.fuse != "auto":
        sv.fuse = bool(int(args.fuse))
    elif args.autotune and not args.tiles and sv.fused and sv.mode != "field6":
        # small grids: the fused single-step backward vs the time-blocked one
        fuse, _, _, bt = tune_backward(sv, params_true, field_true, torch.tensor(sv.wavelets(), device=dev),
                                    steps=max(16, min(96, 40000 // cells_k)))
        sv.fuse = fuse
        log(f"autotune: backward {'fused' if fuse else 'unfused'} (fused {bt['fused'] * 1e3:.2f} vs unfused "
            f"time-blocked {bt['blocked'] * 1e3:.2f} us per reverse step)")

    model = ernn.ERNN(sv, C_init, D_init=D_init, device=dev, learn_constants=not sv.fixed_c0)
    eng = model.engine
    tune = None
    if args.tiles:
        vf, rf, va, ra, bf, ba = (int(x) for x in args.tiles.split(","))
        eng.set_tiling(0, vf, rf, bf)
        eng.set_tiling(1, va, ra, ba)
        log(f"tiles fixed: fwd V={vf} R={rf} B={bf}, adj V={va} R={ra} B={ba}")
    elif args.autotune:
        tune = tune_tiles(eng, params_true, field_true, model.wavelets, steps=max(20, min(100, 40000 // cells_k)),
                          which=(0, 2) if eng.info.fuse else ((0,) if eng.reverse_u else (0, 1)))
        log("autotune: " + ", ".join(f"{('fwd', 'adj', 'fused')[w]} V={v} R={r} B={b} {ms * 1e3:.1f} us/step"
                                     for w, (v, r, b, ms, _) in tune.items()))
    if args.blocking not in ("auto", "0"):
        eng.set_blocking(*(int(x) for x in args.blocking.split(",")))
        log(f"time-blocked forward fixed: (V, RW, KS) = {eng.blocking}")
    elif args.blocking == "auto" and args.autotune:
        bv, brw, bks, bms, btimes = tune_blocking(eng, params_true, field_true, model.wavelets,
                                                  steps=max(16, min(96, 40000 // cells_k)))
        log(f"autotune: forward blocking (V, RW, KS) = ({bv}, {brw}, {bks}) {bms * 1e3:.2f} us/step "
            f"(one step per launch {btimes[(0, 0, 0)] * 1e3:.2f}; {_ranked(btimes)})")
    if args.adj_blocking not in ("auto", "0"):
        eng.set_blocking(*(int(x) for x in args.adj_blocking.split(",")), which=1)
        log(f"time-blocked adjoint fixed: (V, RW, KS) = {eng.adj_blocking}")
    elif args.adj_blocking == "auto" and args.autotune and sv.mode != "field6":
        bv, brw, bks, bms, btimes = tune_blocking(eng, params_true, field_true, model.wavelets,
                                                  steps=max(16, min(96, 40000 // cells_k)), which=1)
        one = btimes.get((0, 0, 0))
        log(f"autotune: adjoint blocking (V, RW, KS) = ({bv}, {brw}, {bks}) {bms * 1e3:.2f} us/step ("
            + (f"one step per launch {one * 1e3:.2f}; " if one else "") + f"{_ranked(btimes)})")
    if eng.info.fuse and eng.adj_blocking[2] > 1:
        if args.ff_rw != "auto":
            v, _, ks = eng.adj_blocking
            eng.set_blocking(v, int(args.ff_rw), ks, which=2)
            log(f"fused recompute rows fixed: {eng.ff_rw}")
        elif args.autotune:
            frw, ftimes = tune_fused_recompute(eng, params_true, field_true, model.wavelets,
                                               steps=max(16, min(96, 40000 // cells_k)))
            if ftimes:
                log(f"autotune: fused recompute rows per wavefront = {frw} ("
                    + ", ".join(f"RW {c}: {t * 1e3:.2f}" for c, t in ftimes.items()) + " us per reverse step)")
    with torch.no_grad():
        obs = ernn.ernn_traces(eng, params=params_true, field=field_true, wavelets=model.wavelets)
        gen = torch.Generator(device="cpu").manual_seed(1)
        noise = torch.randn(obs.shape, generator=gen).to(dev)
        for b in range(len(sv.shots)):
            blk = eng.shot_traces(obs, b)
            nblk = eng.shot_traces(noise, b)
            for c in range(2):
                blk[c] += 0.01 * blk[c].pow(2).mean().sqrt() * nblk[c]
    opt = torch.optim.Adam(model.parameters(), lr=1e-2)

    def adam_iter():
        opt.zero_grad()
        tr = model()
        loss = torch.mean((tr - obs) ** 2)
        loss.backward()
        opt.step()
        return loss

    for _ in range(args.warmup):
        adam_iter()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        loss = adam_iter()
    torch.cuda.synchronize()
    dt_iter = (time.perf_counter() - t0) / args.steps
    value = sv.cell_steps() / dt_iter / 1e9
    log(f"HIP: {dt_iter * 1e3:.2f} ms/Adam-iter, {value:.2f} Gcell-steps/s fwd+adjoint, loss {loss.item():.4e}")

    # per-kernel launch durations (hipEvents around back-to-back launches on the launch stream)
    ws = eng.acquire_workspace()
    tr0, gr0 = timing_buffers(eng)
    nk = min(200, sv.nt_max)
    p = model.stiffness().detach().reshape(-1).contiguous()
    f = None if sv.mode == "const6" else model.damage.detach()
    eng.time_steps(0, 10, p, f, model.wavelets, ws, tr0, gr0)
    fwd_ms = eng.time_steps(0, nk, p, f, model.wavelets, ws, tr0, gr0) / nk
    adj_ms = None if eng.reverse_u else eng.time_steps(1, nk, p, f, model.wavelets, ws, tr0, gr0) / nk
    fused_ms = eng.time_steps(2, nk, p, f, model.wavelets, ws, tr0, gr0) / nk if eng.info.fuse else None
    KS = eng.blocking[2]
    blk_ms = blk_bank_ms = None
    nt, K = sv.nt_max, int(eng.info.ckpt_every)
    S = math.ceil(nt / K)
    rev = eng.reverse_u
    # the forward pass banks from level B on (the last checkpoint segment); a
    # reverse_u plan banks nothing
    B = nt if rev else (S - 1) * K
    if KS > 1:   # per time-blocked launch of KS steps
        # the first launches of the forward pass (no banking while L KS <= B),
        # then the whole pass, whose launches past B also bank: their average
        # is the banking launch's duration (as in the segment recompute)
        L = max(2, min(nk // KS, B // KS)) if B >= 2 * KS else max(2, nk // KS)
        eng.time_steps(3, 2, p, f, model.wavelets, ws, tr0, gr0)
        blk_ms = eng.time_steps(3, L, p, f, model.wavelets, ws, tr0, gr0) / L
        n_fwd, n_nb = math.ceil(nt / KS), B // KS
        if rev:   # no launch banks
            blk_bank_ms = blk_ms
        elif B >= 2 * KS and n_fwd > n_nb:
            t_all = eng.time_steps(3, n_fwd, p, f, model.wavelets, ws, tr0, gr0)
            blk_bank_ms = max(t_all - n_nb * blk_ms, 0.0) / (n_fwd - n_nb)
        else:   # one checkpoint segment: every launch of the pass banks
            blk_bank_ms = blk_ms
    AKS = eng.adj_blocking[2]
    ablk_ms = None
    if AKS > 1:
        L = max(2, nk // AKS)
        eng.time_steps(4, 2, p, f, model.wavelets, ws, tr0, gr0)
        ablk_ms = eng.time_steps(4, L, p, f, model.wavelets, ws, tr0, gr0) / L
    eng.release_workspace(ws)
    cells = sum(s.nx * s.ny for s in sv.shots)
    bf, ba = BYTES_PER_CELL[(sv.mode, 0)], BYTES_PER_CELL[(sv.mode, 1)]
    # (average ms per launch, launches per iteration, algorithmic bytes per
    # cell per launch, recompute bytes per cell per launch not counted): a
    # launch of KS steps is credited KS steps of algorithmic bytes; recompute
    # work (a segment recompute launch, the recompute half of a fused launch)
    # is overhead and is reported apart.
    kern = {}
    # recomputed steps: every segment but the last (the forward pass keeps the
    # last segment's states; a fixed_c0 plan runs one step past each
    # segment); the fused schedule recomputes the last one alone
    seg = K + (1 if sv.fixed_c0 else 0)
    n_rec = 0 if rev else (K if eng.info.fuse else (S - 1) * seg)
    blk_split = None
    if KS > 1:
        n_fwd = math.ceil(nt / KS)
        n_nb = n_fwd if rev else min(B // KS, n_fwd)
        n_rc = 0 if rev else (math.ceil(K / KS) if eng.info.fuse else (S - 1) * math.ceil(seg / KS))
        n_bank = n_fwd - n_nb + n_rc
        total = n_nb * blk_ms + n_bank * blk_bank_ms
        n_blk = n_fwd + n_rc
        # launches of the forward pass carry its algorithmic bytes, recompute launches none
        kern["fwd_block_kernel"] = (total / n_blk, n_blk, bf * KS * n_fwd / n_blk, bf * KS * n_rc / n_blk)
        blk_split = {"launches_no_bank": n_nb, "launches_bank": n_bank, "us_no_bank": round(blk_ms * 1e3, 3),
                     "us_bank": round(blk_bank_ms * 1e3, 3), "forward_pass_launches": n_fwd,
                     "recompute_launches": n_rc}
    else:
        kern["fwd_step_kernel"] = (fwd_ms, nt + n_rec, bf, 0)
    if eng.info.fuse and AKS > 1:   # fused time-blocked backward: AKS adjoint steps + AKS recompute steps
        kern["fused_block_kernel"] = (ablk_ms, math.ceil(nt / AKS), ba * AKS, bf * AKS)
    elif eng.info.fuse:   # backward: ~nt fused adjoint + recompute launches (+ the last segment alone)
        kern["fused_step_kernel"] = (fused_ms, nt, ba, bf)
    elif AKS > 1:
        kern["rev_block_kernel" if rev else "adj_block_kernel"] = (
            ablk_ms, sum(math.ceil(min(K, nt - s * K) / AKS) for s in range(math.ceil(nt / K))),
            (REV_BYTES_PER_CELL if rev else ba) * AKS, 0)
    else:
        kern["adj_step_kernel"] = (adj_ms, nt, ba, 0)
    name = max(kern, key=lambda k: kern[k][0] * kern[k][1])
    ms, launches, bpc, rbpc = kern[name]
    achieved = cells * bpc / (ms * 1e-3) / 1e9
    # HBM bytes per launch from the committed counter passes (FETCH_SIZE and
    # WRITE_SIZE, calibrated on a copy kernel), taken with the tile shapes
    # recorded beside them (traffic_tiles; within a few percent across tile
    # shapes, since every shape reads each field once per launch).
    # The entry is used only when it was taken with the shape the kernel has
    # now (time-blocked forward: blocking; time-blocked adjoint: adj_blocking;
    # fused time-blocked: adj_blocking + recompute rows; one-step: the tiles).
    kshape = {"fwd_block_kernel": list(eng.blocking), "adj_block_kernel": list(eng.adj_blocking),
              "rev_block_kernel": list(eng.adj_blocking),
              "fused_block_kernel": list(eng.adj_blocking) + [eng.ff_rw]}.get(name, list(eng.tiles))
    traffic, traffic_tiles, valu = None, None, None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        # entries per kernel shape ("<workload>/<kernel>@<shape>"), or one entry with its shape inside
        entry = (tj.get(f"{wl.name}{variant}/{name}@" + ",".join(map(str, kshape)))
                 or tj.get(f"{wl.name}{variant}/{name}"))
        if entry and entry.get("shape") == kshape:
            traffic, traffic_tiles = entry["bytes"], entry["tiles"]
            vi = entry.get("valu_insts")
            if name == "fwd_block_kernel" and blk_split and "bytes_no_bank" in entry:
                # the iteration's mix of launches that bank u (or L) and launches that do not
                nb, bk = blk_split["launches_no_bank"], blk_split["launches_bank"]
                traffic = int((nb * entry["bytes_no_bank"] + bk * entry["bytes"]) / (nb + bk))
                vi = int((nb * entry["valu_insts_no_bank"] + bk * entry["valu_insts"]) / (nb + bk))
            if vi:   # SQ_INSTS_VALU per launch, same counter runs
                entry = dict(entry, valu_insts=vi)
                rate = entry["valu_insts"] / (ms * 1e-3)
                valu = {"insts_per_launch": entry["valu_insts"], "issue_rate": round(rate / 1e9, 1),
                        "peak": VALU_PEAK_WIPS / 1e9, "unit": "G wave-instructions/s",
                        "frac": round(rate / VALU_PEAK_WIPS, 4)}
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_tiles": traffic_tiles, "kernel_shape": kshape, "valu": valu,
                "bytes_per_cell": bpc, "recompute_bytes_per_cell": rbpc, "cells_per_launch": cells,
                "avg_launch_us": round(ms * 1e3, 3), "launches_per_iter": launches,
                "fwd_step_us": round(fwd_ms * 1e3, 3), "adj_step_us": round(adj_ms * 1e3, 3) if adj_ms else None,
                "fused_step_us": round(fused_ms * 1e3, 3) if fused_ms else None,
                "fwd_block_us_per_step": round(blk_ms / KS * 1e3, 3) if blk_ms else None,
                "adj_block_us_per_step": round(ablk_ms / AKS * 1e3, 3) if ablk_ms else None,
                "fwd_block_launches": blk_split,
                "kernel_ms_per_iter": {k: round(v[0] * v[1], 2) for k, v in kern.items()},
                # the whole iteration against the algorithmic bytes of one forward and one adjoint
                # step per cell with u^m at hand (recompute and u's reversal excluded): the HBM
                # bound of the fwd + adjoint rate, the same for every schedule
                "iteration": {"bytes_per_cell_step": bf + ba,
                              "achieved": round((bf + ba) * sv.cell_steps() / dt_iter / 1e9, 1),
                              "frac": round((bf + ba) * sv.cell_steps() / dt_iter / 1e9 / HBM_PEAK_GBS, 4),
                              "bound_gcell_steps": round(HBM_PEAK_GBS / (bf + ba), 1)}}
    log(f"kernels: fwd {fwd_ms * 1e3:.2f} us, adj {adj_ms * 1e3 if adj_ms else float('nan'):.2f} us, "
        f"fused {fused_ms * 1e3 if fused_ms else float('nan'):.2f} us per launch, time-blocked fwd "
        f"{blk_ms / KS * 1e3 if blk_ms else float('nan'):.2f} us per step, time-blocked adj "
        f"{ablk_ms / AKS * 1e3 if ablk_ms else float('nan'):.2f} us per step; "
        f"{name}: {achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.2%} of HBM peak)")

    # the pooled workspaces (all of a segment's states) go back before the
    # baselines allocate their own
    del ws, tr0, gr0
    eng.trim()
    torch.cuda.empty_cache()
    baseline = None
    if args.baseline and sv.n_sets == 1:
        theta0 = torch.tensor(C_init / model.scale.cpu().numpy(), dtype=torch.float32, device=dev)
        obs_shots = [eng.shot_traces(obs, b) for b in range(len(sv.shots))]
        D0 = None if sv.mode == "const6" else torch.tensor(D_init, device=dev)
        ag = time_autograd(torch, wl, theta0, model.scale, D0, obs_shots, args.baseline_window,
                           learn_c=not sv.fixed_c0)
        best = min(ag["faithful"], ag["hoisted"])
        baseline = {"s_per_adam_iter_faithful": round(ag["faithful"], 4),
                    "s_per_adam_iter_hoisted": round(ag["hoisted"], 4),
                    "value": round(sv.cell_steps() / best / 1e9, 4), "unit": "Gcell-steps/s",
                    "window_steps": ag["window_steps"], "extrapolated": ag["extrapolated"]}

    per_instance = None
    if args.per_instance and len(sv.shots) > 1:
        per_instance = time_instances(torch, ernn, wl, C_init, D_init, dev, args.steps)

    cpu = None
    if args.cpu:
        torch.set_num_threads(min(os.cpu_count() or 1, 16))
        cpu = cpu_baseline(torch, wl, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']} Gcell-steps/s on {cpu['cores']} threads")
        cpu["fp64_oracle"] = cpu_oracle(wl, args.cpu_seconds / 3)
        log(f"cpu FP64 oracle: {cpu['fp64_oracle']['value']} Gcell-steps/s on one core")

    hbm = None
    if args.hbm_run and sv.ckpt_budget and not sv.reverse_u:   # reverse_u: the interval is not memory-bound
        hbm = time_hbm_budget(torch, ernn, sv, eng, C_init, D_init, obs, dev, args)

    s0 = sv.shots[0]
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt_iter * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name + variant, "grid": f"{s0.nx}x{s0.ny}", "nt": sv.nt_max,
                   "shots": len(sv.shots), "receivers": len(s0.receivers) if s0.receivers is not None else 0,
                   "mode": sv.mode, "fixed_c0": bool(sv.fixed_c0), "reverse_u": bool(sv.reverse_u),
                   "orthotropic": bool(sv.orthotropic),
                   "ckpt_every": int(eng.info.ckpt_every),
                   "ckpt_cap_bytes": sv.ckpt_budget, "ckpt_state_bytes": state_bytes(sv, int(eng.info.ckpt_every)),
                   "graph": bool(sv.use_graph),
                   "fused_backward": bool(eng.info.fuse),
                   "blocking": list(eng.blocking),
                   "adj_blocking": list(eng.adj_blocking), "fused_recompute_rw": eng.ff_rw,
                   "tiles": list(eng.tiles),
                   "parallelism": "none"},
        "s_per_adam_iter": round(dt_iter, 5),
        "autograd_baseline": baseline,
        "speedup_vs_autograd": (round(value / baseline["value"], 2) if baseline else None),
        "per_instance_alone": per_instance,
        "hbm_budget_run": hbm,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
