"""The kernels bench.py times, at the benchmark sizes.

The time-blocked and fused kernels index hundreds to tens of thousands of
blocks at the benchmark sizes (block tables, XCD banding of the block index,
the adjoint / recompute split of the fused grid), which small grids never
reach.  Each case runs the one-step kernels once as the reference schedule
and then every time-blocked shape the autotuner can choose, on the full grid:

  * traces bitwise equal to the one-step path (every cell's arithmetic is
    the same in every kernel shape);
  * gradients within 1e-5 (max relative) of the one-step path (only the
    order of the gradient sums differs);
  * the one-step path, and the shapes bench.py times, against the reference's
    own op sequence with autograd (oracle/plain_torch.py) on the same GPU:
    traces within 1e-5 relative L2, gradients within 1e-4.

Reference loop replaced: Parameter_inversion/Inversion_adam.py:268-301 with
the autograd backward at :447; Damage_inversion/Inversion_adam.py:100-176,
:355-387 with the backward at :511.
"""
import dataclasses

import numpy as np
import pytest
import torch

import ernn
from ernn import autotune, configs
from oracle import plain_torch as pt
from gpu_helpers import DEV, rel_l2, run_hip

pytestmark = pytest.mark.gpu

# The shapes bench.py times on each workload (its autotuner's picks on MI355X).
TIMED = {
    "single-512": dict(tiles=(1, 1, 1, 4, 8, 2), blocking=(1, 4, 4), fuse=True, adj_blocking=(1, 4, 4),
                       adj_ff_rw=8),
    "damage-1024": dict(tiles=(1, 8, 1, 4, 0, 0), blocking=(1, 8, 5), fuse=False, adj_blocking=(1, 8, 4)),
    "batched-uneven": dict(tiles=(1, 2, 1, 4, 0, 0), blocking=(1, 8, 5), fuse=False, adj_blocking=(1, 8, 5)),
    "aniso-2048": dict(tiles=(1, 2, 1, 4, 0, 0), blocking=(1, 8, 5), fuse=False, adj_blocking=(1, 6, 4)),
}
ONE_STEP = dict(tiles=(0, 0, 0, 0, 0, 0), blocking=(0, 0, 0), fuse=False, adj_blocking=(0, 0, 0), adj_ff_rw=0)


def _assert_gradient_bar(e_hip_plain, e_hip_fp64, e_plain_fp64):
    """The 1e-4 gradient bar against the reference's FP32 autograd.  The HIP
    recurrences carry compensated (Kahan) state sums, so over thousands of steps
    the HIP gradient is closer to the FP64 gradient than the FP32 autograd
    gradient is.  The HIP gradient must then be within 1e-4 of FP64, and its
    distance to FP32 autograd within 1e-4 or within 1.5 times autograd's own
    FP64 error (the disagreement is autograd's FP32 rounding)."""
    assert e_hip_fp64 < 1e-4
    assert e_hip_plain < max(1e-4, 1.5 * e_plain_fp64)


def _max_rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _variant(sv, **kw):
    """A copy of survey sv with the given schedule fields (use_graph on)."""
    base = dict(ONE_STEP)
    base.update(kw)
    return dataclasses.replace(sv, use_graph=True, **base)


def _fwd_shapes(mode):
    return [s for s in autotune.BLOCK_SHAPES if mode == "const6" or s[:2] != (2, 8)]


def _adj_shapes(mode):
    return [s for s in autotune.ADJ_BLOCK_SHAPES if mode == "const6" or (s[0] == 1 and s[2] != 5)]


def _run(sv, params, field, G):
    eng, tr, grads = run_hip(sv, params=params, field=field, G=G)
    info = (eng.blocking, eng.adj_blocking, eng.ff_rw, int(eng.info.fuse), int(eng.info.ckpt_every))
    eng.close()
    torch.cuda.empty_cache()
    return tr, grads, info


def _sweep(sv, K, params, field, G, ref_tr, ref_g, label):
    """Every forward shape (one-step adjoint), every unfused adjoint shape (one-step
    forward) and, for fuse-capable grids, every fused shape with both recompute heights."""
    mode = sv.mode
    cases = [dict(blocking=b) for b in _fwd_shapes(mode)]
    cases += [dict(blocking=(1, 4, 4), adj_blocking=a) for a in _adj_shapes(mode)]
    if sum(s.nx * s.ny for s in sv.shots) <= ernn.problem.FUSE_MAX_CELLS:
        for a in autotune.FUSED_BLOCK_SHAPES:
            if mode != "const6" and a[1] == 8:
                continue
            for rwf in ((0, 8) if a[1] == 4 else (0,)):
                cases.append(dict(blocking=(1, 4, 4), fuse=True, adj_blocking=a, adj_ff_rw=rwf))
    worst = 0.0
    for c in cases:
        print(f"{label}: shape case {c}", flush=True)
        tr, g, info = _run(_variant(sv, ckpt_every=K, **c), params, field, G)
        assert info[4] == K
        assert np.array_equal(tr, ref_tr), c
        for x, y in zip(g, ref_g):
            if x is not None:
                e = _max_rel(x, y)
                worst = max(worst, e)
                assert e < 1e-5, (c, e)
    print(f"{label}: {len(cases)} time-blocked shapes: traces bitwise, gradients within {worst:.2e} of one step")


# ----------------------------------------------------------------- single-512
def _plain_const(wl, C, G):
    """Reference op sequence + autograd for a one-shot receiver survey."""
    s = wl.survey.shots[0]
    sig = torch.tensor(s.wavelet, device=DEV).unsqueeze(0)
    spatial = torch.zeros(s.nx, s.ny, device=DEV)
    for (i, j, w, comp) in s.sources:
        spatial[i, j] += w
    Ct = [torch.tensor(float(c), device=DEV, requires_grad=True) for c in C]
    ux, uy = pt.forward(Ct, s.h, s.dt, s.nt, 1, sig, spatial, torch.tensor(s.rho, device=DEV), hoist_weights=True)
    r = torch.as_tensor(s.receivers, device=DEV).long()
    tr = torch.stack([ux[0][:, r[:, 0], r[:, 1]], uy[0][:, r[:, 0], r[:, 1]]])
    (tr * torch.tensor(G.reshape(tr.shape), dtype=torch.float32, device=DEV)).sum().backward()
    out = tr.detach().cpu().numpy(), torch.stack([c.grad for c in Ct]).cpu().numpy()
    del ux, uy, tr
    torch.cuda.empty_cache()
    return out


def test_single_512_every_shape_full_size():
    """512^2 x 2000, one shot, 256 receivers, K = 50 (the benchmark's single-shot
    configuration), with C16 and C26 nonzero so every gradient term is live."""
    wl = configs.single_512()
    sv = wl.survey
    C = wl.C_true[0] * np.array([1.05, 0.97, 1.1, 1.0, 1.0, 0.93]) + np.array([0, 0, 0, 0.4e9, -0.3e9, 0])
    rng = np.random.default_rng(31)
    G = rng.standard_normal(2 * sv.shots[0].nt * len(sv.shots[0].receivers))
    K = 50
    ref_tr, ref_g, _ = _run(_variant(sv, ckpt_every=K), C, None, G)
    tr_t, g_t, info = _run(_variant(sv, ckpt_every=K, **TIMED["single-512"]), C, None, G)
    assert info[:4] == ((1, 4, 4), (1, 4, 4), 8, 1)
    assert np.array_equal(tr_t, ref_tr)
    e_ts = _max_rel(g_t[0], ref_g[0])
    p_tr, p_g = _plain_const(wl, C, G)
    e_tr = rel_l2(ref_tr.reshape(p_tr.shape), p_tr)
    e_g1, e_gt = _max_rel(ref_g[0], p_g), _max_rel(g_t[0], p_g)
    print(f"single-512 full size: timed shapes grad vs one step {e_ts:.2e}; vs plain autograd: traces {e_tr:.2e}, "
          f"grad one step {e_g1:.2e}, timed {e_gt:.2e}")
    assert e_ts < 1e-5
    assert e_tr < 1e-5
    assert e_g1 < 1e-4 and e_gt < 1e-4
    _sweep(sv, K, C, None, G, ref_tr, ref_g, "single-512")


# ---------------------------------------------------------------- damage-1024
def _plain_dfield(sv, C0, D, G, trace_off):
    """Per-cell reference op sequence (DI:100-176 with PI force semantics) on
    C_p = D C0_p, autograd for dJ/dC0 and dJ/dD, shots one at a time."""
    Dr = torch.tensor(D, device=DEV).reshape(sv.shots[0].nx, sv.shots[0].ny).requires_grad_()
    Cr = torch.tensor(np.asarray(C0, np.float32), device=DEV).requires_grad_()
    trs = []
    for b, s in enumerate(sv.shots):
        sig = torch.tensor(s.wavelet, device=DEV).unsqueeze(0)
        spatial = torch.zeros(s.nx, s.ny, device=DEV)
        for (i, j, w, comp) in s.sources:
            spatial[i, j] += w
        fields = [(Dr * Cr[k]).unsqueeze(0) for k in range(6)]
        ux, uy = pt.forward(fields, s.h, s.dt, s.nt, 1, sig, spatial, torch.tensor(s.rho, device=DEV),
                            per_cell=True, hoist_weights=True)
        r = torch.as_tensor(s.receivers, device=DEV).long()
        t = torch.stack([ux[0][:, r[:, 0], r[:, 1]], uy[0][:, r[:, 0], r[:, 1]]])
        Gb = torch.tensor(G[trace_off[b]:trace_off[b + 1]], dtype=torch.float32, device=DEV).view(t.shape)
        (t * Gb).sum().backward()
        trs.append(t.detach().cpu().numpy())
        del ux, uy, t, fields
        torch.cuda.empty_cache()
    return trs, Cr.grad.cpu().numpy(), Dr.grad.reshape(-1).cpu().numpy()


@pytest.mark.parametrize("name", ["damage-1024", "aniso-2048"])
def test_damage_field_every_shape_full_grid(name):
    """damage-1024 at its full 1024^2 grid (and the aniso-2048 physics at
    1024^2: 30-degree fibre frame, C15, C35 != 0, five ellipses), two shots x
    300 steps, K = 24: every time-blocked shape vs the one-step path, the
    one-step path and the timed shapes vs per-cell plain autograd."""
    if name == "damage-1024":
        wl = configs.damage_1024(nt=300, shots=2, fixed_c0=False)
    else:
        wl = configs.aniso_2048(nt=300, n=1024, shots=2)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32).reshape(-1)
    K = 24
    probe = ernn.Engine(_variant(sv, ckpt_every=K), device=DEV)
    off = list(probe.trace_off)
    probe.close()
    G = np.random.default_rng(9).standard_normal(off[-1])
    ref_tr, ref_g, _ = _run(_variant(sv, ckpt_every=K), C0, D, G)
    tr_t, g_t, info = _run(_variant(sv, ckpt_every=K, **TIMED[name]), C0, D, G)
    assert np.array_equal(tr_t, ref_tr)
    e_ts = max(_max_rel(g_t[0], ref_g[0]), rel_l2(g_t[1], ref_g[1]))
    trs, gC, gD = _plain_dfield(sv, C0, D, G, off)
    e_tr = max(rel_l2(ref_tr[off[b]:off[b + 1]].reshape(t.shape), t) for b, t in enumerate(trs))
    e_c1, e_d1 = _max_rel(ref_g[0], gC), rel_l2(ref_g[1], gD)
    e_ct, e_dt = _max_rel(g_t[0], gC), rel_l2(g_t[1], gD)
    print(f"{name} 1024^2 x 300, 2 shots: timed shapes {info[:2]} grads vs one step {e_ts:.2e}; vs plain autograd: "
          f"traces {e_tr:.2e}, dJ/dC0 {e_c1:.2e} / {e_ct:.2e}, dJ/dD rel-L2 {e_d1:.2e} / {e_dt:.2e} (one step / timed)")
    assert e_ts < 1e-5
    assert e_tr < 1e-5
    assert max(e_c1, e_ct, e_d1, e_dt) < 1e-4
    # K = nt, the interval a whole-HBM budget gives the full workload: the
    # forward pass keeps every state and the backward recomputes nothing
    tr_n, g_n, info_n = _run(_variant(sv, ckpt_every=sv.nt_max, **TIMED[name]), C0, D, G)
    assert info_n[4] == sv.nt_max
    assert np.array_equal(tr_n, ref_tr)
    e_n = max(_max_rel(g_n[0], ref_g[0]), rel_l2(g_n[1], ref_g[1]))
    print(f"{name}: K = nt (nothing recomputed) grads vs one step K = {K}: {e_n:.2e}")
    assert e_n < 1e-5
    _sweep(sv, K, C0, D, G, ref_tr, ref_g, name)


# ------------------------------------------------------------------- di-301
DI_TRUE = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)   # DI:838-843
DI_NAMES = ("C11", "C22", "C12", "C16", "C26", "C66")


def test_di_301_timed_shapes_full_grid():
    """The damage inversion at its own grid (301^2 over 0.4 m, dt = 30 ns, the
    centred Gaussian force, rho 1610, frames every 5 steps, six per-cell fields
    with three damage ellipses), 600 of its 4001 steps: the shapes bench.py
    times on di-301 (forward (1, 4, 4), adjoint (1, 2, 4)) vs the one-step
    kernels (frames bitwise, gradients within 1e-5) and vs the reference's op
    sequence with autograd through clamp(exp(theta)) (DI:304-322,
    DI:100-176, DI:331-393): frames within 1e-5 relative L2, gradients of the
    six log-fields within 1e-4."""
    n, nt, P = 301, 600, 5
    h, dt = 0.4 / (n - 1), 120e-6 / 4000
    D = ernn.ellipse_damage(n, n, 3, np.random.default_rng(5), semi=(12, 30), level=(0.5, 0.8), taper=3)
    true = [(D * c).astype(np.float32) for c in DI_TRUE]
    nf = (nt + P - 1) // P
    rng = np.random.default_rng(29)
    Gx = torch.tensor(rng.standard_normal((1, nf, n, n)), dtype=torch.float32, device=DEV)
    Gy = torch.tensor(rng.standard_normal((1, nf, n, n)), dtype=torch.float32, device=DEV)

    def make():
        sim = ernn.SpatiallyVaryingAnisotropicSimulator((n, n), h, dt, nt, P, total_time=(nt - 1) * dt,
                                                        rho_known=1610.0).to(DEV)
        with torch.no_grad():
            for k, f in zip(DI_NAMES, true):
                getattr(sim, "log_" + k).copy_(torch.tensor(f, device=DEV).log().unsqueeze(0))
        return sim

    def hip(fwd, adj):
        sim = make()
        eng = sim.engine(False)
        eng.set_blocking(*fwd)
        eng.set_blocking(*adj, which=1)
        assert eng.info.ckpt_every == nt and not eng.info.fuse   # every state kept, nothing recomputed
        ux, uy = sim()
        ((ux * Gx).sum() + (uy * Gy).sum()).backward()
        grads = [getattr(sim, "log_" + k).grad.detach().double().cpu().numpy() for k in DI_NAMES]
        out = ux.detach().cpu().numpy(), uy.detach().cpu().numpy(), grads
        eng.close()
        return out

    x1, y1, g1 = hip((0, 0, 0), (0, 0, 0))
    xt, yt, gt = hip((1, 4, 4), (1, 2, 4))
    assert np.array_equal(x1, xt) and np.array_equal(y1, yt)
    e_ts = max(rel_l2(a, b) for a, b in zip(gt, g1))

    sim = make()
    thetas = [getattr(sim, "log_" + k) for k in DI_NAMES]
    C = sim.get_material_params()[:6]
    ux, uy = pt.forward(list(C), h, dt, nt, P, sim.source_signal, sim.gaussian_dist, sim.rho, per_cell=True)
    ((ux * Gx).sum() + (uy * Gy).sum()).backward()
    gref = [t.grad.detach().double().cpu().numpy() for t in thetas]
    e_tr = max(rel_l2(x1, ux.detach().cpu().numpy()), rel_l2(y1, uy.detach().cpu().numpy()))
    e_g1 = max(rel_l2(a, b) for a, b in zip(g1, gref))
    e_gt = max(rel_l2(a, b) for a, b in zip(gt, gref))
    print(f"di-301 301^2 x {nt}: timed shapes grads vs one step {e_ts:.2e}; vs plain autograd: frames {e_tr:.2e}, "
          f"log-field grads one step {e_g1:.2e}, timed {e_gt:.2e}")
    assert e_ts < 1e-5
    assert e_tr < 1e-5
    assert e_g1 < 1e-4 and e_gt < 1e-4


# ------------------------------------------------------------- batched-uneven
def test_batched_uneven_full_workload():
    """The batched-uneven workload itself: 24 seeded instances of
    {256, 512, 768, 1024}^2 x 2000 steps, each with its own constants and
    source, 256 receivers, in one plan with the timed shapes.  Every
    instance's traces and gradient equal its lone run (same shapes and K) bit
    for bit; one 256^2 and one 512^2 instance against plain autograd.  The
    reference runs one instance per call (batch_size = 1,
    Parameter_inversion/Inversion_adam.py:252)."""
    wl = configs.batched_uneven()
    sv = wl.survey
    Cs = wl.C_true * (1.0 + np.random.default_rng(5).uniform(-0.05, 0.05, wl.C_true.shape))
    Cs[:, 3] = 0.3e9
    Cs[:, 4] = -0.2e9
    sizes = [s.nx for s in sv.shots]
    K = autotune.choose_ckpt(_variant(sv, **TIMED["batched-uneven"]))
    bsv = _variant(sv, ckpt_every=K, **TIMED["batched-uneven"])
    probe = ernn.Engine(bsv, device=DEV)
    off = list(probe.trace_off)
    probe.close()
    G = np.random.default_rng(17).standard_normal(off[-1])
    tr, (gp, _), info = _run(bsv, Cs, None, G)
    assert info[0] == (1, 8, 5) and info[1] == (1, 8, 5)
    gp = gp.reshape(len(sv.shots), 6)
    for b, s in enumerate(sv.shots):
        one = ernn.Survey([dataclasses.replace(s, set=0)], mode="const6", n_sets=1, ckpt_every=K,
                          use_graph=True, **TIMED["batched-uneven"])
        t1, (g1, _), _ = _run(one, Cs[b], None, G[off[b]:off[b + 1]])
        assert np.array_equal(t1, tr[off[b]:off[b + 1]]), b
        assert np.array_equal(g1, gp[b]), (b, g1, gp[b])
    print(f"batched-uneven: 24 instances ({sorted(set(sizes))} cells per side, K = {K}) equal their lone runs "
          f"bit for bit")
    from oracle import ernn_fp64 as orc
    for n in (256, 512):
        b = sizes.index(n)
        s = sv.shots[b]
        lone = configs.Workload("one", ernn.Survey([dataclasses.replace(s, set=0)], mode="const6"), Cs[b:b + 1])
        Gb = G[off[b]:off[b + 1]]
        p_tr, p_g = _plain_const(lone, Cs[b], Gb)
        e_tr = rel_l2(tr[off[b]:off[b + 1]].reshape(p_tr.shape), p_tr)
        e_g = _max_rel(gp[b], p_g)
        # the FP64 forward and discrete adjoint of the same instance decide between the two FP32 paths
        C64 = np.asarray(Cs[b], np.float32).astype(np.float64)
        ref, st = orc.forward(C64, s.h, s.dt, s.rho, s.nt, s.nx, s.ny, s.sources, s.wavelet.astype(np.float64), 1,
                              s.receivers, keep_states=True)
        g64 = orc.const_gradient(orc.adjoint_gradient(C64, s.h, s.dt, s.rho, s.nt, s.nx, s.ny, st,
                                                      Gb.reshape(ref.shape), 1, s.receivers))
        del st
        e_h64, e_p64 = _max_rel(gp[b], g64), _max_rel(p_g, g64)
        print(f"batched-uneven instance {b} ({n}^2 x {s.nt}): traces vs plain autograd {e_tr:.2e}; gradient: HIP vs "
              f"plain FP32 autograd {e_g:.2e}, HIP vs FP64 {e_h64:.2e}, plain FP32 autograd vs FP64 {e_p64:.2e}")
        assert e_tr < 1e-5
        _assert_gradient_bar(e_g, e_h64, e_p64)


# ----------------------------------------------- full time length, real K
@pytest.mark.parametrize("name,nt,K", [("damage-1024", 3000, 375), ("aniso-2048", 4000, 364)])
def test_full_length_gradient_vs_fp64_oracle(name, nt, K):
    """Gradient accuracy over the workloads' full step counts with the
    checkpoint intervals their plans use (damage-1024: 3000 steps, K = 375
    under its 32 GB cap; aniso-2048: 4000 steps, K = 364 within 80 % of the
    HBM), on a 128^2 crop of their physics (h,
    dt, C0, damage ellipses), two shots, in damage-field mode: HIP (timed
    shapes and one step per launch) vs the FP64 forward and discrete adjoint
    (oracle/ernn_fp64.py) on the fields C_p = D C0_p.  Bars: dJ/dC0 max
    relative and dJ/dD relative L2 below 1e-4."""
    from oracle import ernn_fp64 as orc
    n = 128
    if name == "damage-1024":
        wl = configs.damage_1024(nt=nt, n=n, shots=2, fixed_c0=False)
    else:
        wl = configs.aniso_2048(nt=nt, n=n, shots=2)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32)
    probe = ernn.Engine(_variant(sv, ckpt_every=K), device=DEV)
    off = list(probe.trace_off)
    probe.close()
    G = np.random.default_rng(23).standard_normal(off[-1])
    cells = [D.astype(np.float64) * float(c) for c in C0]
    gcell, e_or = 0.0, 0.0
    results = {}
    for label, kw in (("one step", {}), ("timed", TIMED[name])):
        print(f"{name} {nt} steps K {K}: HIP {label}", flush=True)
        tr, g, info = _run(_variant(sv, ckpt_every=K, **kw), C0, D.reshape(-1), G)
        assert info[4] == K
        results[label] = (tr, g)
    for b, s in enumerate(sv.shots):
        print(f"{name} {nt} steps: FP64 oracle shot {b}", flush=True)
        ref, st = orc.forward(cells, s.h, s.dt, s.rho, s.nt, n, n, s.sources, s.wavelet.astype(np.float64), 1,
                              s.receivers, keep_states=True)
        e_or = max(e_or, rel_l2(results["timed"][0][off[b]:off[b + 1]].reshape(ref.shape), ref))
        gcell = gcell + orc.adjoint_gradient(cells, s.h, s.dt, s.rho, s.nt, n, n, st,
                                             G[off[b]:off[b + 1]].reshape(ref.shape), 1, s.receivers)
        del st
    gC, gD = orc.dfield_gradient(gcell, C0.astype(np.float64), D.astype(np.float64))
    errs = []
    for label, (tr, (gc, gd)) in results.items():
        e_c, e_d = _max_rel(gc, gC), rel_l2(gd.reshape(n, n), gD)
        print(f"{name} crop {n}^2 x {nt}, K {K}, {label}: traces vs FP64 {e_or:.2e}, dJ/dC0 max rel {e_c:.2e}, "
              f"dJ/dD rel-L2 {e_d:.2e}")
        errs += [e_c, e_d]
    assert max(errs) < 1e-4
    assert np.array_equal(results["one step"][0], results["timed"][0])
    assert e_or < 1e-5
