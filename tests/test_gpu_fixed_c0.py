"""Damage field over known background constants (Survey.fixed_c0).

Stiffness-degradation imaging inverts D(x) with C_p = D C0_p and C0 known, so
the adjoint needs dJ/dD(cell) = sum_m lam^{m+1} . (dt^2/rho) A(C0) u^m only.
A fixed_c0 plan banks L^m = (dt^2/rho) A(C0) u^m (the forward forms it for
its own update) instead of u^m and reads it pointwise.  Against the joint
damage plan of the same survey:

  * traces bitwise equal (the forward arithmetic is the same);
  * dJ/dD within 1e-5 (relative L2): the products lam . L^m round once more
    or less than lam . A(C0) u^m taken in the adjoint;
  * dJ/dD within 1e-4 of the FP64 discrete adjoint (oracle/ernn_fp64.py);
  * no constant gradient.

Cases cover every forward / adjoint shape pair on a grid with sources and
receivers on block edges, checkpoint intervals that leave short launches and
short segments, one checkpoint segment, uneven batches whose instances stop
inside earlier segments, whole-field frames, the stored last segment, graph
replays, the autograd module, and the damage-1024 shapes at full size and at
full length.  Reference path replaced: Damage_inversion/Inversion_adam.py:
100-176 and :355-387 with its backward at :511 (D's share).
"""
import dataclasses

import numpy as np
import pytest
import torch

import ernn
from ernn import configs
from ernn.function import ernn_traces
from gpu_helpers import C_ANISO, DEV, make_shot, oracle_run, rel_l2, run_hip
from test_gpu_blocking import _edge_sources, _receivers

pytestmark = pytest.mark.gpu

FWD = [(0, 0, 0), (1, 4, 2), (1, 4, 4), (1, 8, 5), (1, 8, 8), (2, 4, 4)]
ADJ = [(0, 0, 0), (1, 4, 2), (1, 4, 4), (1, 6, 4), (1, 8, 4), (1, 8, 8)]


def _pair(shots, fwd, adj, K, n_sets=1, snapshot=False, stride=1, graph=False, **kw):
    base = ernn.Survey(shots, mode="dfield", ckpt_every=K, tiles=(2, 2, 1, 4, 0, 0), snapshot=snapshot,
                       rec_stride=stride, fuse=False, blocking=fwd, adj_blocking=adj, n_sets=n_sets,
                       use_graph=graph)
    fixed = dataclasses.replace(base, fixed_c0=True)
    a = run_hip(base, **kw)
    b = run_hip(fixed, **kw)
    assert b[0].fixed_c0 and not a[0].fixed_c0
    return a, b


def _oracle_dD(shots, C0, field, G, off, stride=1, snapshot=False):
    """FP64 dJ/dD over the shots of one damage field (one set)."""
    from oracle import ernn_fp64 as orc
    D = np.asarray(field, np.float64)
    gD = 0.0
    for b, s in enumerate(shots):
        cells = [(np.float32(field) * np.float32(c)).astype(np.float64) for c in C0]
        _, gcell = oracle_run(s, cells, stride, snapshot, G[off[b]:off[b + 1]])
        gD = gD + orc.dfield_gradient(gcell, np.asarray(C0, np.float32).astype(np.float64), D)[1]
    return gD


@pytest.mark.parametrize("adj", ADJ)
@pytest.mark.parametrize("fwd", FWD)
def test_fixed_c0_matches_joint_plan_and_oracle(fwd, adj):
    """Every forward x adjoint shape; K = 7 with 47 steps leaves a short last
    segment and launches shorter than KS at segment ends; the recompute of each
    segment runs one step past it."""
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(31)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, 1, 4, 4),
                     receivers=_receivers(nx, ny, 1, 8, 4))
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    (e0, t0, g0), (e1, t1, g1) = _pair([shot], fwd, adj, K=7, params=C_ANISO, field=D, G=G)
    assert np.array_equal(t0, t1)
    assert g1[0] is None
    e_j = rel_l2(g1[1], g0[1])
    e_o = rel_l2(g1[1].reshape(nx, ny), _oracle_dD([shot], C_ANISO, D, G.reshape(-1), [0, G.size]))
    print(f"fixed_c0 fwd {fwd} adj {adj}: dJ/dD vs joint plan {e_j:.2e}, vs FP64 {e_o:.2e}")
    assert e_j < 1e-5
    assert e_o < 1e-4
    e0.close(); e1.close()


@pytest.mark.parametrize("K", [2, 5, 60])
def test_fixed_c0_segments_uneven_batch_and_snapshot(K):
    """Three damage fields (sets) of different grids and step counts, whole-field
    frames every 3 steps; K = 2 (one-step segments), 5, and 60 >= nt (one
    segment: the forward pass stores everything).  Two instances stop inside
    earlier segments."""
    h = 0.2e-3
    shots = [make_shot(nx, ny, nt, h, set_=k) for k, (nx, ny, nt) in enumerate(((5, 3, 20), (40, 90, 41),
                                                                                 (33, 70, 30)))]
    rng = np.random.default_rng(K)
    sv = ernn.Survey(shots, mode="dfield", n_sets=3, ckpt_every=K, snapshot=True, rec_stride=3)
    probe = ernn.Engine(sv, device=DEV)
    off = list(probe.trace_off)
    nfield = int(probe.info.field_floats)
    probe.close()
    D = rng.uniform(0.4, 1.0, nfield).astype(np.float32)
    G = rng.standard_normal(off[-1])
    P = np.tile(np.asarray(C_ANISO, np.float32), 3)
    for fwd, adj in (((0, 0, 0), (0, 0, 0)), ((1, 4, 4), (1, 4, 4)), ((1, 8, 5), (1, 8, 4))):
        (e0, t0, g0), (e1, t1, g1) = _pair(shots, fwd, adj, K, n_sets=3, snapshot=True, stride=3,
                                           params=P, field=D, G=G)
        assert np.array_equal(t0, t1)
        e_j = rel_l2(g1[1], g0[1])
        fo, go = 0, []
        for b, s in enumerate(shots):
            n = s.nx * s.ny
            go.append(_oracle_dD([s], C_ANISO, D[fo:fo + n].reshape(s.nx, s.ny), G[off[b]:off[b + 1]],
                                 [0, off[b + 1] - off[b]], stride=3, snapshot=True).reshape(-1))
            fo += n
        e_o = rel_l2(g1[1], np.concatenate(go))
        print(f"fixed_c0 K={K} fwd {fwd} adj {adj} uneven snapshot batch: dJ/dD vs joint {e_j:.2e}, vs FP64 {e_o:.2e}")
        assert e_j < 1e-5
        assert e_o < 1e-4
        e0.close(); e1.close()


@pytest.mark.parametrize("shapes", [((1, 8, 5), (1, 8, 4)), ((0, 0, 0), (0, 0, 0))])
def test_fixed_c0_stored_segment_and_graph_replay(shapes):
    """The first adjoint after a forward starts on the stored last segment, the
    next ones recompute it; with captured graphs, three replays with unrelated
    eager work between them.  Every gradient equals the first bit for bit."""
    fwd, adj = shapes
    nx, ny, nt, h = 40, 90, 37, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32))
    sv = ernn.Survey([shot], mode="dfield", ckpt_every=7, blocking=fwd, adj_blocking=adj, use_graph=True,
                     fixed_c0=True)
    eng = ernn.Engine(sv, device=DEV)
    wav = torch.tensor(sv.wavelets(), device=DEV)
    rng = np.random.default_rng(41)
    p = torch.tensor(np.asarray(C_ANISO, np.float32), device=DEV)
    f = torch.tensor(rng.uniform(0.4, 1.0, nx * ny).astype(np.float32), device=DEV)
    G = torch.tensor(rng.standard_normal(eng.trace_floats).astype(np.float32), device=DEV)
    ws = eng.acquire_workspace()
    t1 = eng.forward(p, f, wav, ws).cpu().numpy()
    gs = []
    for k in range(3):
        gp, gf = eng.adjoint(p, f, wav, G, ws)
        assert gp is None
        gs.append(gf.cpu().numpy())
        torch.randn(1000, device=DEV).sum().item()   # unrelated eager work between replays
    t2 = eng.forward(p, f, wav, ws).cpu().numpy()
    gs.append(eng.adjoint(p, f, wav, G, ws)[1].cpu().numpy())
    assert np.abs(gs[0]).max() > 0
    assert np.array_equal(t1, t2)
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])
    print(f"fixed_c0 {shapes}: stored and recomputed last segment, graph replays: equal gradients")
    eng.release_workspace(ws)
    eng.close()


def test_fixed_c0_module_autograd():
    """ERNN with learn_constants=False on a fixed_c0 survey: the damage field's
    .grad equals the joint plan's; asking for the constants' gradient is an
    error (the plan does not form it)."""
    nx, ny, nt, h = 48, 96, 60, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=ernn.row_receivers(nx, ny - 1, 12))
    D0 = np.random.default_rng(43).uniform(0.5, 1.0, (nx, ny)).astype(np.float32)
    grads = {}
    for fixed in (False, True):
        sv = ernn.Survey([shot], mode="dfield", ckpt_every=8, blocking=(1, 4, 4), adj_blocking=(1, 4, 4),
                         fixed_c0=fixed)
        m = ernn.ERNN(sv, np.asarray(C_ANISO)[None], D_init=D0, learn_constants=not fixed, device=DEV)
        tr = m()
        (tr * tr).sum().backward()
        grads[fixed] = m.damage.grad.detach().cpu().numpy()
        if fixed:
            assert m.theta.grad is None
            m.theta.requires_grad_(True)
            with pytest.raises(RuntimeError):
                m()
        m.engine.close()
    with pytest.raises(ValueError):
        ernn.ERNN(ernn.Survey([shot], mode="dfield", fixed_c0=True, ckpt_every=8), np.asarray(C_ANISO)[None],
                  D_init=D0, device=DEV)
    e = rel_l2(grads[True], grads[False])
    print(f"fixed_c0 module: dJ/dD vs joint module {e:.2e}")
    assert e < 1e-5


def test_fixed_c0_damage_1024_full_grid():
    """damage-1024's grid and physics (1024^2, two shots x 300 steps, three
    ellipses), the shapes its bench line times and the other adjoint shapes
    the tuner ranks: traces bitwise equal to the joint plan, dJ/dD within 1e-5
    of it; K = 24 (eleven recomputed segments) and K = nt (none)."""
    from test_gpu_fullsize import _run, _variant
    wl = configs.damage_1024(nt=300, shots=2, fixed_c0=False)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32).reshape(-1)
    probe = ernn.Engine(_variant(sv, ckpt_every=24), device=DEV)
    G = np.random.default_rng(47).standard_normal(probe.trace_floats)
    probe.close()
    tiles = (1, 8, 1, 4, 0, 0)
    ref_tr, ref_g, _ = _run(_variant(sv, ckpt_every=24, tiles=tiles, blocking=(1, 8, 5), adj_blocking=(1, 8, 4)),
                            C0, D, G)
    for K in (24, 300):
        for adj in ((1, 8, 4), (1, 8, 8), (1, 6, 4), (0, 0, 0)):
            sv1 = dataclasses.replace(_variant(sv, ckpt_every=K, tiles=tiles, blocking=(1, 8, 5), adj_blocking=adj),
                                      fixed_c0=True)
            print(f"damage-1024 fixed_c0 K {K} adj {adj}", flush=True)
            tr, g, info = _run(sv1, C0, D, G)
            assert info[4] == K
            assert np.array_equal(tr, ref_tr)
            assert g[0] is None
            e = rel_l2(g[1], ref_g[1])
            print(f"damage-1024 1024^2 x 300, fixed_c0, K {K}, adjoint {adj}: dJ/dD vs joint plan {e:.2e}")
            assert e < 1e-5


def test_fixed_c0_full_length_vs_fp64_oracle():
    """damage-1024's physics over its full 3000 steps with the interval its
    32 GB cap gives (K = 375), on a 128^2 crop, two shots: dJ/dD of the
    fixed_c0 plan (timed shapes and one step per launch) against the FP64
    discrete adjoint, relative L2 below 1e-4."""
    from oracle import ernn_fp64 as orc
    from test_gpu_fullsize import _run, _variant
    n, nt, K = 128, 3000, 375
    wl = configs.damage_1024(nt=nt, n=n, shots=2)
    sv = wl.survey
    assert sv.fixed_c0
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32)
    probe = ernn.Engine(_variant(sv, ckpt_every=K), device=DEV)
    off = list(probe.trace_off)
    probe.close()
    G = np.random.default_rng(53).standard_normal(off[-1])
    res = {}
    for label, kw in (("one step", {}), ("timed", dict(tiles=(1, 8, 1, 4, 0, 0), blocking=(1, 8, 5),
                                                       adj_blocking=(1, 8, 4)))):
        print(f"fixed_c0 {nt} steps K {K}: HIP {label}", flush=True)
        tr, g, info = _run(_variant(sv, ckpt_every=K, **kw), C0, D.reshape(-1), G)
        res[label] = (tr, g[1])
    cells = [D.astype(np.float64) * float(c) for c in C0]
    gcell = 0.0
    for b, s in enumerate(sv.shots):
        print(f"fixed_c0 {nt} steps: FP64 oracle shot {b}", flush=True)
        ref, st = orc.forward(cells, s.h, s.dt, s.rho, s.nt, n, n, s.sources, s.wavelet.astype(np.float64), 1,
                              s.receivers, keep_states=True)
        gcell = gcell + orc.adjoint_gradient(cells, s.h, s.dt, s.rho, s.nt, n, n, st,
                                             G[off[b]:off[b + 1]].reshape(ref.shape), 1, s.receivers)
        del st
    gD = orc.dfield_gradient(gcell, C0.astype(np.float64), D.astype(np.float64))[1]
    assert np.array_equal(res["one step"][0], res["timed"][0])
    for label, (tr, gd) in res.items():
        e = rel_l2(gd.reshape(n, n), gD)
        print(f"fixed_c0 crop {n}^2 x {nt}, K {K}, {label}: dJ/dD rel-L2 vs FP64 {e:.2e}")
        assert e < 1e-4
