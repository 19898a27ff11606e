"""Damage field over known constants without a bank (Survey.reverse_u).

A reverse_u plan stores (u, c, v) every K steps and at each instance's last
level, and nothing else.  Its adjoint runs the forward recurrence backward in
time beside lam, from the checkpoint at the top of each segment:

  L^m = (dt^2/rho) A(C0) u^m,   v^m = v^{m+1} - (D L^m + (dt^2/rho) f^m),
  (u, c)^{m-1} = (u, c)^m - v^m (compensated),   dJ/dD += lam^{m+1} . L^m.

The leapfrog step is time-symmetric, so the reversed states differ from the
forward's by rounding only (v's subtraction rounds once per step).  Against
the fixed_c0 plan that banks L^m (tests/test_gpu_fixed_c0.py) and the joint
damage plan:

  * traces bitwise equal (same forward arithmetic; the plan only stores the
    instance's last level besides);
  * dJ/dD within 1e-5 (relative L2) of the banked plan, and within 1e-4 of the
    FP64 discrete adjoint (oracle/ernn_fp64.py), for segments up to 375 steps.

Cases: every reversed shape with several forward shapes on a grid with
sources and receivers on block edges and K = 7 (short launches at segment
ends, a short last segment); uneven batches whose instances stop inside
earlier segments (their last level is a checkpoint of its own) with
whole-field frames; one segment; repeated adjoints and graph replays; the
autograd module; the damage-1024 grid at full size, and its full 3000 steps.
Reference path replaced: Damage_inversion/Inversion_adam.py:100-176 and
:355-387 with its backward at :511 (D's share).
"""
import dataclasses

import numpy as np
import pytest
import torch

import ernn
from ernn import autotune, configs
from gpu_helpers import C_ANISO, DEV, make_shot, rel_l2, run_hip
from test_gpu_blocking import _edge_sources, _receivers
from test_gpu_fixed_c0 import _oracle_dD

pytestmark = pytest.mark.gpu

REV = list(autotune.REV_BLOCK_SHAPES)
# banked damage-mode adjoint shapes (V = 1, KS != 5)
DFIELD_ADJ = {s for s in autotune.ADJ_BLOCK_SHAPES if s[0] == 1 and s[2] != 5}


def _trio(shots, fwd, adj, K, n_sets=1, snapshot=False, stride=1, **kw):
    """(banked fixed_c0 run, reverse_u run): each (engine, traces, grads).  The
    banked plan takes the reversed shape where it has that shape, else (1, 4, 4)."""
    banked = adj if adj in DFIELD_ADJ else (1, 4, 4)
    base = ernn.Survey(shots, mode="dfield", ckpt_every=K, tiles=(2, 2, 1, 4, 0, 0), snapshot=snapshot,
                       rec_stride=stride, fuse=False, blocking=fwd, adj_blocking=banked, n_sets=n_sets,
                       fixed_c0=True)
    rev = dataclasses.replace(base, reverse_u=True, adj_blocking=adj)
    a = run_hip(base, **kw)
    b = run_hip(rev, **kw)
    assert b[0].reverse_u and not a[0].reverse_u and b[0].adj_blocking == adj
    return a, b


@pytest.mark.parametrize("adj", REV)
@pytest.mark.parametrize("fwd", [(0, 0, 0), (1, 4, 4), (1, 8, 5), (2, 4, 4)])
def test_reverse_u_matches_banked_plan_and_oracle(fwd, adj):
    """Every reversed shape x four forward shapes; K = 7 with 47 steps: seven
    segments of 7 and one of 5, launches of fewer than KS steps at segment
    ends, sources and receivers on block edges."""
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(61)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, 1, adj[1], adj[2]),
                     receivers=_receivers(nx, ny, 1, adj[1], adj[2]))
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    (e0, t0, g0), (e1, t1, g1) = _trio([shot], fwd, adj, K=7, params=C_ANISO, field=D, G=G)
    assert np.array_equal(t0, t1)
    assert g1[0] is None and np.abs(g1[1]).max() > 0
    e_b = rel_l2(g1[1], g0[1])
    e_o = rel_l2(g1[1].reshape(nx, ny), _oracle_dD([shot], C_ANISO, D, G.reshape(-1), [0, G.size]))
    print(f"reverse_u fwd {fwd} adj {adj}: dJ/dD vs banked plan {e_b:.2e}, vs FP64 {e_o:.2e}")
    assert e_b < 1e-5
    assert e_o < 1e-4
    e0.close(); e1.close()


@pytest.mark.parametrize("K", [2, 5, 12, 60])
def test_reverse_u_uneven_batch_and_snapshot(K):
    """Three damage fields (sets) of different grids and step counts (20, 41,
    30), whole-field frames every 3 steps.  K = 2 and 5: the shorter instances
    stop inside earlier segments, some at a segment's last level, some inside
    one; K = 12: the 20-step instance ends in segment 1 (level 20 is its own
    checkpoint); K = 60 >= nt: one segment, every instance starts from its last
    level."""
    h = 0.2e-3
    shots = [make_shot(nx, ny, nt, h, set_=k) for k, (nx, ny, nt) in enumerate(((5, 3, 20), (40, 90, 41),
                                                                                 (33, 70, 30)))]
    rng = np.random.default_rng(K + 100)
    sv = ernn.Survey(shots, mode="dfield", n_sets=3, ckpt_every=K, snapshot=True, rec_stride=3)
    probe = ernn.Engine(sv, device=DEV)
    off = list(probe.trace_off)
    nfield = int(probe.info.field_floats)
    probe.close()
    D = rng.uniform(0.4, 1.0, nfield).astype(np.float32)
    G = rng.standard_normal(off[-1])
    P = np.tile(np.asarray(C_ANISO, np.float32), 3)
    for fwd, adj in (((0, 0, 0), (1, 4, 2)), ((1, 4, 4), (1, 4, 4)), ((1, 8, 5), (1, 8, 4)), ((1, 8, 8), (1, 6, 2))):
        (e0, t0, g0), (e1, t1, g1) = _trio(shots, fwd, adj, K, n_sets=3, snapshot=True, stride=3,
                                           params=P, field=D, G=G)
        assert np.array_equal(t0, t1)
        e_b = rel_l2(g1[1], g0[1])
        fo, go = 0, []
        for b, s in enumerate(shots):
            n = s.nx * s.ny
            go.append(_oracle_dD([s], C_ANISO, D[fo:fo + n].reshape(s.nx, s.ny), G[off[b]:off[b + 1]],
                                 [0, off[b + 1] - off[b]], stride=3, snapshot=True).reshape(-1))
            fo += n
        e_o = rel_l2(g1[1], np.concatenate(go))
        print(f"reverse_u K={K} fwd {fwd} adj {adj} uneven snapshot batch: dJ/dD vs banked {e_b:.2e}, "
              f"vs FP64 {e_o:.2e}")
        assert e_b < 1e-5
        assert e_o < 1e-4
        e0.close(); e1.close()


def _oracle_joint(shots, C0, field, G, off, stride=1, snapshot=False):
    """FP64 (dJ/dC0, dJ/dD) over the shots of one damage field (one set)."""
    from oracle import ernn_fp64 as orc
    from gpu_helpers import oracle_run
    D = np.asarray(field, np.float64)
    gcell = 0.0
    for b, s in enumerate(shots):
        cells = [(np.float32(field) * np.float32(c)).astype(np.float64) for c in C0]
        gcell = gcell + oracle_run(s, cells, stride, snapshot, G[off[b]:off[b + 1]])[1]
    return orc.dfield_gradient(gcell, np.asarray(C0, np.float32).astype(np.float64), D)


def _max_rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


REV_JOINT = [s for s in REV if s[1] != 10]


@pytest.mark.parametrize("adj", REV_JOINT)
@pytest.mark.parametrize("fwd", [(0, 0, 0), (1, 8, 5)])
def test_reverse_u_joint_matches_banked_joint_plan_and_oracle(fwd, adj):
    """The joint damage gradient (dJ/dC0 and dJ/dD) with u reversed: against the
    banked joint plan and the FP64 adjoint, every joint reversed shape, K = 7,
    47 steps, sources and receivers on block edges."""
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(97)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, 1, adj[1], adj[2]),
                     receivers=_receivers(nx, ny, 1, adj[1], adj[2]))
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    banked = adj if adj in DFIELD_ADJ else (1, 4, 4)
    base = ernn.Survey([shot], mode="dfield", ckpt_every=7, tiles=(2, 2, 1, 4, 0, 0), fuse=False, blocking=fwd,
                       adj_blocking=banked)
    e0, t0, g0 = run_hip(base, params=C_ANISO, field=D, G=G)
    e1, t1, g1 = run_hip(dataclasses.replace(base, reverse_u=True, adj_blocking=adj), params=C_ANISO, field=D, G=G)
    assert e1.reverse_u and not e1.fixed_c0 and np.array_equal(t0, t1)
    gC0, gD = _oracle_joint([shot], C_ANISO, D, G.reshape(-1), [0, G.size])
    e_c, e_d = _max_rel(g1[0], g0[0]), rel_l2(g1[1], g0[1])
    o_c, o_d = _max_rel(g1[0], gC0), rel_l2(g1[1].reshape(nx, ny), gD)
    print(f"reverse_u joint fwd {fwd} adj {adj}: vs banked dJ/dC0 {e_c:.2e} dJ/dD {e_d:.2e}; "
          f"vs FP64 {o_c:.2e} / {o_d:.2e}")
    assert e_c < 1e-5 and e_d < 1e-5
    assert o_c < 1e-4 and o_d < 1e-4
    e0.close(); e1.close()


@pytest.mark.parametrize("K", [5, 60])
def test_reverse_u_joint_uneven_batch(K):
    """Three damage fields (sets) of different grids and step counts, whole-field
    frames every 3 steps, joint gradient: per-set dJ/dC0 and every dJ/dD against
    the banked joint plan."""
    h = 0.2e-3
    shots = [make_shot(nx, ny, nt, h, set_=k) for k, (nx, ny, nt) in enumerate(((5, 3, 20), (40, 90, 41),
                                                                                 (33, 70, 30)))]
    rng = np.random.default_rng(K + 300)
    base = ernn.Survey(shots, mode="dfield", n_sets=3, ckpt_every=K, snapshot=True, rec_stride=3, fuse=False,
                       blocking=(1, 8, 5), adj_blocking=(1, 8, 4))
    probe = ernn.Engine(base, device=DEV)
    off = list(probe.trace_off)
    nfield = int(probe.info.field_floats)
    probe.close()
    D = rng.uniform(0.4, 1.0, nfield).astype(np.float32)
    G = rng.standard_normal(off[-1])
    P = np.tile(np.asarray(C_ANISO, np.float32), 3)
    e0, t0, g0 = run_hip(base, params=P, field=D, G=G)
    e1, t1, g1 = run_hip(dataclasses.replace(base, reverse_u=True, adj_blocking=(1, 8, 6)), params=P, field=D, G=G)
    assert np.array_equal(t0, t1)
    e_c, e_d = _max_rel(g1[0], g0[0]), rel_l2(g1[1], g0[1])
    print(f"reverse_u joint K={K} uneven snapshot batch: vs banked dJ/dC0 {e_c:.2e}, dJ/dD {e_d:.2e}")
    assert e_c < 1e-5 and e_d < 1e-5
    e0.close(); e1.close()


@pytest.mark.parametrize("stride", [2, 3])
def test_reverse_u_receiver_traces_with_stride(stride):
    """Receiver traces recorded every `stride` steps (frame k = u^{kP+1}), so the
    residuals enter lam only on recorded steps; K = 5 and (1,8,6) launches put
    recorded and unrecorded steps at every position within a launch."""
    nx, ny, nt, h = 40, 120, 53, 0.2e-3
    rng = np.random.default_rng(89 + stride)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, 1, 8, 6), receivers=_receivers(nx, ny, 1, 8, 6))
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    nf = (nt + stride - 1) // stride
    G = rng.standard_normal((2, nf, len(shot.receivers)))
    (e0, t0, g0), (e1, t1, g1) = _trio([shot], (1, 8, 5), (1, 8, 6), K=5, stride=stride, params=C_ANISO, field=D,
                                       G=G)
    assert np.array_equal(t0, t1) and t1.size == G.size
    e_b = rel_l2(g1[1], g0[1])
    e_o = rel_l2(g1[1].reshape(nx, ny), _oracle_dD([shot], C_ANISO, D, G.reshape(-1), [0, G.size], stride=stride))
    print(f"reverse_u stride {stride}: dJ/dD vs banked {e_b:.2e}, vs FP64 {e_o:.2e}")
    assert e_b < 1e-5
    assert e_o < 1e-4
    e0.close(); e1.close()


@pytest.mark.parametrize("fixed_c0", [True, False])
@pytest.mark.parametrize("graph", [False, True])
def test_reverse_u_repeated_adjoints_and_graph_replay(graph, fixed_c0):
    """Adjoints after one forward, with unrelated eager work between them,
    and after a second forward on the same workspace: every gradient equals
    the first bit for bit (the adjoint reads the checkpoints only; the joint
    gradient's partial slots are cleared by a fill kernel at each adjoint)."""
    nx, ny, nt, h = 40, 90, 37, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32))
    sv = ernn.Survey([shot], mode="dfield", ckpt_every=7, blocking=(1, 8, 5), adj_blocking=(1, 8, 4),
                     use_graph=graph, fixed_c0=fixed_c0, reverse_u=True)
    eng = ernn.Engine(sv, device=DEV)
    wav = torch.tensor(sv.wavelets(), device=DEV)
    rng = np.random.default_rng(67)
    p = torch.tensor(np.asarray(C_ANISO, np.float32), device=DEV)
    f = torch.tensor(rng.uniform(0.4, 1.0, nx * ny).astype(np.float32), device=DEV)
    G = torch.tensor(rng.standard_normal(eng.trace_floats).astype(np.float32), device=DEV)
    ws = eng.acquire_workspace()
    t1 = eng.forward(p, f, wav, ws).cpu().numpy()
    gs = []
    for k in range(3):
        gp, gf = eng.adjoint(p, f, wav, G, ws)
        assert (gp is None) == fixed_c0
        gs.append(np.concatenate([gf.cpu().numpy()] + ([] if fixed_c0 else [gp.cpu().numpy()])))
        torch.randn(1000, device=DEV).sum().item()   # unrelated eager work between replays
    t2 = eng.forward(p, f, wav, ws).cpu().numpy()
    gp, gf = eng.adjoint(p, f, wav, G, ws)
    gs.append(np.concatenate([gf.cpu().numpy()] + ([] if fixed_c0 else [gp.cpu().numpy()])))
    assert np.abs(gs[0]).max() > 0
    assert np.array_equal(t1, t2)
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])
    print(f"reverse_u graph={graph} fixed_c0={fixed_c0}: repeated adjoints equal bit for bit")
    eng.release_workspace(ws)
    eng.close()


def test_reverse_u_module_autograd_and_default_shape():
    """ERNN on a reverse_u survey that names no adjoint shape (the plan takes
    (1, 4, 4) and the interval REV_SEGMENT): the damage field's .grad matches
    the banked fixed_c0 module's."""
    nx, ny, nt, h = 48, 96, 150, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=ernn.row_receivers(nx, ny - 1, 12))
    D0 = np.random.default_rng(71).uniform(0.5, 1.0, (nx, ny)).astype(np.float32)
    grads = {}
    for rev in (False, True):
        sv = ernn.Survey([shot], mode="dfield", blocking=(1, 4, 4), fixed_c0=True, reverse_u=rev,
                         ckpt_every=0 if rev else 8, adj_blocking=(0, 0, 0) if rev else (1, 4, 4))
        m = ernn.ERNN(sv, np.asarray(C_ANISO)[None], D_init=D0, learn_constants=False, device=DEV)
        if rev:
            assert m.engine.reverse_u and m.engine.adj_blocking == (1, 4, 4)
            assert int(m.engine.info.ckpt_every) == autotune.REV_SEGMENT
        tr = m()
        (tr * tr).sum().backward()
        grads[rev] = m.damage.grad.detach().cpu().numpy()
        m.engine.close()
    e = rel_l2(grads[True], grads[False])
    print(f"reverse_u module: dJ/dD vs banked module {e:.2e}")
    assert e < 1e-5


def test_reverse_u_tuner_ranks_the_reversed_shapes():
    """tune_blocking on a reverse_u plan times the six reversed shapes (and no
    single-step kernel) on the real problem and applies the fastest."""
    nx, ny, nt, h = 256, 256, 64, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=ernn.row_receivers(nx, ny - 1, 16))
    sv = ernn.Survey([shot], mode="dfield", ckpt_every=32, blocking=(1, 8, 5), fixed_c0=True, reverse_u=True)
    eng = ernn.Engine(sv, device=DEV)
    wav = torch.tensor(sv.wavelets(), device=DEV)
    p = torch.tensor(np.asarray(C_ANISO, np.float32), device=DEV)
    f = torch.ones(nx * ny, device=DEV)
    v, rw, ks, ms, times = autotune.tune_blocking(eng, p, f, wav, steps=16, which=1, rounds=2)
    assert set(times) == set(autotune.REV_BLOCK_SHAPES) and eng.adj_blocking == (v, rw, ks)
    assert all(t > 0 for t in times.values())
    print("reverse_u tuner (ms per step): " + ", ".join(f"{s}: {t:.4f}" for s, t in sorted(times.items())))
    eng.close()


def test_reverse_u_damage_1024_full_grid():
    """damage-1024's grid and physics (1024^2, two shots x 300 steps, three
    ellipses): traces bitwise equal to the joint plan, dJ/dD within 1e-5 of
    it, for K = 24, REV_SEGMENT and 300 (one segment: every level rebuilt from
    the last), with the reversed shapes the bench ranks."""
    from test_gpu_fullsize import _run, _variant
    wl = configs.damage_1024(nt=300, shots=2, fixed_c0=False)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32).reshape(-1)
    probe = ernn.Engine(_variant(sv, ckpt_every=24), device=DEV)
    G = np.random.default_rng(73).standard_normal(probe.trace_floats)
    probe.close()
    tiles = (1, 8, 1, 4, 0, 0)
    ref_tr, ref_g, _ = _run(_variant(sv, ckpt_every=24, tiles=tiles, blocking=(1, 8, 5), adj_blocking=(1, 8, 4)),
                            C0, D, G)
    for K, adjs in ((24, ((1, 4, 4), (1, 8, 4))), (autotune.REV_SEGMENT, REV), (300, ((1, 4, 4),))):
        for adj in adjs:
            sv1 = dataclasses.replace(_variant(sv, ckpt_every=K, tiles=tiles, blocking=(1, 8, 5), adj_blocking=adj),
                                      fixed_c0=True, reverse_u=True)
            print(f"damage-1024 reverse_u K {K} adj {adj}", flush=True)
            tr, g, info = _run(sv1, C0, D, G)
            assert info[4] == K and info[1] == adj
            assert np.array_equal(tr, ref_tr)
            assert g[0] is None
            e = rel_l2(g[1], ref_g[1])
            print(f"damage-1024 1024^2 x 300, reverse_u, K {K}, adjoint {adj}: dJ/dD vs joint plan {e:.2e}")
            assert e < 1e-5
    # the joint gradient with u reversed
    for adj in ((1, 8, 6), (1, 8, 4), (1, 6, 4)):
        sv1 = dataclasses.replace(_variant(sv, ckpt_every=autotune.REV_SEGMENT, tiles=tiles, blocking=(1, 8, 5),
                                           adj_blocking=adj), reverse_u=True)
        print(f"damage-1024 joint reverse_u adj {adj}", flush=True)
        tr, g, info = _run(sv1, C0, D, G)
        assert np.array_equal(tr, ref_tr)
        e_c, e_d = _max_rel(g[0], ref_g[0]), rel_l2(g[1], ref_g[1])
        print(f"damage-1024 1024^2 x 300, joint reverse_u, K {autotune.REV_SEGMENT}, adjoint {adj}: vs joint plan dJ/dC0 {e_c:.2e}, "
              f"dJ/dD {e_d:.2e}")
        assert e_c < 1e-5 and e_d < 1e-5


def test_reverse_u_full_length_vs_fp64_oracle():
    """damage-1024's physics over its full 3000 steps on a 128^2 crop, two
    shots: dJ/dD of reverse_u plans with K = REV_SEGMENT and K = 375 (each
    reconstruction runs 375 steps from its checkpoint) against the FP64
    discrete adjoint, relative L2 below 1e-4; traces bitwise equal to the
    banked fixed_c0 plan's."""
    from oracle import ernn_fp64 as orc
    from test_gpu_fullsize import _run, _variant
    n, nt = 128, 3000
    wl = configs.damage_1024(nt=nt, n=n, shots=2)
    sv = wl.survey
    C0 = np.asarray(wl.C_true[0], np.float32)
    D = np.asarray(wl.D_true, np.float32)
    probe = ernn.Engine(_variant(sv, ckpt_every=375), device=DEV)
    off = list(probe.trace_off)
    probe.close()
    G = np.random.default_rng(79).standard_normal(off[-1])
    timed = dict(tiles=(1, 8, 1, 4, 0, 0), blocking=(1, 8, 5), adj_blocking=(1, 8, 4))
    print(f"reverse_u {nt} steps: banked fixed_c0 plan K 375", flush=True)
    ref_tr, ref_g, _ = _run(_variant(sv, ckpt_every=375, **timed), C0, D.reshape(-1), G)
    res = {}
    for K in (autotune.REV_SEGMENT, 375):
        print(f"reverse_u {nt} steps K {K}", flush=True)
        tr, g, info = _run(dataclasses.replace(_variant(sv, ckpt_every=K, **timed), reverse_u=True), C0,
                           D.reshape(-1), G)
        assert np.array_equal(tr, ref_tr)
        res[K] = g[1]
    print(f"reverse_u {nt} steps: joint gradient, K {autotune.REV_SEGMENT}", flush=True)
    trj, gj, _ = _run(dataclasses.replace(_variant(sv, ckpt_every=autotune.REV_SEGMENT, **timed), fixed_c0=False, orthotropic=False,
                                          reverse_u=True), C0, D.reshape(-1), G)
    assert np.array_equal(trj, ref_tr)
    cells = [D.astype(np.float64) * float(c) for c in C0]
    gcell = 0.0
    for b, s in enumerate(sv.shots):
        print(f"reverse_u {nt} steps: FP64 oracle shot {b}", flush=True)
        ref, st = orc.forward(cells, s.h, s.dt, s.rho, s.nt, n, n, s.sources, s.wavelet.astype(np.float64), 1,
                              s.receivers, keep_states=True)
        gcell = gcell + orc.adjoint_gradient(cells, s.h, s.dt, s.rho, s.nt, n, n, st,
                                             G[off[b]:off[b + 1]].reshape(ref.shape), 1, s.receivers)
        del st
    gC0, gD = orc.dfield_gradient(gcell, C0.astype(np.float64), D.astype(np.float64))
    e_jc, e_jd = _max_rel(gj[0], gC0), rel_l2(gj[1].reshape(n, n), gD)
    print(f"joint reverse_u crop {n}^2 x {nt}, K {autotune.REV_SEGMENT}: dJ/dC0 max rel vs FP64 {e_jc:.2e}, dJ/dD rel-L2 {e_jd:.2e}")
    assert e_jc < 1e-4 and e_jd < 1e-4
    e_bank = rel_l2(ref_g[1].reshape(n, n), gD)
    print(f"banked fixed_c0 crop {n}^2 x {nt}, K 375: dJ/dD rel-L2 vs FP64 {e_bank:.2e}")
    for K, gd in res.items():
        e = rel_l2(gd.reshape(n, n), gD)
        print(f"reverse_u crop {n}^2 x {nt}, K {K}: dJ/dD rel-L2 vs FP64 {e:.2e}, vs banked {rel_l2(gd, ref_g[1]):.2e}")
        assert e < 1e-4
