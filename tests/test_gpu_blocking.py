"""Time-blocked forward (KS steps per launch) vs the single-step kernel.

The blocked kernel runs each cell's arithmetic exactly as the single-step
kernel does, so traces, the checkpoints it leaves for the adjoint and the
states of the blocked segment recompute are bitwise the same; gradients are
therefore bitwise the same at equal adjoint tiles.  Cases put sources and
receivers on block edges and in block halos, cover grids smaller than one
block, step counts that are not multiples of KS or of K, uneven batches whose
instances stop at different steps, and whole-field (snapshot) recording.
"""
import numpy as np
import pytest
import torch

import ernn
from gpu_helpers import C_ANISO, DEV, make_shot, rel_l2, run_hip

pytestmark = pytest.mark.gpu

# (V, RW, KS): output tiles of (4 RW - 2 KS) x (64 V - 2 KS) cells
SHAPES = [(1, 4, 2), (1, 4, 4), (1, 8, 8), (2, 4, 4), (2, 8, 4), (1, 8, 2), (1, 8, 5)]


def _edge_sources(nx, ny, h, V, RW, KS):
    """Point sources on and next to the output-tile edges of the block grid."""
    w = float(np.float32(1.0 / h ** 2))
    OR, OC = 4 * RW - 2 * KS, 64 * V - 2 * KS
    cells = {(nx // 2, ny - 1), (0, 0), (nx - 1, ny - 1), (nx // 3, ny // 2)}
    for i in (OR - 1, OR, OR + 1, 2 * OR - 1):
        for j in (OC - 1, OC, 1, ny - 2):
            if 0 <= i < nx and 0 <= j < ny:
                cells.add((i, j))
    return [(i, j, w * (0.5 + 0.1 * k), k % 2) for k, (i, j) in enumerate(sorted(cells))]


def _receivers(nx, ny, V, RW, KS):
    OR, OC = 4 * RW - 2 * KS, 64 * V - 2 * KS
    r = {(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)}
    for i in (OR - 1, OR, nx - 1):
        for j in (OC - 1, OC, ny - 1):
            if 0 <= i < nx and 0 <= j < ny:
                r.add((i, j))
    return np.array(sorted(r), dtype=np.int32)


def _pair(shots, mode, blocking, tiles=(2, 2, 1, 4, 0, 0), K=7, snapshot=False, stride=1, fuse=False, **kw):
    base = ernn.Survey(shots, mode=mode, ckpt_every=K, tiles=tiles, snapshot=snapshot, rec_stride=stride,
                       fuse=fuse, n_sets=kw.pop("n_sets", 1))
    blk = ernn.Survey(shots, mode=mode, ckpt_every=K, tiles=tiles, snapshot=snapshot, rec_stride=stride,
                      fuse=fuse, n_sets=base.n_sets, blocking=blocking)
    a = run_hip(base, **kw)
    b = run_hip(blk, **kw)
    assert b[0].blocking == tuple(blocking)
    return a, b


# (2, 8, ·) is built for six constants only (the field modes' registers would spill)
@pytest.mark.parametrize("mode,shape", [(m, s) for m in ("const6", "dfield", "field6") for s in SHAPES
                                        if m == "const6" or s[:2] != (2, 8)])
def test_blocked_forward_and_gradients_are_bitwise_single_step(mode, shape):
    V, RW, KS = shape
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(7)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, V, RW, KS),
                     receivers=_receivers(nx, ny, V, RW, KS))
    kw = {}
    if mode == "const6":
        kw["params"] = C_ANISO
    elif mode == "dfield":
        kw["params"] = C_ANISO
        kw["field"] = rng.uniform(0.4, 1.0, (nx, ny))
    else:
        kw["field"] = np.stack([rng.uniform(0.6, 1.0, (nx, ny)) * c for c in C_ANISO])
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    tiles = (2, 2, 1, 2, 0, 0) if mode == "field6" else (2, 2, 1, 4, 0, 0)
    (e0, t0, g0), (e1, t1, g1) = _pair([shot], mode, shape, tiles=tiles, G=G, **kw)
    print(f"{mode} {shape}: blocks {e1.info.blk_blocks}, traces max |diff| {np.abs(t1 - t0).max():.3e}")
    assert np.array_equal(t0, t1)
    for x, y in zip(g0, g1):
        if x is not None:
            assert np.array_equal(x, y)
    e0.close(); e1.close()


def test_blocked_snapshot_small_grid_and_uneven_batch():
    """Whole-field frames with stride 3; a grid smaller than one block; instances
    of different sizes and step counts (each stops at its own nt) in one launch."""
    h = 0.2e-3
    shots = [make_shot(5, 3, 29, h), make_shot(40, 90, 41, h), make_shot(33, 70, 35, h)]
    G = np.random.default_rng(3).standard_normal(
        sum(2 * ((s.nt + 2) // 3) * s.nx * s.ny for s in shots))
    (e0, t0, g0), (e1, t1, g1) = _pair(shots, "const6", (1, 4, 4), snapshot=True, stride=3, K=5,
                                       params=C_ANISO, G=G)
    assert np.array_equal(t0, t1)
    assert np.array_equal(g0[0], g1[0])
    e0.close(); e1.close()


def test_blocked_forward_pass_with_fused_backward():
    """The fused backward recomputes one step per launch; the blocked forward
    pass must leave it the same checkpoints."""
    nx, ny, nt, h = 64, 128, 61, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=_receivers(nx, ny, 2, 8, 4))
    G = np.random.default_rng(5).standard_normal((2, nt, len(shot.receivers)))
    (e0, t0, g0), (e1, t1, g1) = _pair([shot], "const6", (2, 8, 4), K=6, fuse=True, params=C_ANISO, G=G)
    assert e1.info.fuse == 1
    assert np.array_equal(t0, t1)
    assert np.array_equal(g0[0], g1[0])
    e0.close(); e1.close()


def test_set_blocking_switches_and_replays_graphs():
    """Switching the shape on a live plan (as the autotuner does) and graph replay."""
    nx, ny, nt, h = 50, 200, 40, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=_receivers(nx, ny, 1, 4, 2))
    sv = ernn.Survey([shot], mode="const6", ckpt_every=8, use_graph=True)
    eng = ernn.Engine(sv, device=DEV)
    ref = run_hip(sv, params=C_ANISO, engine=eng)[1]
    for shape in [(1, 4, 2), (2, 4, 4), (1, 8, 8), (0, 0, 0)]:
        eng.set_blocking(*shape)
        for _ in range(2):
            tr = run_hip(sv, params=C_ANISO, engine=eng)[1]
            assert np.array_equal(tr, ref), shape
    eng.close()


# ------------------------------------------------------------- blocked adjoint
ADJ_SHAPES = [(1, 4, 2), (1, 4, 4), (1, 6, 4), (1, 8, 4), (1, 8, 5), (1, 8, 8), (2, 4, 4)]


def _adj_pair(shots, mode, ablk, K, snapshot=False, stride=1, **kw):
    def sv(a):
        return ernn.Survey(shots, mode=mode, ckpt_every=K, tiles=(2, 2, 1, 4, 0, 0), snapshot=snapshot,
                           rec_stride=stride, fuse=False, adj_blocking=a)
    return run_hip(sv((0, 0, 0)), **kw), run_hip(sv(ablk), **kw)


# V = 2 and KS = 5 adjoint shapes are built for six constants only
@pytest.mark.parametrize("mode,shape", [(m, s) for m in ("const6", "dfield") for s in ADJ_SHAPES
                                        if m == "const6" or (s[0] == 1 and s[2] != 5)])
def test_blocked_adjoint_matches_single_step_and_oracle(mode, shape):
    """Gradients of the time-blocked adjoint vs the single-step adjoint (same
    recurrence, gradient sums taken in another order) and vs the FP64 oracle;
    K = 7 leaves one-step launches (7 mod KS = 1 or 3) at segment ends."""
    from gpu_helpers import oracle_run
    from oracle import ernn_fp64 as orc
    V, RW, KS = shape
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(11)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, V, RW, KS),
                     receivers=_receivers(nx, ny, V, RW, KS))
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    kw = {"params": C_ANISO}
    if mode == "dfield":
        kw["field"] = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    (e0, t0, g0), (e1, t1, g1) = _adj_pair([shot], mode, shape, K=7, G=G, **kw)
    assert e1.adj_blocking == shape
    assert np.array_equal(t0, t1)
    gp0, gp1 = g0[0], g1[0]
    err_ss = np.abs(gp1 - gp0).max() / np.abs(gp0).max()
    if mode == "const6":
        _, gcell = oracle_run(shot, np.asarray(C_ANISO, np.float32).astype(np.float64), 1, False, G)
        gref = orc.const_gradient(gcell)
        err_or = np.abs(gp1 - gref).max() / np.abs(gref).max()
        err_f = 0.0
    else:
        D = kw["field"].astype(np.float64)
        cells = [(np.float32(kw["field"]) * np.float32(c)).astype(np.float64) for c in C_ANISO]
        _, gcell = oracle_run(shot, cells, 1, False, G)
        gC0, gD = orc.dfield_gradient(gcell, np.asarray(C_ANISO, np.float32).astype(np.float64), D)
        err_or = np.abs(gp1 - gC0).max() / np.abs(gC0).max()
        err_f = rel_l2(g1[1].reshape(nx, ny), gD)
        assert rel_l2(g1[1], g0[1]) < 1e-5
    print(f"{mode} adj {shape}: dJ/dC vs single-step {err_ss:.2e}, vs FP64 {err_or:.2e}, dJ/dD rel-L2 {err_f:.2e}")
    assert err_ss < 1e-5
    assert err_or < 1e-4
    assert err_f < 1e-4
    e0.close(); e1.close()


FIELD6_ADJ_SHAPES = [(1, 2, 2), (1, 2, 4), (1, 4, 2), (1, 4, 4)]


def _field6_pair(shots, ablk, K, snapshot=False, stride=1, **kw):
    def sv(a):
        return ernn.Survey(shots, mode="field6", ckpt_every=K, tiles=(2, 2, 1, 2, 0, 0), snapshot=snapshot,
                           rec_stride=stride, fuse=False, adj_blocking=a)
    return run_hip(sv((0, 0, 0)), **kw), run_hip(sv(ablk), **kw)


def _field6_errors(g_blk, g_one, gcell, nx, ny):
    a = g_blk.reshape(6, nx, ny)
    b = g_one.reshape(6, nx, ny)
    e_ss = max(rel_l2(a[p], b[p]) for p in range(6))
    e_or = max(np.abs(a[p] - gcell[p]).max() / np.abs(gcell[p]).max() for p in range(6))
    return e_ss, e_or


@pytest.mark.parametrize("shape", FIELD6_ADJ_SHAPES)
def test_blocked_six_field_adjoint_matches_single_step_and_oracle(shape):
    """The six-field time-blocked adjoint (per-cell C_p) vs the single-step
    adjoint (the same lam recurrence; the per-cell gradients are summed per
    launch before their FP64 update) and vs the FP64 oracle, with sources and
    receivers on output-tile edges and in halos (the walked receiver list);
    K = 7 leaves one-step launches at segment ends."""
    from gpu_helpers import oracle_run
    V, RW, KS = shape
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(17)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, 1, 4, KS),
                     receivers=_receivers(nx, ny, 1, 4, KS))
    fields = np.stack([rng.uniform(0.6, 1.0, (nx, ny)) * c for c in C_ANISO]).astype(np.float32)
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    (e0, t0, g0), (e1, t1, g1) = _field6_pair([shot], shape, K=7, field=fields, G=G)
    assert e1.adj_blocking == shape
    assert np.array_equal(t0, t1)
    assert g1[0] is None
    _, gcell = oracle_run(shot, [f.astype(np.float64) for f in fields], 1, False, G)
    e_ss, e_or = _field6_errors(g1[1], g0[1], gcell, nx, ny)
    print(f"field6 adj {shape}: dJ/dC_p(cell) vs single-step rel-L2 {e_ss:.2e}, vs FP64 worst max rel {e_or:.2e}")
    assert e_ss < 1e-5
    assert e_or < 1e-4
    e0.close(); e1.close()


@pytest.mark.parametrize("shape", [(1, 2, 4), (1, 4, 4)])
def test_blocked_six_field_adjoint_snapshot(shape):
    """The damage inversion's layout: whole-field residuals every P = 5 steps, a
    Gaussian force spread over several cells, a grid of a few blocks with an
    odd size; blocked vs single-step and vs the FP64 oracle."""
    from gpu_helpers import oracle_run
    nx, ny, nt, h, P = 53, 97, 43, 0.2e-3, 5
    rng = np.random.default_rng(19)
    ci, cj = nx // 2, ny // 2
    src = [(ci + di, cj + dj, float(np.exp(-0.5 * (di * di + dj * dj))) / h ** 2, 1)
           for di in (-1, 0, 1) for dj in (-1, 0, 1)]
    shot = make_shot(nx, ny, nt, h, sources=src)
    fields = np.stack([rng.uniform(0.6, 1.0, (nx, ny)) * c for c in C_ANISO]).astype(np.float32)
    nf = (nt + P - 1) // P
    G = rng.standard_normal((2, nf, nx, ny))
    (e0, t0, g0), (e1, t1, g1) = _field6_pair([shot], shape, K=9, snapshot=True, stride=P, field=fields, G=G)
    assert np.array_equal(t0, t1)
    _, gcell = oracle_run(shot, [f.astype(np.float64) for f in fields], P, True, G)
    e_ss, e_or = _field6_errors(g1[1], g0[1], gcell, nx, ny)
    print(f"field6 snapshot adj {shape}: vs single-step rel-L2 {e_ss:.2e}, vs FP64 worst max rel {e_or:.2e}")
    assert e_ss < 1e-5
    assert e_or < 1e-4
    e0.close(); e1.close()


def test_blocked_adjoint_snapshot_and_uneven_batch():
    """Whole-field residuals with stride 2; instances of different sizes and step
    counts (each starts its reverse recurrence at its own nt) in one launch."""
    h = 0.2e-3
    shots = [make_shot(5, 3, 29, h), make_shot(40, 90, 41, h), make_shot(33, 70, 36, h)]
    G = np.random.default_rng(4).standard_normal(sum(2 * ((s.nt + 1) // 2) * s.nx * s.ny for s in shots))
    (e0, t0, g0), (e1, t1, g1) = _adj_pair(shots, "const6", (1, 4, 4), K=6, snapshot=True, stride=2,
                                           params=C_ANISO, G=G)
    assert np.array_equal(t0, t1)
    err = np.abs(g1[0] - g0[0]).max() / np.abs(g0[0]).max()
    print(f"snapshot uneven batch, adj (1, 4, 4): dJ/dC vs single-step {err:.2e}")
    assert err < 1e-5
    e0.close(); e1.close()


@pytest.mark.parametrize("mode", ["const6", "dfield"])
def test_receiver_line_and_shared_cells(mode):
    """A receiver on every other row of the last column (the benchmark layout:
    served through the per-launch cell map of the blocked kernels) plus two
    receivers on one cell with a third two rows away (the blocks holding them
    walk their receiver lists instead).  Blocked forward bitwise equal to the
    single-step kernel, the shared cell's two traces equal, blocked adjoint
    within 1e-5 of the single-step adjoint."""
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rec = [(i, ny - 1) for i in range(0, nx, 2)] + [(22, 140), (22, 140), (24, 144), (30, 100)]
    shot = make_shot(nx, ny, nt, h, receivers=np.array(rec, np.int32))
    rng = np.random.default_rng(21)
    kw = {"params": C_ANISO}
    if mode == "dfield":
        kw["field"] = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    G = rng.standard_normal((2, nt, len(rec)))
    (e0, t0, g0), (e1, t1, g1) = _pair([shot], mode, (1, 4, 4), G=G, **kw)
    assert np.array_equal(t0, t1)
    tr = t1.reshape(2, nt, len(rec))
    d = len(rec) - 4
    assert np.array_equal(tr[:, :, d], tr[:, :, d + 1])
    assert np.abs(tr[:, :, d]).max() > 0
    (a0, _, h0), (a1, _, h1) = _adj_pair([shot], mode, (1, 4, 4), K=7, G=G, **kw)
    err = np.abs(h1[0] - h0[0]).max() / np.abs(h0[0]).max()
    print(f"{mode} receiver line + shared cell: traces bitwise, dJ/dC blocked vs single-step {err:.2e}")
    assert err < 1e-5
    if mode == "dfield":
        assert rel_l2(h1[1], h0[1]) < 1e-5
    for e in (e0, e1, a0, a1):
        e.close()


# ------------------------------------------------------ fused time-blocked backward
# fused (1, 8, 4) is built for six constants only (the damage kernel would spill)
@pytest.mark.parametrize("mode,shape", [("const6", (1, 4, 2)), ("const6", (1, 4, 4)), ("const6", (1, 8, 4)),
                                        ("dfield", (1, 4, 2)), ("dfield", (1, 4, 4))])
def test_fused_blocked_backward_matches_unfused_and_oracle(mode, shape):
    """Each launch runs KS reverse steps of segment s beside KS recompute steps
    of segment s - 1.  K = 7 with 47 steps leaves a short last segment (its
    adjoint launches cannot carry all of segment s - 1's recompute, which then
    runs alone) and one-step launches (7 mod KS); gradients vs the unfused
    single-step backward (same recurrence) and the FP64 oracle."""
    from gpu_helpers import oracle_run
    from oracle import ernn_fp64 as orc
    V, RW, KS = shape
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(13)
    shot = make_shot(nx, ny, nt, h, sources=_edge_sources(nx, ny, h, V, RW, KS),
                     receivers=_receivers(nx, ny, V, RW, KS))
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    kw = {"params": C_ANISO}
    if mode == "dfield":
        kw["field"] = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)

    def sv(fuse, ablk):
        return ernn.Survey([shot], mode=mode, ckpt_every=7, tiles=(2, 2, 1, 4, 0, 0), fuse=fuse,
                           adj_blocking=ablk, blocking=(1, 4, 4))
    e0, t0, g0 = run_hip(sv(False, (0, 0, 0)), G=G, **kw)
    e1, t1, g1 = run_hip(sv(True, shape), G=G, **kw)
    assert e1.info.fuse == 1 and e1.adj_blocking == shape
    assert np.array_equal(t0, t1)
    err_ss = np.abs(g1[0] - g0[0]).max() / np.abs(g0[0]).max()
    if mode == "const6":
        _, gcell = oracle_run(shot, np.asarray(C_ANISO, np.float32).astype(np.float64), 1, False, G)
        gref = orc.const_gradient(gcell)
        err_f = 0.0
    else:
        D = kw["field"].astype(np.float64)
        cells = [(np.float32(kw["field"]) * np.float32(c)).astype(np.float64) for c in C_ANISO]
        _, gcell = oracle_run(shot, cells, 1, False, G)
        gref, gD = orc.dfield_gradient(gcell, np.asarray(C_ANISO, np.float32).astype(np.float64), D)
        err_f = rel_l2(g1[1].reshape(nx, ny), gD)
        assert rel_l2(g1[1], g0[1]) < 1e-5
    err_or = np.abs(g1[0] - gref).max() / np.abs(gref).max()
    print(f"{mode} fused blocked {shape}: dJ/dC vs unfused single-step {err_ss:.2e}, vs FP64 {err_or:.2e}, "
          f"dJ/dD rel-L2 vs FP64 {err_f:.2e}")
    assert err_ss < 1e-5
    assert err_or < 1e-4
    assert err_f < 1e-4
    e0.close(); e1.close()


def test_fused_blocked_backward_uneven_batch_and_graph_replay():
    """Instances of different sizes and step counts in one plan; captured
    graph replays equal the eager pass."""
    h = 0.2e-3
    shots = [make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32))
             for nx, ny, nt in ((5, 3, 29), (40, 90, 41), (33, 70, 36))]
    G = np.random.default_rng(6).standard_normal(sum(2 * s.nt * len(s.receivers) for s in shots))

    def sv(fuse, ablk, graph=False):
        return ernn.Survey(shots, mode="const6", ckpt_every=6, fuse=fuse, adj_blocking=ablk, blocking=(1, 4, 4),
                           use_graph=graph)
    _, t0, g0 = run_hip(sv(False, (0, 0, 0)), params=C_ANISO, G=G)
    eng = ernn.Engine(sv(True, (1, 4, 4), graph=True), device=DEV)
    for _ in range(3):
        _, t1, g1 = run_hip(eng.survey, params=C_ANISO, G=G, engine=eng)
        assert np.array_equal(t0, t1)
        err = np.abs(g1[0] - g0[0]).max() / np.abs(g0[0]).max()
        assert err < 1e-5, err
    print(f"fused blocked uneven batch, graph replays: dJ/dC vs unfused single-step {err:.2e}")
    eng.close()


@pytest.mark.parametrize("K", [2, 3])
def test_fused_blocked_backward_short_segments_and_snapshot(K):
    """Checkpoint intervals shorter than KS = 4 (every launch of both halves is
    partial) with whole-field residuals at stride 3."""
    nx, ny, nt, h = 40, 90, 31, 0.2e-3
    shot = make_shot(nx, ny, nt, h)
    nf = (nt + 2) // 3
    G = np.random.default_rng(K).standard_normal(2 * nf * nx * ny)

    def sv(fuse, ablk):
        return ernn.Survey([shot], mode="const6", ckpt_every=K, snapshot=True, rec_stride=3, fuse=fuse,
                           adj_blocking=ablk, blocking=(1, 4, 4))
    _, t0, g0 = run_hip(sv(False, (0, 0, 0)), params=C_ANISO, G=G)
    e1, t1, g1 = run_hip(sv(True, (1, 4, 4)), params=C_ANISO, G=G)
    assert np.array_equal(t0, t1)
    err = np.abs(g1[0] - g0[0]).max() / np.abs(g0[0]).max()
    print(f"fused blocked K={K}, snapshot stride 3: dJ/dC vs unfused single-step {err:.2e}")
    assert err < 1e-5
    e1.close()


# ------------------------------------------- stored last segment (forward -> adjoint)
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("blocking,fuse,ablk", [((1, 4, 4), True, (1, 4, 4)), ((1, 4, 4), False, (1, 4, 4)),
                                                ((0, 0, 0), True, (0, 0, 0)), ((0, 0, 0), False, (0, 0, 0))])
def test_stored_last_segment_is_used_once(blocking, fuse, ablk, graph):
    """The forward pass leaves every state of the last checkpoint segment in
    its bank; the first adjoint on that workspace starts from them, a second
    adjoint (the bank was reused) or one on a workspace whose forward is not
    the latest recomputes the segment.  Every order gives the gradient of the
    fresh forward -> adjoint bit for bit.  K = 7 with 37 steps: the time-blocked
    forward's launch before the last segment is partial (33..35)."""
    nx, ny, nt, h = 40, 90, 37, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32))
    sv = ernn.Survey([shot], mode="const6", ckpt_every=7, fuse=fuse, blocking=blocking, adj_blocking=ablk,
                     use_graph=graph)
    eng = ernn.Engine(sv, device=DEV)
    wav = torch.tensor(sv.wavelets(), device=DEV)
    p1 = torch.tensor(np.asarray(C_ANISO, np.float32), device=DEV)
    p2 = p1 * 1.1
    G = torch.tensor(np.random.default_rng(8).standard_normal(eng.trace_floats).astype(np.float32), device=DEV)
    wa, wb = eng.acquire_workspace(), eng.acquire_workspace()

    def grad(p, ws):
        return eng.adjoint(p, None, wav, G, ws)[0].cpu().numpy()

    t1 = eng.forward(p1, None, wav, wa).cpu().numpy()
    g_first = grad(p1, wa)           # starts on the stored segment
    g_again = grad(p1, wa)           # recomputes it
    eng.forward(p1, None, wav, wa)
    eng.forward(p2, None, wav, wb)
    g_older = grad(p1, wa)           # wa's forward is not the latest: recomputes
    g_b = grad(p2, wb)               # stored
    g_b2 = grad(p2, wb)              # recomputed
    torch.cuda.synchronize()
    assert np.abs(t1).max() > 0 and np.abs(g_first).max() > 0
    assert np.array_equal(g_first, g_again)
    assert np.array_equal(g_first, g_older)
    assert np.array_equal(g_b, g_b2)
    assert not np.array_equal(g_b, g_first)
    print(f"blocking {blocking} fuse {fuse} adj {ablk} graph {graph}: stored / recomputed gradients equal")
    eng.release_workspace(wa)
    eng.release_workspace(wb)
    eng.close()


@pytest.mark.parametrize("mode", ["const6", "dfield"])
@pytest.mark.parametrize("fuse", [True, False])
def test_stored_last_segment_uneven_batch(mode, fuse):
    """Instances of different step counts in one plan (two stop before the last
    checkpoint segment, 5 x 3 is smaller than one block): the forward's
    last-segment launches leave them alone, and the first adjoint (stored
    segment) equals the second (recomputed) bit for bit, in both modes."""
    h = 0.2e-3
    # one parameter set per instance (a set's damage field has its grid's size)
    shots = [make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32),
                       set_=k)
             for k, (nx, ny, nt) in enumerate(((5, 3, 20), (40, 90, 41), (33, 70, 30)))]
    sv = ernn.Survey(shots, mode=mode, ckpt_every=6, fuse=fuse, blocking=(1, 4, 4), adj_blocking=(1, 4, 4),
                     use_graph=True, n_sets=3)
    eng = ernn.Engine(sv, device=DEV)
    wav = torch.tensor(sv.wavelets(), device=DEV)
    p = torch.tensor(np.tile(np.asarray(C_ANISO, np.float32), 3), device=DEV)
    rng = np.random.default_rng(12)
    f = None
    if mode == "dfield":
        f = torch.tensor(rng.uniform(0.4, 1.0, eng.info.field_floats).astype(np.float32), device=DEV)
    G = torch.tensor(rng.standard_normal(eng.trace_floats).astype(np.float32), device=DEV)
    ws = eng.acquire_workspace()
    eng.forward(p, f, wav, ws)
    g1 = [None if x is None else x.cpu().numpy() for x in eng.adjoint(p, f, wav, G, ws)]
    g2 = [None if x is None else x.cpu().numpy() for x in eng.adjoint(p, f, wav, G, ws)]
    torch.cuda.synchronize()
    for a, b in zip(g1, g2):
        if a is not None:
            assert np.abs(a).max() > 0
            assert np.array_equal(a, b)
    print(f"{mode} fuse {fuse} uneven batch: stored and recomputed last segment give equal gradients")
    eng.release_workspace(ws)
    eng.close()


@pytest.mark.parametrize("blocking,fuse,ablk", [((1, 4, 4), True, (1, 4, 4)), ((0, 0, 0), False, (0, 0, 0))])
def test_stored_last_segment_single_segment(blocking, fuse, ablk):
    """nt <= K: one checkpoint segment, so the whole forward pass runs as the
    stored last segment.  Traces equal the plan with K < nt bit for bit; the
    stored and recomputed adjoints are equal, and match the K < nt plan's."""
    nx, ny, nt, h = 40, 90, 23, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=np.array([(0, ny - 1), (nx - 1, 0), (nx // 2, ny // 2)], np.int32))
    G = np.random.default_rng(9).standard_normal((2, nt, 3))

    def sv(K):
        return ernn.Survey([shot], mode="const6", ckpt_every=K, fuse=fuse, blocking=blocking, adj_blocking=ablk,
                           use_graph=True)
    e0, t0, g0 = run_hip(sv(5), params=C_ANISO, G=G)
    eng = ernn.Engine(sv(32), device=DEV)
    assert eng.info.n_ckpt == 1
    wav = torch.tensor(eng.survey.wavelets(), device=DEV)
    p = torch.tensor(np.asarray(C_ANISO, np.float32), device=DEV)
    Gt = torch.tensor(G.astype(np.float32).reshape(-1), device=DEV)
    ws = eng.acquire_workspace()
    t1 = eng.forward(p, None, wav, ws).cpu().numpy()
    g1 = eng.adjoint(p, None, wav, Gt, ws)[0].cpu().numpy()
    g2 = eng.adjoint(p, None, wav, Gt, ws)[0].cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(t0, t1)
    assert np.array_equal(g1, g2)
    err = np.abs(g1 - g0[0]).max() / np.abs(g0[0]).max()
    print(f"one segment, blocking {blocking} fuse {fuse}: dJ/dC vs K = 5 plan {err:.2e}")
    assert err < 1e-5
    eng.release_workspace(ws)
    eng.close(); e0.close()


@pytest.mark.parametrize("shape", [(1, 4, 2), (1, 4, 4)])
@pytest.mark.parametrize("mode", ["const6", "dfield"])
def test_fused_blocked_backward_tall_recompute_blocks(mode, shape):
    """The fused kernel's recompute half on its own, taller block table (8 rows
    per wavefront beside the adjoint's 4): gradients equal the fused kernel
    with one shape bit for bit (the recomputed states are the same), and the
    unfused single-step backward within 1e-5; graph replays included.  K = 7
    with 47 steps leaves recompute launches that run alone."""
    V, RW, KS = shape
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    rng = np.random.default_rng(14)
    # sources and receivers on and next to the output-tile edges of both block tables
    # (the adjoint's 4-row bands and the recompute half's 8-row bands)
    src = sorted(set(_edge_sources(nx, ny, h, V, RW, KS)) | set(_edge_sources(nx, ny, h, V, 8, KS)))
    rec = np.unique(np.concatenate([_receivers(nx, ny, V, RW, KS), _receivers(nx, ny, V, 8, KS)]), axis=0)
    shot = make_shot(nx, ny, nt, h, sources=src, receivers=rec)
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    kw = {"params": C_ANISO}
    if mode == "dfield":
        kw["field"] = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)

    def sv(fuse, ablk, graph=False):
        return ernn.Survey([shot], mode=mode, ckpt_every=7, tiles=(2, 2, 1, 4, 0, 0), fuse=fuse,
                           adj_blocking=ablk, blocking=(1, 4, 4), use_graph=graph)
    _, t0, g0 = run_hip(sv(False, (0, 0, 0)), G=G, **kw)
    _, t1, g1 = run_hip(sv(True, shape), G=G, **kw)
    eng = ernn.Engine(sv(True, shape, graph=True), device=DEV)
    eng.set_blocking(V, 8, KS, which=2)
    assert eng.ff_rw == 8
    for _ in range(2):
        _, t2, g2 = run_hip(eng.survey, G=G, engine=eng, **kw)
        assert np.array_equal(t0, t2)
        for x, y in zip(g1, g2):
            if x is not None:
                assert np.array_equal(x, y)
    err = np.abs(g2[0] - g0[0]).max() / np.abs(g0[0]).max()
    print(f"{mode} fused {shape} with 8-row recompute blocks: equal to one-shape fused; vs single-step {err:.2e}")
    assert err < 1e-5
    eng.close()
