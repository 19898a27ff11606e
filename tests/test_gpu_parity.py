"""HIP kernels vs the FP64 oracle and vs each other, through the C ABI.

Tolerances: FP32 traces within 2e-6 relative L2 of the FP64 oracle at these
sizes (the FP32 rounding drift itself); gradients within 1e-4 relative (the
bar for the adjoint); bitwise equality where the arithmetic per cell is the
same by construction (forward tilings, checkpoint intervals, batching, graph
replay).  The adjoint contracts multiply-adds, so its rounding may depend on
the adjoint tile; adjoint bit-equality is asserted only at equal adjoint tiles.
"""
import numpy as np
import pytest
import torch

import ernn
from oracle import ernn_fp64 as orc
from gpu_helpers import C_ANISO, DEV, make_shot, oracle_run, rel_l2, run_hip

pytestmark = pytest.mark.gpu

F32 = lambda a: np.asarray(a, np.float32).astype(np.float64)


def _receivers(nx, ny):
    r = [(0, ny - 1), (nx - 1, ny - 1), (nx // 2, ny - 1), (nx // 2, ny - 1), (0, 0), (nx - 1, 0),
         (nx // 3, ny // 2), (nx - 2, 1)]
    r += [(i, ny - 1) for i in range(1, nx - 1, 3)]
    return np.array(r, dtype=np.int32)


def _sources(nx, ny, h):
    w = float(np.float32(1.0 / h ** 2))
    return [(nx // 2, ny - 1, w, 1), (0, 0, 0.5 * w, 0), (nx - 1, ny - 1, -0.25 * w, 1),
            (nx // 2, ny - 1, 0.1 * w, 0), (nx // 3, ny // 2, 0.3 * w, 1)]


GRIDS = [(37, 70), (20, 130), (64, 64), (5, 3)]
# (V, R, blocks per CU): one tile per wavefront, and persistent wavefronts looping over tiles
TILES = [(1, 1, 0), (1, 2, 1), (2, 8, 0), (4, 4, 2), (1, 8, 0), (2, 1, 1)]


@pytest.mark.parametrize("nx,ny", GRIDS)
@pytest.mark.parametrize("stride,snapshot", [(1, False), (3, False), (4, True)])
def test_const6_forward_matches_oracle_and_is_tiling_invariant(nx, ny, stride, snapshot):
    h, nt = 0.2e-3, 61
    rec = None if snapshot else _receivers(nx, ny)
    shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=rec)
    ref, _ = oracle_run(shot, F32(C_ANISO), stride, snapshot)
    first = None
    for (v, r, b) in TILES:
        sv = ernn.Survey([shot], mode="const6", rec_stride=stride, snapshot=snapshot, ckpt_every=4,
                         tiles=(v, r, v, r, b, b))
        eng, tr, _ = run_hip(sv, params=C_ANISO)
        got = tr.reshape(ref.shape)
        assert rel_l2(got, ref) < 2e-6, (v, r, b)
        if first is None:
            first = tr
        else:
            assert np.array_equal(tr, first), f"tiling {(v, r, b)} changed the bits"
        eng.close()


def test_const6_gradient_matches_oracle_and_is_checkpoint_invariant():
    nx, ny, nt, h = 33, 70, 41, 0.2e-3
    shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny))
    G = np.random.default_rng(0).standard_normal((2, nt, len(shot.receivers)))
    ref, gcell = oracle_run(shot, F32(C_ANISO), 1, False, G)
    gref = orc.const_gradient(gcell)
    first = None
    for K, tiles in [(2, (2, 8, 2, 8)), (3, (1, 2, 4, 4, 1, 1)), (7, (4, 4, 1, 8)), (41, (2, 8, 2, 8, 1, 1)),
                     (64, (2, 8, 2, 8))]:
        sv = ernn.Survey([shot], mode="const6", ckpt_every=K, tiles=tiles)
        eng, tr, (gp, gf) = run_hip(sv, params=C_ANISO, G=G)
        assert gf is None
        e_tr = rel_l2(tr.reshape(ref.shape), ref)
        err = np.abs(gp - gref).max() / np.abs(gref).max()
        print(f"const6 K={K} tiles={tiles}: traces rel-L2 vs FP64 {e_tr:.2e}, dJ/dC max rel {err:.2e}")
        assert e_tr < 2e-6
        assert err < 1e-4, (K, err)
        if first is None:
            first = (tiles, gp)
        elif tiles[2:4] == first[0][2:4]:
            assert np.array_equal(gp, first[1]), f"K={K} changed the gradient bits"
        eng.close()


def test_snapshot_gradient_matches_oracle():
    nx, ny, nt, h, P = 30, 27, 33, 0.2e-3, 5
    shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h))
    nf = orc.n_frames(nt, P)
    G = np.random.default_rng(1).standard_normal((2, nf, nx, ny))
    ref, gcell = oracle_run(shot, F32(C_ANISO), P, True, G)
    sv = ernn.Survey([shot], mode="const6", rec_stride=P, snapshot=True, ckpt_every=6)
    eng, tr, (gp, _) = run_hip(sv, params=C_ANISO, G=G)
    e_tr = rel_l2(tr.reshape(ref.shape), ref)
    gref = orc.const_gradient(gcell)
    err = np.abs(gp - gref).max() / np.abs(gref).max()
    print(f"snapshot P={P}: traces rel-L2 vs FP64 {e_tr:.2e}, dJ/dC max rel {err:.2e}")
    assert e_tr < 2e-6
    assert err < 1e-4


def test_dfield_forward_and_gradients_match_oracle():
    nx, ny, nt, h = 40, 72, 45, 0.2e-3
    rng = np.random.default_rng(2)
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    shots = [make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny)),
             make_shot(nx, ny, nt, h, sources=ernn.point_source(3, ny - 1, h), receivers=_receivers(nx, ny))]
    sv = ernn.Survey(shots, mode="dfield", ckpt_every=5, tiles=(2, 4, 1, 4, 1, 1))
    Gs = [rng.standard_normal((2, nt, len(s.receivers))) for s in shots]
    C32 = F32(C_ANISO)
    cells = [np.float32(D) * np.float32(c) for c in C_ANISO]     # fl32(D * C0) like the kernel
    cells = [c.astype(np.float64) for c in cells]
    refs, gcells = zip(*[oracle_run(s, cells, 1, False, G) for s, G in zip(shots, Gs)])
    gcell = sum(gcells)
    gC0, gD = orc.dfield_gradient(gcell, C32, D.astype(np.float64))
    eng, tr, (gp, gf) = run_hip(sv, params=C_ANISO, field=D, G=np.concatenate([g.ravel() for g in Gs]))
    for b, ref in enumerate(refs):
        got = tr[eng.trace_off[b]:eng.trace_off[b + 1]].reshape(ref.shape)
        assert rel_l2(got, ref) < 2e-6
    print(f"dfield: traces rel-L2 vs FP64 {max(rel_l2(tr[eng.trace_off[b]:eng.trace_off[b + 1]].reshape(r.shape), r) for b, r in enumerate(refs)):.2e}, "
          f"dJ/dC0 max rel {np.abs(gp - gC0).max() / np.abs(gC0).max():.2e}, "
          f"dJ/dD rel-L2 {rel_l2(gf.reshape(nx, ny), gD):.2e}, max rel {np.abs(gf.reshape(nx, ny) - gD).max() / np.abs(gD).max():.2e}")
    assert np.abs(gp - gC0).max() / np.abs(gC0).max() < 1e-4
    assert rel_l2(gf.reshape(nx, ny), gD) < 1e-4
    assert np.abs(gf.reshape(nx, ny) - gD).max() / np.abs(gD).max() < 1e-4


def test_field6_forward_and_gradients_match_oracle():
    nx, ny, nt, h = 36, 66, 40, 0.2e-3
    rng = np.random.default_rng(3)
    fields = np.stack([rng.uniform(0.6, 1.0, (nx, ny)) * c for c in C_ANISO]).astype(np.float32)
    shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny))
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    ref, gcell = oracle_run(shot, [f.astype(np.float64) for f in fields], 1, False, G)
    for tiles in [(1, 2, 1, 2), (2, 4, 2, 2), (4, 2, 4, 1, 1, 1), (1, 1, 2, 1, 2, 2), (1, 8, 1, 2)]:
        sv = ernn.Survey([shot], mode="field6", ckpt_every=6, tiles=tiles)
        eng, tr, (gp, gf) = run_hip(sv, field=fields, G=G)
        assert gp is None
        assert rel_l2(tr.reshape(ref.shape), ref) < 2e-6
        gf = gf.reshape(6, nx, ny)
        worst = max(np.abs(gf[p] - gcell[p]).max() / np.abs(gcell[p]).max() for p in range(6))
        print(f"field6 tiles={tiles}: traces rel-L2 vs FP64 {rel_l2(tr.reshape(ref.shape), ref):.2e}, "
              f"dJ/dC_p(cell) worst max rel {worst:.2e}")
        for p in range(6):
            assert np.abs(gf[p] - gcell[p]).max() / np.abs(gcell[p]).max() < 1e-4, (tiles, p)
        eng.close()


def test_uneven_batch_equals_single_instance_runs():
    """Instances of different sizes, step counts and parameter sets in one launch
    give exactly what each gives alone; shots of one set sum their gradients."""
    h = 0.2e-3
    rng = np.random.default_rng(4)
    spec = [(40, 70, 50, 0), (64, 64, 37, 1), (23, 130, 50, 2), (40, 70, 44, 0)]
    shots = [make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny), set_=s)
             for nx, ny, nt, s in spec]
    Cs = np.stack([C_ANISO * rng.uniform(0.8, 1.2, 6) for _ in range(3)]).astype(np.float32)
    Gs = [rng.standard_normal((2, s.nt, len(s.receivers))) for s in shots]
    sv = ernn.Survey(shots, mode="const6", n_sets=3, ckpt_every=6, tiles=(1, 4, 2, 2, 1, 1))
    eng, tr, (gp, _) = run_hip(sv, params=Cs, G=np.concatenate([g.ravel() for g in Gs]))
    gp = gp.reshape(3, 6)
    gsum = np.zeros((3, 6))
    for b, s in enumerate(shots):
        one = ernn.Shot(**{**s.__dict__, "set": 0})
        e1, t1, (g1, _) = run_hip(ernn.Survey([one], mode="const6", ckpt_every=6, tiles=sv.tiles),
                                  params=Cs[s.set], G=Gs[b])
        assert np.array_equal(tr[eng.trace_off[b]:eng.trace_off[b + 1]], t1), f"instance {b}"
        gsum[s.set] += g1
        e1.close()
    np.testing.assert_allclose(gp, gsum, rtol=1e-12, atol=0)


def test_graph_replay_is_bitwise_equal_to_eager():
    nx, ny, nt, h = 48, 100, 60, 0.2e-3
    rng = np.random.default_rng(5)
    D = rng.uniform(0.5, 1.0, (nx, ny)).astype(np.float32)
    shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny))
    G = rng.standard_normal((2, nt, len(shot.receivers)))
    eager = run_hip(ernn.Survey([shot], mode="dfield", ckpt_every=7), params=C_ANISO, field=D, G=G)
    geng = ernn.Engine(ernn.Survey([shot], mode="dfield", ckpt_every=7, use_graph=True), device=DEV)
    for rep in range(3):
        junk = torch.randn(1 << 20, device=DEV).sum()      # unrelated eager work between replays
        _, tr, (gp, gf) = run_hip(geng.survey, params=C_ANISO, field=D, G=G, engine=geng)
        assert np.array_equal(tr, eager[1]), rep
        assert np.array_equal(gp, eager[2][0]) and np.array_equal(gf, eager[2][1]), rep
        torch.cuda.synchronize()
        del junk


def test_autograd_function_and_workspace_reuse():
    nx, ny, nt, h = 32, 64, 30, 0.2e-3
    shot = make_shot(nx, ny, nt, h, receivers=_receivers(nx, ny))
    eng = ernn.Engine(ernn.Survey([shot], mode="const6", ckpt_every=5), device=DEV)
    wav = torch.tensor(eng.survey.wavelets(), device=DEV)
    p = torch.tensor(C_ANISO, dtype=torch.float32, device=DEV, requires_grad=True)
    g1 = None
    for it in range(3):       # pooled workspaces must not leak state between calls
        p.grad = None
        tr = ernn.ernn_traces(eng, params=p, wavelets=wav)
        (tr ** 2).sum().backward()
        if g1 is None:
            g1 = p.grad.clone()
        else:
            assert torch.equal(p.grad, g1)
    # two forwards before one backward each keep their own checkpoints
    ta = ernn.ernn_traces(eng, params=p, wavelets=wav)
    tb = ernn.ernn_traces(eng, params=p * 1.01, wavelets=wav)
    p.grad = None
    (ta ** 2).sum().backward()
    assert torch.equal(p.grad, g1)
    with torch.no_grad():
        t0 = ernn.ernn_traces(eng, params=p, wavelets=wav)
    assert torch.equal(t0, ta.detach())
    del tb


def test_linearity_in_source_amplitude_is_exact():
    """Scaling the wavelet by 2 scales every FP32 operation exactly, except
    where values fall into the denormal range ahead of the wavefront (flushed
    to zero on the GPU); those differences stay at the denormal floor."""
    shot = make_shot(64, 64, 80, receivers=_receivers(64, 64))
    sv = ernn.Survey([shot], mode="const6", ckpt_every=8)
    _, t1, _ = run_hip(sv, params=C_ANISO)
    shot2 = ernn.Shot(**{**shot.__dict__, "wavelet": shot.wavelet * np.float32(2.0)})
    _, t2, _ = run_hip(ernn.Survey([shot2], mode="const6", ckpt_every=8), params=C_ANISO)
    d = np.abs(t2.astype(np.float64) - 2 * t1.astype(np.float64))
    assert d.max() <= 1e-6 * np.abs(t1).max()
    big = np.abs(t1) > 1e-6 * np.abs(t1).max()
    assert np.array_equal(t2[big], 2 * t1[big])


@pytest.mark.parametrize("mode", ["const6", "dfield"])
def test_persistent_wavefronts_are_bitwise_equal(mode):
    """A 512^2 grid has 4096 one-row tiles; capping the grid at 1 or 2 blocks
    per CU makes every wavefront loop over 2-4 tiles of its XCD band.  Traces
    and gradients must equal the one-tile-per-wavefront launch bit for bit."""
    n, nt, h = 512, 30, 0.05e-3
    rng = np.random.default_rng(6)
    rec = ernn.row_receivers(n, n - 1, 128)
    shots = [make_shot(n, n, nt, h, C=C_ANISO, receivers=rec),
             make_shot(n, n, nt, h, C=C_ANISO, sources=ernn.point_source(40, n - 1, h), receivers=rec)]
    D = rng.uniform(0.5, 1.0, (n, n)).astype(np.float32) if mode == "dfield" else None
    G = rng.standard_normal((len(shots), 2, nt, len(rec)))
    out = {}
    for bpc in (0, 1, 2):
        sv = ernn.Survey(shots, mode=mode, ckpt_every=7, tiles=(1, 1, 1, 1, bpc, bpc))
        eng, tr, grads = run_hip(sv, params=C_ANISO, field=D, G=G)
        assert int(eng.info.fwd_blocks) <= (bpc * eng.info.n_cu if bpc else 10 ** 9)
        out[bpc] = (tr, grads)
        eng.close()
    for bpc in (1, 2):
        assert np.array_equal(out[bpc][0], out[0][0]), bpc
        for a, b in zip(out[bpc][1], out[0][1]):
            assert (a is None and b is None) or np.array_equal(a, b), bpc


@pytest.mark.parametrize("mode", ["const6", "dfield", "field6"])
def test_fused_backward_equals_unfused(mode):
    """Sharing each adjoint launch with a recompute step of the next segment
    (two segment banks) gives the same traces and gradients as the plain
    schedule, including a short last segment (nt % K != 0) and K > nt / 2."""
    nx, ny, h = 40, 70, 0.2e-3
    rng = np.random.default_rng(7)
    fields = np.stack([rng.uniform(0.6, 1.0, (nx, ny)) * c for c in C_ANISO]).astype(np.float32)
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    for nt, K in [(47, 5), (40, 8), (30, 16), (25, 40)]:
        shot = make_shot(nx, ny, nt, h, sources=_sources(nx, ny, h), receivers=_receivers(nx, ny))
        G = rng.standard_normal((2, nt, len(shot.receivers)))
        kw = dict(params=None if mode == "field6" else C_ANISO,
                  field=fields if mode == "field6" else (D if mode == "dfield" else None), G=G)
        outs = []
        for fuse in (False, True):
            sv = ernn.Survey([shot], mode=mode, ckpt_every=K, tiles=(2, 2, 2, 2), fuse=fuse)
            eng, tr, grads = run_hip(sv, **kw)
            outs.append((tr, grads))
            eng.close()
        assert np.array_equal(outs[0][0], outs[1][0]), (nt, K)
        for a, b in zip(outs[0][1], outs[1][1]):
            assert (a is None and b is None) or np.array_equal(a, b), (mode, nt, K)
