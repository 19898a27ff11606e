"""Damage field over orthotropic known constants (Survey.orthotropic).

With C16 = C26 = 0 the operator sum of each component keeps three of its six
products; the other three multiply exact zeros and add +-0, which leaves
every rounded sum unchanged.  The orthotropic plan's kernels (the
time-blocked forward's sliding-row path and the reversed adjoint) drop those
terms, so against the same plan without the promise:

  * traces bitwise equal;
  * dJ/dD bitwise equal (reversed adjoint, and the banked adjoint, whose bank
    holds the forward's operator values).

Cases: every reversed shape beside three forward shapes, with sources and
receivers on block edges and short segments (K = 7 over 47 steps), and the
banked fixed_c0 plan over every damage forward shape.
"""
import dataclasses

import numpy as np
import pytest

import ernn
from ernn import autotune, configs
from gpu_helpers import make_shot, run_hip
from test_gpu_blocking import _edge_sources, _receivers

pytestmark = pytest.mark.gpu

C_O = np.array([configs.C_ORTHO], dtype=np.float64)


def _pair(shot, fwd, adj, reverse_u, K=7):
    base = ernn.Survey([shot], mode="dfield", ckpt_every=K, tiles=(2, 2, 1, 4, 0, 0), fuse=False,
                       blocking=fwd, adj_blocking=adj, fixed_c0=True, reverse_u=reverse_u)
    return base, dataclasses.replace(base, orthotropic=True)


def _run_both(shot, fwd, adj, reverse_u, seed):
    nx, ny = shot.nx, shot.ny
    rng = np.random.default_rng(seed)
    D = rng.uniform(0.4, 1.0, (nx, ny)).astype(np.float32)
    G = rng.standard_normal((2, shot.nt, len(shot.receivers)))
    s0, s1 = _pair(shot, fwd, adj, reverse_u)
    e0, t0, g0 = run_hip(s0, params=C_O, field=D, G=G)
    e1, t1, g1 = run_hip(s1, params=C_O, field=D, G=G)
    assert e1.ortho and not e0.ortho
    assert np.array_equal(t0, t1)
    assert np.abs(g1[1]).max() > 0
    assert np.array_equal(g0[1], g1[1])


@pytest.mark.parametrize("adj", list(autotune.REV_BLOCK_SHAPES))
@pytest.mark.parametrize("fwd", [(1, 4, 4), (1, 8, 5), (2, 4, 4)])
def test_orthotropic_reversed_plan_gives_the_same_bits(fwd, adj):
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    shot = make_shot(nx, ny, nt, h, C=configs.C_ORTHO, sources=_edge_sources(nx, ny, h, 1, adj[1], adj[2]),
                     receivers=_receivers(nx, ny, 1, adj[1], adj[2]))
    _run_both(shot, fwd, adj, True, seed=71)


@pytest.mark.parametrize("fwd", [s for s in autotune.BLOCK_SHAPES if s in
                                 {(1, 4, 2), (1, 4, 4), (1, 8, 2), (1, 8, 4), (1, 8, 5), (1, 8, 8),
                                  (2, 4, 2), (2, 4, 4)}])
def test_orthotropic_banked_plan_gives_the_same_bits(fwd):
    nx, ny, nt, h = 45, 150, 47, 0.2e-3
    shot = make_shot(nx, ny, nt, h, C=configs.C_ORTHO, sources=_edge_sources(nx, ny, h, 1, fwd[1], fwd[2]),
                     receivers=_receivers(nx, ny, 1, fwd[1], fwd[2]))
    _run_both(shot, fwd, (1, 4, 4), False, seed=72)
