"""Host logic of the autotuner (ernn/autotune.py) on a stub engine, CPU only:
candidate sets follow the built kernel shapes per mode and schedule, the
median over interleaved rounds picks the winner, and the winner is applied."""
import types

import pytest
import torch

from ernn import autotune


class StubEngine:
    """Records shape switches; time_steps returns a cost model of the current shape."""

    def __init__(self, mode, fuse=False, cost=None, reverse_u=False):
        self.survey = types.SimpleNamespace(mode=mode)
        self.info = types.SimpleNamespace(fuse=int(fuse))
        self.reverse_u = reverse_u
        self.trace_floats = 4
        self.device = torch.device("cpu")
        self.tiling = {0: None, 1: None}
        self.blocking = {0: (0, 0, 0), 1: (0, 0, 0)}
        self.calls = []
        self.cost = cost or (lambda which, shape: 1.0)

    def acquire_workspace(self):
        return object()

    def release_workspace(self, ws):
        pass

    def set_tiling(self, kernel, v, r, b):
        self.tiling[kernel] = (v, r, b)

    def set_blocking(self, v, rw, ks, which=0):
        self.blocking[which] = (v, rw, ks)

    def time_steps(self, which, n, *args):
        self.calls.append(which)
        if which in (3, 4):
            shape = self.blocking[which - 3]
            return n * shape[2] * self.cost(which, shape)
        shape = self.tiling[1 if which == 2 else which]
        return n * self.cost(which, shape)


def _tried(eng, which):
    seen = []
    orig = eng.set_blocking

    def rec(v, rw, ks, which=0):
        seen.append(((v, rw, ks), which))
        orig(v, rw, ks, which=which)
    eng.set_blocking = rec
    return seen


@pytest.mark.parametrize("mode,fuse,expect", [
    ("const6", False, set(autotune.ADJ_BLOCK_SHAPES)),
    ("dfield", False, {s for s in autotune.ADJ_BLOCK_SHAPES if s[0] == 1 and s[2] != 5}),
    ("const6", True, set(autotune.FUSED_BLOCK_SHAPES)),
    ("dfield", True, {s for s in autotune.FUSED_BLOCK_SHAPES if s[1] == 4}),
    ("field6", False, set(autotune.ADJ_BLOCK_SHAPES_FIELD6)),
    ("field6", True, set()),
])
def test_adjoint_blocking_candidates_follow_the_built_shapes(mode, fuse, expect):
    eng = StubEngine(mode, fuse)
    seen = _tried(eng, 1)
    autotune.tune_blocking(eng, None, None, None, steps=8, which=1, rounds=2)
    shapes = {s for s, w in seen if w == 1} - {(0, 0, 0)}
    assert shapes == expect
    # on a fuse plan the one-step-per-launch candidate is the fused single-step kernel
    assert (2 in eng.calls) == fuse and (1 in eng.calls) == (not fuse)


def test_reverse_u_adjoint_candidates_are_the_reversed_shapes_only():
    eng = StubEngine("dfield", reverse_u=True, cost=lambda which, s: 0.5 if s == (1, 6, 4) else 1.0)
    seen = _tried(eng, 1)
    v, rw, ks, ms, times = autotune.tune_blocking(eng, None, None, None, steps=8, which=1, rounds=2)
    # no single-step candidate: a reverse_u plan has no one-step adjoint
    assert {s for s, w in seen if w == 1} == set(autotune.REV_BLOCK_SHAPES) and set(times) == set(autotune.REV_BLOCK_SHAPES)
    assert (v, rw, ks) == (1, 6, 4) and eng.blocking[1] == (1, 6, 4) and set(eng.calls) == {4}


def test_forward_blocking_candidates_exclude_wide_rows_for_field_modes():
    eng = StubEngine("dfield")
    seen = _tried(eng, 0)
    autotune.tune_blocking(eng, None, None, None, steps=8, which=0, rounds=1)
    shapes = {s for s, w in seen if w == 0} - {(0, 0, 0)}
    assert shapes == {s for s in autotune.BLOCK_SHAPES if s[:2] != (2, 8)}


def test_blocking_keeps_the_median_fastest_and_applies_it():
    best = (1, 8, 5)
    eng = StubEngine("const6", cost=lambda which, s: 0.5 if s == best else (1.0 if s is None else 1.0 + 0.01 * sum(s)))
    v, rw, ks, ms, times = autotune.tune_blocking(eng, None, None, None, steps=8, which=0, rounds=3)
    assert (v, rw, ks) == best and eng.blocking[0] == best
    assert ms == pytest.approx(0.5) and times[(0, 0, 0)] == pytest.approx(1.0)


def test_tiles_keep_the_fastest_per_kernel_and_respect_register_limits():
    eng = StubEngine("field6", cost=lambda which, s: 1.0 / (1 + s[0] * s[1]) if s else 1.0)
    res = autotune.tune_tiles(eng, None, None, None, steps=4, which=(0, 1), rounds=1)
    for w in (0, 1):
        v, r, b, ms, times = res[w]
        assert eng.tiling[w] == (v, r, b)
        assert ms == min(times.values())
    # six-field limits: forward V*R <= 8, adjoint V*R <= 4 and R <= 2
    assert all(c[0] * c[1] <= 8 for c in res[0][4])
    assert all(c[0] * c[1] <= 4 and c[1] <= 2 for c in res[1][4])
