"""CPU tests of the host side: the C-ABI library loads and exports every
declared symbol, descriptor validation rejects bad input without touching a
GPU, and the problem / autotune helpers compute what they document."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import ernn
from ernn import _lib, autotune, configs
from ernn.problem import Survey

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "ernn.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ernn_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol():
    lib = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS), "ctypes binding and header disagree"
    assert lib.ernn_abi_version() == _lib.ERNN_ABI_VERSION


def _desc(inst, src=(), rec=(), **kw):
    n = len(inst)
    I = (_lib.Instance * n)()
    for k, d in enumerate(inst):
        for f, v in d.items():
            setattr(I[k], f, v)
    S = (_lib.Source * max(len(src), 1))()
    for k, (i, j, w, c) in enumerate(src):
        S[k].i, S[k].j, S[k].w, S[k].comp = i, j, w, c
    R = (_lib.Receiver * max(len(rec), 1))()
    for k, (i, j) in enumerate(rec):
        R[k].i, R[k].j = i, j
    d = _lib.Desc()
    d.n_inst, d.n_sets, d.param_mode, d.rec_stride, d.ckpt_every = n, 1, 0, 1, 4
    d.fwd_v, d.fwd_r, d.adj_v, d.adj_r, d.fwd_bpc, d.adj_bpc = 2, 4, 2, 2, 0, 0
    for f, v in kw.items():
        setattr(d, f, v)
    d.inst, d.src, d.n_src, d.rec, d.n_rec = I, S, len(src), R, len(rec)
    return d, (I, S, R)


GOOD = dict(nx=16, ny=16, nt=10, set=0, h=1e-3, dt=1e-8, rho=1600.0,
            src_begin=0, src_count=1, rec_begin=0, rec_count=1)


@pytest.mark.parametrize("patch,kw,msg", [
    ({"nx": 0}, {}, "grid"),
    ({"nt": 0}, {}, "nt"),
    ({"set": 1}, {}, "set"),
    ({"dt": -1.0}, {}, "finite and positive"),
    ({"src_count": 2}, {}, "source slice"),
    ({"rec_count": 3}, {}, "receiver slice"),
    ({}, {"ckpt_every": 1}, "ckpt_every"),
    ({}, {"rec_stride": 0}, "rec_stride"),
    ({}, {"param_mode": 7}, "param_mode"),
    ({}, {"fwd_v": 3}, "tile (V, R)"),
    ({}, {"adj_r": 3}, "tile (V, R)"),
    ({}, {"fwd_v": 4, "fwd_r": 8}, "tile (V, R)"),
    ({}, {"adj_bpc": -1}, "blocks per CU"),
    ({}, {"param_mode": 2, "adj_v": 2, "adj_r": 4}, "adjoint tile"),
    ({}, {"param_mode": 2, "adj_v": 1, "adj_r": 8}, "adjoint tile"),
    ({}, {"param_mode": 2, "adj_v": 1, "adj_r": 4}, "adjoint tile"),
    ({}, {"param_mode": 1, "adj_v": 1, "adj_r": 8}, "adjoint tile"),
    ({}, {"param_mode": 1, "adj_v": 4, "adj_r": 4}, "adjoint tile"),
    ({}, {"blk_v": 4, "blk_rw": 4, "blk_ks": 2}, "time-blocked"),
    ({}, {"blk_v": 1, "blk_rw": 4, "blk_ks": 8}, "time-blocked"),
    ({}, {"blk_v": 1, "blk_rw": 6, "blk_ks": 2}, "time-blocked"),
    ({}, {"blk_v": 1, "blk_rw": 4, "blk_ks": 3}, "time-blocked"),
    ({}, {"param_mode": 1, "blk_v": 2, "blk_rw": 8, "blk_ks": 4}, "time-blocked"),
    ({}, {"ablk_v": 2, "ablk_rw": 8, "ablk_ks": 4}, "time-blocked adjoint"),
    ({}, {"param_mode": 2, "fuse": 1, "ablk_v": 1, "ablk_rw": 4, "ablk_ks": 2}, "time-blocked adjoint"),
    ({}, {"param_mode": 1, "ablk_v": 2, "ablk_rw": 4, "ablk_ks": 2}, "time-blocked adjoint"),
    ({}, {"blk_v": 1, "blk_rw": 6, "blk_ks": 4}, "time-blocked forward"),
    ({}, {"ablk_v": 1, "ablk_rw": 6, "ablk_ks": 2}, "time-blocked adjoint"),
    ({}, {"ablk_v": 2, "ablk_rw": 6, "ablk_ks": 4}, "time-blocked adjoint"),
    ({}, {"param_mode": 2, "ablk_v": 1, "ablk_rw": 6, "ablk_ks": 4}, "time-blocked adjoint"),
    ({}, {"param_mode": 2, "ablk_v": 1, "ablk_rw": 8, "ablk_ks": 4}, "time-blocked adjoint"),
    ({}, {"param_mode": 2, "ablk_v": 1, "ablk_rw": 2, "ablk_ks": 8}, "time-blocked adjoint"),
    ({}, {"param_mode": 2, "ablk_v": 2, "ablk_rw": 2, "ablk_ks": 2}, "time-blocked adjoint"),
    ({}, {"fixed_c0": 1}, "damage-field mode"),
    ({}, {"param_mode": 2, "fixed_c0": 1}, "damage-field mode"),
    ({}, {"param_mode": 1, "fixed_c0": 2}, "fixed_c0 must be 0 or 1"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "fuse": 1}, "fuse = 0"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "reverse_u": 2}, "reverse_u must be 0 or 1"),
    ({}, {"param_mode": 0, "reverse_u": 1, "ablk_v": 1, "ablk_rw": 4, "ablk_ks": 4}, "damage-field mode"),
    ({}, {"param_mode": 1, "reverse_u": 1, "fuse": 1, "ablk_v": 1, "ablk_rw": 4, "ablk_ks": 4}, "fuse = 0"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "reverse_u": 1}, "needs a time-blocked adjoint of shape (1, RW, KS)"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "reverse_u": 1, "ablk_v": 1, "ablk_rw": 8, "ablk_ks": 8},
     "needs a time-blocked adjoint of shape (1, RW, KS)"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "reverse_u": 1, "ablk_v": 2, "ablk_rw": 4, "ablk_ks": 4},
     "needs a time-blocked adjoint of shape (1, RW, KS)"),
    ({}, {"param_mode": 1, "fixed_c0": 1, "ortho": 2}, "ortho must be 0 or 1"),
    ({}, {"param_mode": 1, "ortho": 1}, "fixed_c0 = 1"),
    ({}, {"param_mode": 0, "ortho": 1}, "fixed_c0 = 1"),
])
def test_plan_create_rejects_bad_descriptors(patch, kw, msg):
    inst = dict(GOOD, **patch)
    d, keep = _desc([inst], src=[(1, 2, 1.0, 1)], rec=[(0, 15)], **kw)
    h = C.c_void_p()
    rc = _lib.lib().ernn_plan_create(C.byref(d), C.byref(h))
    assert rc == -1
    assert msg in _lib.lib().ernn_last_error().decode()
    assert not h.value


def test_plan_create_rejects_out_of_grid_source_and_receiver():
    for src, rec, msg in [([(16, 2, 1.0, 1)], [(0, 15)], "source"),
                          ([(1, 2, 1.0, 2)], [(0, 15)], "source"),
                          ([(1, 2, 1.0, 1)], [(0, 16)], "receiver")]:
        d, keep = _desc([GOOD], src=src, rec=rec)
        h = C.c_void_p()
        assert _lib.lib().ernn_plan_create(C.byref(d), C.byref(h)) == -1
        assert msg in _lib.lib().ernn_last_error().decode()


def test_field_sets_must_share_grid():
    a = dict(GOOD)
    b = dict(GOOD, nx=20)
    d, keep = _desc([a, b], src=[(1, 2, 1.0, 1)], rec=[(0, 15)], param_mode=1)
    h = C.c_void_p()
    assert _lib.lib().ernn_plan_create(C.byref(d), C.byref(h)) == -1
    assert "share the grid" in _lib.lib().ernn_last_error().decode()


def test_null_arguments_are_errors():
    lib = _lib.lib()
    assert lib.ernn_plan_create(None, None) == -1
    info = _lib.PlanInfo()
    assert lib.ernn_plan_info(None, C.byref(info)) == -1
    assert lib.ernn_forward(None, None, None) == -1
    assert lib.ernn_adjoint(None, None, None) == -1
    assert lib.ernn_plan_destroy(None) == 0


def test_engine_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    wl = configs.plain_128(nt=10)
    with pytest.raises(RuntimeError, match="GPU"):
        ernn.Engine(wl.survey)


# ------------------------------------------------------------------ helpers
def test_checkpoint_interval_minimizes_state_memory():
    """fields(K) = 6 (4 + ceil(nt / K)) + 2 banks K padded fields per instance is
    smallest near K = sqrt(3 nt / banks)."""
    wl = configs.single_512(nt=2000)
    for fuse, banks in ((False, 1), (True, 2)):
        wl.survey.fuse = fuse
        assert wl.survey.fused == fuse
        K = autotune.choose_ckpt(wl.survey)
        assert autotune.fields(2000, K, banks) == min(autotune.fields(2000, k, banks) for k in range(2, 2001))
        assert abs(K - math.sqrt(3 * 2000 / banks)) < 8
    with pytest.raises(MemoryError):
        autotune.choose_ckpt(wl.survey, mem_cap_bytes=1 << 20)


def test_checkpoint_interval_fills_the_memory_budget():
    """With a budget the interval is the one that fits it and recomputes the
    fewest steps: the whole run (nothing recomputed) when it fits, the
    least-memory K when nothing larger fits, and in between the K brute force
    over every interval picks."""
    wl = configs.damage_1024()
    sv = wl.survey
    nt = sv.nt_max
    k_min = autotune.choose_ckpt(sv)
    assert autotune.choose_ckpt(sv, budget_bytes=1 << 20) == k_min
    full = autotune.state_bytes(sv, nt)
    assert full < 230 << 30                              # 8 x 1024^2 x 3000 steps of u: about 210 GB
    assert autotune.choose_ckpt(sv, budget_bytes=full) == nt
    assert autotune.recomputed_steps(nt, nt) == 0
    assert autotune.recomputed_steps(3000, 70) == 42 * 70     # 43 segments, the last (60 steps) kept
    for frac in (0.3, 0.5, 0.9):
        budget = int(full * frac)
        K = autotune.choose_ckpt(sv, budget_bytes=budget)
        assert autotune.state_bytes(sv, K) <= budget
        fit = [k for k in range(k_min, nt + 1) if autotune.state_bytes(sv, k) <= budget]
        assert autotune.recomputed_steps(nt, K) == min(autotune.recomputed_steps(nt, k) for k in fit)
        assert autotune.recomputed_steps(nt, K) <= nt - K   # the kept last segment is a full one


def test_damage_1024_fits_the_32_gb_cap():
    wl = configs.damage_1024()
    sv = wl.survey
    assert sv.ckpt_budget == configs.DAMAGE_MEM_CAP == 32 * 10**9
    # the interval the engine takes for the workload (a stated budget is a cap)
    K = autotune.choose_ckpt(sv, mem_cap_bytes=sv.ckpt_budget, budget_bytes=sv.ckpt_budget)
    assert autotune.state_bytes(sv, K) <= 32 * 10**9
    assert K == 375 and autotune.recomputed_steps(sv.nt_max, K) == 2625   # 8 segments of 375 steps
    # nothing larger fits, and no K that fits recomputes fewer steps
    fit = [k for k in range(2, sv.nt_max + 1) if autotune.state_bytes(sv, k) <= 32 * 10**9]
    assert min(autotune.recomputed_steps(sv.nt_max, k) for k in fit) == 2625
    # the least-memory interval is well under the cap; a cap below it is refused
    k_min = autotune.choose_ckpt(sv)
    assert autotune.state_bytes(sv, k_min) < 16 * 10**9
    with pytest.raises(MemoryError):
        autotune.choose_ckpt(sv, mem_cap_bytes=autotune.state_bytes(sv, k_min) - 1)


def test_reverse_u_checkpoints_every_rev_segment_steps_without_a_bank():
    """A reverse_u plan keeps S + 1 checkpoints and four ring states, no bank:
    its interval is REV_SEGMENT unless the budget forces a longer one."""
    wl = configs.damage_1024()
    sv = wl.survey
    sv.reverse_u = True
    nt = sv.nt_max
    assert autotune.fields(nt, 64, reverse_u=True) == 6 * (7 + 47)
    per_field = sum(autotune.padded_field_floats(s.nx, s.ny) * 4 for s in sv.shots)
    assert autotune.state_bytes(sv, 64) == 6 * (7 + 47) * per_field
    K = autotune.choose_ckpt(sv, mem_cap_bytes=sv.ckpt_budget, budget_bytes=sv.ckpt_budget)
    assert K == autotune.REV_SEGMENT and autotune.state_bytes(sv, K) < 12 * 10**9
    tight = autotune.state_bytes(sv, 200)
    K = autotune.choose_ckpt(sv, budget_bytes=tight)
    assert autotune.state_bytes(sv, K) <= tight < autotune.state_bytes(sv, K - 1)
    with pytest.raises(MemoryError):
        autotune.choose_ckpt(sv, mem_cap_bytes=autotune.state_bytes(sv, nt) - 1,
                             budget_bytes=autotune.state_bytes(sv, nt) - 1)
    with pytest.raises(ValueError, match="dfield"):
        ernn.Survey(sv.shots, mode="const6", reverse_u=True)
    with pytest.raises(ValueError, match="unfused"):
        ernn.Survey(sv.shots, mode="dfield", reverse_u=True, fuse=True)
    assert not ernn.Survey(sv.shots, mode="dfield", reverse_u=True).fused


def test_padded_layout_matches_library_convention():
    # pitch = 32 + round_up(ny, 256); rows round_up(nx, 16) + 3
    assert autotune.padded_field_floats(128, 128) == 131 * 288
    assert autotune.padded_field_floats(10, 256) == 19 * 288
    assert autotune.padded_field_floats(10, 257) == 19 * 544


def test_workload_presets():
    wl = configs.plain_128()
    s = wl.survey.shots[0]
    assert (s.nx, s.ny, s.nt) == (128, 128, 600)
    assert s.dt == pytest.approx(1.1094e-8, rel=1e-4)
    assert len(s.receivers) == 64 and set(s.receivers[:, 1]) == {127}
    assert s.sources[0][:2] == (64, 127)
    assert wl.survey.cell_steps() == 128 * 128 * 600
    wl = configs.batched_uneven()
    assert len(wl.survey.shots) == 24 and wl.survey.n_sets == 24
    assert {s.nx for s in wl.survey.shots} <= {256, 512, 768, 1024}
    assert np.all(np.abs(wl.C_true[:, [0, 1, 2, 5]] / np.array(configs.C_ORTHO)[[0, 1, 2, 5]] - 1) <= 0.3)
    wl = configs.damage_1024()
    assert wl.D_true.shape == (1024, 1024)
    assert 0.29 <= wl.D_true.min() <= 0.71 and wl.D_true.max() == 1.0
    assert len(wl.survey.shots) == 8 and len(wl.survey.shots[0].receivers) == 512


def test_bond_rotation_known_values_and_invariants():
    r = np.array(ernn.bond_rotate((130, 12, 6, 0, 0, 5), 30.0))
    np.testing.assert_allclose(r, [79.875, 20.875, 26.625, 37.456, 13.640, 25.625], atol=1e-3)
    assert r[0] + r[1] + 2 * r[2] == pytest.approx(154)
    assert r[0] + r[1] + 2 * r[5] == pytest.approx(152)
    # rotating back recovers the original
    np.testing.assert_allclose(ernn.bond_rotate(tuple(r), -30.0), [130, 12, 6, 0, 0, 5], atol=1e-9)
    # the fastest direction speed is rotation invariant
    C = np.array([130e9, 12e9, 6e9, 0, 0, 5e9])
    v0 = ernn.max_phase_velocity(C, 1600)
    v1 = ernn.max_phase_velocity(np.array(ernn.bond_rotate(C, 30.0)), 1600)
    assert v1 == pytest.approx(v0, rel=1e-6)
    assert v0 == pytest.approx(math.sqrt(130e9 / 1600), rel=1e-9)


def test_gaussian_sources_match_reference_formula():
    src, g = ernn.gaussian_sources(21, 21, 1.333e-3, a=5e6)
    assert g[10, 10] == pytest.approx(5e6 / math.pi, rel=1e-6)
    assert all(w > 1e-12 * g.max() for (_, _, w, _) in src)
    assert max(w for (_, _, w, _) in src) == pytest.approx(g.max())


def test_ellipse_damage_properties():
    D = ernn.ellipse_damage(200, 180, 3, np.random.default_rng(0), semi=(20, 40), level=(0.3, 0.7))
    assert D.shape == (200, 180) and D.dtype == np.float32
    assert D.max() == 1.0 and 0.3 <= D.min() <= 0.7
    assert (D < 1).mean() > 0.02


def test_integration_stub_structs_match_the_binding():
    """The ctypes stub in INTEGRATION.md declares the descriptor and plan-info
    structs with the library's layout (the binding in ernn/_lib.py)."""
    import re
    doc = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "INTEGRATION.md")).read()
    code = doc[doc.index("class Instance(C.Structure):"):doc.index("def check(rc):")]
    ns = {"C": C}
    exec(code, ns)
    for name, ref in (("Desc", _lib.Desc), ("Info", _lib.PlanInfo), ("Call", _lib.Call), ("Instance", _lib.Instance)):
        stub = ns[name]
        assert C.sizeof(stub) == C.sizeof(ref), name
        assert [f[0] for f in stub._fields_] == [f[0] for f in ref._fields_], name
    assert re.search(r"Desc\(1, 1, 0, P, 1, K(, \d+){18}, inst,", doc)


def test_orthotropic_survey_needs_known_constants_with_c16_c26_zero():
    """Survey.orthotropic is a promise about the known constants: it needs
    fixed_c0, and ERNN rejects a C_init with C16 or C26 nonzero before any
    device work."""
    from ernn.simulator import ERNN
    w = configs.damage_1024(nt=20, n=64, shots=1)
    assert w.survey.orthotropic and w.survey.fixed_c0
    assert not configs.damage_1024(nt=20, n=64, shots=1, fixed_c0=False).survey.orthotropic
    with pytest.raises(ValueError, match="fixed_c0"):
        Survey(w.survey.shots, mode="dfield", orthotropic=True)
    bad = np.array([configs.C_ORTHO]); bad[0, 3] = 1e9
    with pytest.raises(ValueError, match="C16 = C26 = 0"):
        ERNN(w.survey, bad, learn_constants=False)
