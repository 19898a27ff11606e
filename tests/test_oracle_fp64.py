"""Pin the FP64 oracle (oracle/ernn_fp64.py) against closed forms, exact
identities and invariants, so that a dropped term, a wrong sign, a swapped
axis or an off-by-one recording offset fails a test.  CPU only."""
import math

import numpy as np
import pytest
import torch

from oracle import ernn_fp64 as orc
from oracle import plain_torch as pt

RHO = 1600.0
C_ANISO = np.array([130e9, 12e9, 6e9, 3e9, 1.5e9, 5e9])   # C11 C22 C12 C16 C26 C66


# ---------------------------------------------------------------- sources
def test_ricker_closed_form():
    f0, t0, A = 1e6, 1e-6, 2.5
    assert orc.ricker(t0, f0, t0, A) == pytest.approx(A)                    # peak at tau = 0
    tz = t0 + 1.0 / (math.sqrt(2.0) * math.pi * f0)                          # tau^2 = 1/(2 pi^2 f0^2)
    assert abs(orc.ricker(tz, f0, t0, A)) < 1e-14
    tau = np.linspace(0, 3e-6, 50)
    np.testing.assert_allclose(orc.ricker(t0 + tau, f0, t0, A), orc.ricker(t0 - tau, f0, t0, A))


def test_gaussian_closed_form():
    a, h, n = 5e6, 0.02e-3, 201          # a h^2 = 2e-3 << 1: Riemann sum ~ 1
    g = orc.gaussian_load(n, n, h, 0.0, 0.0, a)
    assert g[n // 2, n // 2] == pytest.approx(a / math.pi)
    assert g.sum() * h * h == pytest.approx(1.0, rel=1e-6)
    # FWHM = 2 sqrt(ln 2 / a) (PF:91-93): value at half width is half the peak
    half = math.sqrt(math.log(2) / a)
    g1 = orc.gaussian_load(3, 1, half, 0.0, 0.0, a)
    assert g1[0, 0] == pytest.approx(0.5 * a / math.pi, rel=1e-12)


# ---------------------------------------------------------------- stencils
def _poly(nx, ny, h, f):
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    return f(I * h, J * h)


def test_second_differences_exact_on_polynomials():
    nx, ny, h = 9, 7, 0.3
    inner = (slice(1, -1), slice(1, -1))
    cases = {  # u -> (Dxx, Dyy, Dxy) in the interior
        "x2": (lambda x, y: x * x, (2, 0, 0)),
        "y2": (lambda x, y: y * y, (0, 2, 0)),
        "xy": (lambda x, y: x * y, (0, 0, 1)),
    }
    for name, (f, want) in cases.items():
        d = orc.second_differences(_poly(nx, ny, h, f), h)
        for k in range(3):
            np.testing.assert_allclose(d[k][inner], want[k], atol=1e-9, err_msg=f"{name} stencil {k}")


def test_operator_coefficient_positions():
    """Polynomial probes give each coefficient at a distinct place (PI:86-98)."""
    nx, ny, h = 9, 8, 0.5
    C11, C22, C12, C16, C26, C66 = [float(c) for c in (11, 22, 12, 16, 26, 66)]
    C = (C11, C22, C12, C16, C26, C66)
    z = np.zeros((nx, ny))
    inner = (slice(1, -1), slice(1, -1))
    x2 = _poly(nx, ny, h, lambda x, y: x * x)
    y2 = _poly(nx, ny, h, lambda x, y: y * y)
    xy = _poly(nx, ny, h, lambda x, y: x * y)
    probes = [
        ((x2, z), (2 * C11, 2 * C16)),
        ((y2, z), (2 * C66, 2 * C26)),
        ((xy, z), (2 * C16, C12 + C66)),
        ((z, x2), (2 * C16, 2 * C66)),
        ((z, y2), (2 * C26, 2 * C22)),
        ((z, xy), (C12 + C66, 2 * C26)),
    ]
    for (ux, uy), (wx, wy) in probes:
        Lx, Ly = orc.apply_A(ux, uy, C, h)
        np.testing.assert_allclose(Lx[inner], wx, rtol=1e-10)
        np.testing.assert_allclose(Ly[inner], wy, rtol=1e-10)


def _dense(op, nx, ny, C, h):
    n = nx * ny
    M = np.zeros((2 * n, 2 * n))
    for k in range(2 * n):
        e = np.zeros(2 * n); e[k] = 1.0
        a, b = op(e[:n].reshape(nx, ny), e[n:].reshape(nx, ny), C, h)
        M[:, k] = np.concatenate([a.ravel(), b.ravel()])
    return M


def test_constant_operator_symmetric_and_negative_definite():
    M = _dense(orc.apply_A, 6, 5, C_ANISO, 1e-3)
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-9 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).max() < 0


def test_per_cell_transpose_dot_product():
    rng = np.random.default_rng(0)
    nx, ny, h = 7, 6, 1e-3
    C = [rng.uniform(1, 2, (nx, ny)) * c for c in (130e9, 12e9, 6e9, 3e9, 1.5e9, 5e9)]
    u = rng.standard_normal((2, nx, ny)); v = rng.standard_normal((2, nx, ny))
    Au = orc.apply_A(u[0], u[1], C, h)
    ATv = orc.apply_AT(v[0], v[1], C, h)
    lhs = (Au[0] * v[0]).sum() + (Au[1] * v[1]).sum()
    rhs = (u[0] * ATv[0]).sum() + (u[1] * ATv[1]).sum()
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # and the dense transpose equals apply_AT column by column
    M = _dense(orc.apply_A, nx, ny, C, h)
    MT = _dense(orc.apply_AT, nx, ny, C, h)
    np.testing.assert_allclose(MT, M.T, atol=1e-9 * np.abs(M).max())


def test_dA_dC_is_the_derivative_of_A():
    rng = np.random.default_rng(1)
    nx, ny, h = 6, 7, 1e-3
    u = rng.standard_normal((2, nx, ny))
    parts = orc.dA_dC(u[0], u[1], h)
    for p, name in enumerate(orc.PARAM_NAMES):
        e = np.zeros(6); e[p] = 1.0
        Lx, Ly = orc.apply_A(u[0], u[1], e, h)          # A is linear in C
        np.testing.assert_allclose(parts[name][0], Lx, atol=1e-9 * np.abs(Lx).max() + 1e-30)
        np.testing.assert_allclose(parts[name][1], Ly, atol=1e-9 * np.abs(Ly).max() + 1e-30)


# ---------------------------------------------------------------- time stepping
def test_first_frame_is_dt2_f_over_rho():
    """Recording offset (PI:294-301): frame 0 is u^1 = dt^2 f^0 / rho."""
    nx, ny, h, nt = 8, 9, 1e-3, 5
    dt = 1e-8
    w = np.array([3.0, 1.0, 1.0, 1.0, 1.0])
    src = [(3, 4, 2.0, 1)]
    tr, _ = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, src, w, stride=2)
    assert tr.shape == (2, 3, nx, ny)
    want = np.zeros((nx, ny)); want[3, 4] = dt * dt * 3.0 * 2.0 / RHO
    np.testing.assert_allclose(tr[1, 0], want, rtol=1e-15)
    assert np.abs(tr[0, 0]).max() == 0.0


def test_standing_wave_closed_form():
    """With C12 = -C66 and C16 = C26 = 0 the x component decouples and a
    Dirichlet sine mode oscillates at the discrete frequency."""
    nx, ny, h, nt = 11, 9, 1e-3, 60
    C = (130e9, 12e9, -5e9, 0.0, 0.0, 5e9)
    dt = 0.5 * h / math.sqrt(130e9 / RHO)
    k, l = 2, 3
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    phi = np.sin(k * np.pi * (I + 1) / (nx + 1)) * np.sin(l * np.pi * (J + 1) / (ny + 1))
    src = [(i, j, phi[i, j], 0) for i in range(nx) for j in range(ny)]
    wav = np.zeros(nt); wav[0] = 1.0
    tr, _ = orc.forward(C, h, dt, RHO, nt, nx, ny, src, wav, stride=1)
    ax = 4 / h ** 2 * math.sin(k * math.pi / (2 * (nx + 1))) ** 2
    ay = 4 / h ** 2 * math.sin(l * math.pi / (2 * (ny + 1))) ** 2
    w2 = (C[0] * ax + C[5] * ay) / RHO
    Om = math.acos(1 - dt * dt * w2 / 2)
    u1 = dt * dt / RHO
    for n in range(1, nt + 1):
        want = u1 * math.sin(n * Om) / math.sin(Om) * phi
        np.testing.assert_allclose(tr[0, n - 1], want, atol=1e-10 * u1)
    assert np.abs(tr[1]).max() < 1e-20


def test_discrete_energy_conserved_after_source_stops():
    nx, ny, h, nt = 10, 9, 1e-3, 80
    dt = 0.5 * h / math.sqrt(130e9 / RHO)
    wav = np.zeros(nt); wav[:10] = np.hanning(10)
    _, st = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, [(4, 6, 1e6, 1), (2, 2, -5e5, 0)], wav,
                        keep_states=True)
    E = []
    for n in range(10, nt):
        (a, b), (c, d) = st[n + 1], st[n]
        Lx, Ly = orc.apply_A(c, d, C_ANISO, h)
        kin = RHO / (2 * dt * dt) * (((a - c) ** 2).sum() + ((b - d) ** 2).sum())
        pot = -0.5 * ((a * Lx).sum() + (b * Ly).sum())
        E.append(kin + pot)
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-9 * abs(E[0])


def test_linear_in_source():
    nx, ny, h, nt, dt = 7, 8, 1e-3, 20, 5e-8
    wav = orc.ricker(np.arange(nt) * dt, 1e6, 3e-7, 1.0)
    a, _ = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, [(3, 7, 1.0, 1)], wav)
    b, _ = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, [(3, 7, 2.5, 1)], wav)
    np.testing.assert_allclose(b, 2.5 * a, rtol=1e-12, atol=1e-30)


# ---------------------------------------------------------------- adjoint
def _setup_small(stride=2, receivers=True, seed=0):
    rng = np.random.default_rng(seed)
    nx, ny, h, nt = 8, 7, 1e-3, 24
    dt = 0.4 * h / math.sqrt(130e9 / RHO)
    wav = orc.ricker(np.arange(nt) * dt, 2e6, 1e-7, 1e9)
    src = [(4, 6, 1.0 / h ** 2, 1), (1, 2, 0.3 / h ** 2, 0)]
    rec = np.array([[0, 6], [3, 6], [7, 6], [5, 0], [3, 6]]) if receivers else None
    nf = orc.n_frames(nt, stride)
    shape = (2, nf, len(rec)) if receivers else (2, nf, nx, ny)
    G = rng.standard_normal(shape)
    return nx, ny, h, nt, dt, wav, src, rec, G


@pytest.mark.parametrize("stride,receivers", [(1, True), (3, True), (2, False)])
def test_const_gradient_matches_central_differences(stride, receivers):
    nx, ny, h, nt, dt, wav, src, rec, G = _setup_small(stride, receivers)
    C = C_ANISO.copy()
    def J(Cv):
        tr, _ = orc.forward(Cv, h, dt, RHO, nt, nx, ny, src, wav, stride, rec)
        return float((tr * G).sum())
    _, st = orc.forward(C, h, dt, RHO, nt, nx, ny, src, wav, stride, rec, keep_states=True)
    g = orc.const_gradient(orc.adjoint_gradient(C, h, dt, RHO, nt, nx, ny, st, G, stride, rec))
    for p in range(6):
        d = 1e-4 * 5e9
        e = np.zeros(6); e[p] = d
        fd = (J(C + e) - J(C - e)) / (2 * d)
        assert g[p] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(g).max())


def test_per_cell_and_damage_gradients_match_central_differences():
    nx, ny, h, nt, dt, wav, src, rec, G = _setup_small(2, True, seed=3)
    rng = np.random.default_rng(4)
    D = rng.uniform(0.5, 1.0, (nx, ny))
    C0 = C_ANISO
    fields = lambda Dv: [Dv * c for c in C0]
    def J(Cf):
        tr, _ = orc.forward(Cf, h, dt, RHO, nt, nx, ny, src, wav, 2, rec)
        return float((tr * G).sum())
    _, st = orc.forward(fields(D), h, dt, RHO, nt, nx, ny, src, wav, 2, rec, keep_states=True)
    gc = orc.adjoint_gradient(fields(D), h, dt, RHO, nt, nx, ny, st, G, 2, rec)
    for (p, i, j) in [(0, 3, 5), (2, 4, 4), (3, 0, 6), (5, 6, 1), (4, 2, 3), (1, 5, 5)]:
        Cf = fields(D)
        d = 1e-4 * C0[0]
        Cp = [c.copy() for c in Cf]; Cp[p][i, j] += d
        Cm = [c.copy() for c in Cf]; Cm[p][i, j] -= d
        fd = (J(Cp) - J(Cm)) / (2 * d)
        assert gc[p, i, j] == pytest.approx(fd, rel=1e-5, abs=1e-8 * np.abs(gc[p]).max())
    gC0, gD = orc.dfield_gradient(gc, C0, D)
    for (i, j) in [(3, 5), (0, 0), (7, 6)]:
        Dp = D.copy(); Dp[i, j] += 1e-5
        Dm = D.copy(); Dm[i, j] -= 1e-5
        fd = (J(fields(Dp)) - J(fields(Dm))) / 2e-5
        assert gD[i, j] == pytest.approx(fd, rel=1e-5, abs=1e-8 * np.abs(gD).max())
    for p in range(6):
        e = np.zeros(6); e[p] = 1e-4 * 5e9
        fd = (J([D * c for c in C0 + e]) - J([D * c for c in C0 - e])) / (2e-4 * 5e9)
        assert gC0[p] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(gC0).max())


def test_adjoint_matches_autograd_of_plain_path_fp64():
    """The hand-derived adjoint equals PyTorch autograd through the op-for-op
    restatement (PI:44-106, PI:240-307) in float64, for scalar and per-cell C."""
    torch.set_default_dtype(torch.float64)
    try:
        nx, ny, h, nt = 8, 8, 1e-3, 20
        dt = 0.4 * h / math.sqrt(130e9 / RHO)
        P = 3
        sig = pt.source_signal(nt, dt, 2e6, 1e-7, 1e9, dtype=torch.float64)
        spatial = torch.zeros(nx, ny); spatial[4, 7] = 1.0 / h ** 2; spatial[2, 3] = 0.2 / h ** 2
        src = [(4, 7, 1.0 / h ** 2, 1), (2, 3, 0.2 / h ** 2, 1)]
        rng = np.random.default_rng(5)
        nf = orc.n_frames(nt, P)
        G = rng.standard_normal((2, nf, nx, ny))
        rho = torch.tensor(RHO)
        # scalar C
        Ct = [torch.tensor(c, requires_grad=True) for c in C_ANISO]
        ux, uy = pt.forward(Ct, h, dt, nt, P, sig, spatial, rho)
        J = (ux[0] * torch.tensor(G[0])).sum() + (uy[0] * torch.tensor(G[1])).sum()
        J.backward()
        ga = np.array([c.grad.item() for c in Ct])
        tr, st = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, src, sig[0].numpy(), P, None, keep_states=True)
        np.testing.assert_allclose(torch.stack([ux[0], uy[0]]).detach().numpy(), tr, rtol=1e-12, atol=1e-25)
        go = orc.const_gradient(orc.adjoint_gradient(C_ANISO, h, dt, RHO, nt, nx, ny, st, G, P, None))
        np.testing.assert_allclose(ga, go, rtol=1e-9)
        # per-cell C (DI:100-176, force not transposed: centred semantics of PI)
        fields = [torch.tensor(rng.uniform(0.7, 1.0, (1, nx, ny)) * c, requires_grad=True) for c in C_ANISO]
        ux, uy = pt.forward(fields, h, dt, nt, P, sig, spatial, rho, per_cell=True)
        J = (ux[0] * torch.tensor(G[0])).sum() + (uy[0] * torch.tensor(G[1])).sum()
        J.backward()
        Cn = [f.detach()[0].numpy() for f in fields]
        _, st = orc.forward(Cn, h, dt, RHO, nt, nx, ny, src, sig[0].numpy(), P, None, keep_states=True)
        gc = orc.adjoint_gradient(Cn, h, dt, RHO, nt, nx, ny, st, G, P, None)
        for p in range(6):
            np.testing.assert_allclose(fields[p].grad[0].numpy(), gc[p], rtol=1e-8,
                                       atol=1e-10 * np.abs(gc[p]).max())
    finally:
        torch.set_default_dtype(torch.float32)


def test_plain_path_fp32_close_to_fp64_oracle():
    nx, ny, h, nt = 32, 32, 0.2e-3, 120
    dt = 0.5 * h / math.sqrt(130e9 / RHO)
    sig = pt.source_signal(nt, dt, 1e6, 1e-6, 1e10)
    spatial = torch.zeros(nx, ny); spatial[16, 31] = float(np.float32(1.0 / h ** 2))
    Ct = [torch.tensor(c, dtype=torch.float32) for c in C_ANISO]
    ux, uy = pt.forward(Ct, h, dt, nt, 5, sig, spatial, torch.tensor(RHO), hoist_weights=True)
    tr, _ = orc.forward(C_ANISO, h, dt, RHO, nt, nx, ny, [(16, 31, float(spatial[16, 31]), 1)],
                        sig[0].double().numpy(), 5, None)
    got = torch.stack([ux[0], uy[0]]).double().numpy()
    assert np.linalg.norm(got - tr) / np.linalg.norm(tr) < 1e-5


def test_total_variation_closed_form():
    f = np.array([[0.0, 1.0, 3.0], [2.0, 2.0, 0.0]])
    # |dx|: |2|,|1|,|3| ; |dy|: 1,2,0,2
    assert orc.total_variation(f) == pytest.approx(6 + 5)
    assert float(pt.total_variation_2d(torch.tensor(f)[None])) == pytest.approx(11)


def test_oracle_adjoint_matches_committed_fp64_finite_differences_single_512_crop():
    """At the single-512 physics on a 128^2 crop (2000 steps, 64 receivers), the
    FP64 discrete adjoint matches FP64 central differences of J = <traces, G>
    at step 1e-4 relative (tests/golden/fd_single512_crop.json, written by
    tests/golden/make_fd_single512_crop.py from the oracle forward)."""
    import importlib.util
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("make_fd", os.path.join(here, "make_fd_single512_crop.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = json.load(open(os.path.join(here, "fd_single512_crop.json")))
    s, C, G = mk.problem()
    assert np.array_equal(C, np.array(fx["C"])) and fx["step_rel"] == 1e-4
    g = mk.adjoint(s, C, G)
    fd = np.array(fx["fd"])
    assert np.abs(g - fd).max() / np.abs(fd).max() < 1e-6
    assert np.allclose(g, fx["adjoint"], rtol=1e-10, atol=0)


# ---------------------------------------------------------------- time reversal
def _reversed_states(C, h, dt, nt, nx, ny, src, wav, kept, K):
    """u^m for m = nt .. 0 rebuilt backward in increment form, the algebra of
    the reversed-u adjoint (rev_block_kernel): each segment starts from the
    kept state of its top level (u^T and v^T = u^T - u^{T-1}), then
    v^m = v^{m+1} - (dt^2/rho)(A u^m + f^m) and u^{m-1} = u^m - v^m."""
    k1 = dt * dt / RHO
    out = [None] * (nt + 1)
    for top in range(nt, 0, -K):
        bot = max(top - K, 0)
        ux, uy = (a.copy() for a in kept[top])
        vx, vy = ux - kept[top - 1][0], uy - kept[top - 1][1]
        out[top] = (ux.copy(), uy.copy())
        for m in range(top, bot, -1):
            if m < top:   # v^m from v^{m+1}; at the top, v^T is the checkpoint's
                fx, fy = orc._force(nx, ny, src, float(wav[m]))
                Lx, Ly = orc.apply_A(ux, uy, C, h)
                vx = vx - k1 * (Lx + fx)
                vy = vy - k1 * (Ly + fy)
            ux, uy = ux - vx, uy - vy
            out[m - 1] = (ux.copy(), uy.copy())
    return out


@pytest.mark.parametrize("K", [7, 64])
def test_time_reversed_states_and_damage_gradient(K):
    """The leapfrog with its force is time-symmetric: rebuilding u backward from
    sparse checkpoints recovers every stored state, and the damage-field
    gradient taken on the rebuilt states equals the one on the stored states
    (what the reverse_u plans rely on).  Per-cell stiffness D * C0, receivers
    and a source inside the grid, 60 steps in FP64."""
    rng = np.random.default_rng(5)
    nx, ny, h, nt = 12, 10, 1e-3, 60
    dt = 0.45 * h / math.sqrt(130e9 / RHO)
    wav = orc.ricker(np.arange(nt) * dt, 2e6, 1e-7, 1e9)
    src = [(5, 9, 1.0 / h ** 2, 1), (2, 3, 0.4 / h ** 2, 0)]
    rec = np.array([[0, 9], [6, 9], [11, 9], [4, 0]])
    D = rng.uniform(0.5, 1.0, (nx, ny))
    C = [D * c for c in C_ANISO]
    tr, kept = orc.forward(C, h, dt, RHO, nt, nx, ny, src, wav, 1, rec, keep_states=True)
    back = _reversed_states(C, h, dt, nt, nx, ny, src, wav, kept, K)
    scale = max(np.abs(kept[n][0]).max() for n in range(nt + 1))
    err = max(np.abs(back[n][q] - kept[n][q]).max() for n in range(nt + 1) for q in (0, 1))
    assert err < 1e-12 * scale
    # a dropped force or a sign slip in the reversal is far outside that bound
    bad = _reversed_states(C, h, dt, nt, nx, ny, [], wav, kept, K)
    assert max(np.abs(bad[n][0] - kept[n][0]).max() for n in range(nt + 1)) > 1e-6 * scale
    G = rng.standard_normal(tr.shape)
    g_kept = orc.adjoint_gradient(C, h, dt, RHO, nt, nx, ny, kept, G, 1, rec)
    g_back = orc.adjoint_gradient(C, h, dt, RHO, nt, nx, ny, back, G, 1, rec)
    gD_kept = orc.dfield_gradient(g_kept, C_ANISO, D)[1]
    gD_back = orc.dfield_gradient(g_back, C_ANISO, D)[1]
    assert np.linalg.norm(gD_back - gD_kept) < 1e-10 * np.linalg.norm(gD_kept)
