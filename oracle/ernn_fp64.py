"""FP64 NumPy oracle for the ERNN forward recurrence and its discrete adjoint.

Test infrastructure only (see oracle/__init__.py): never imported by the
product.

Reference (all paths relative to the reference repository root):
  PI = Parameter_inversion/Inversion_adam.py
  DI = Damage_inversion/Inversion_adam.py

Conventions (PI:186-191, PI:255): a field is an (nx, ny) array, axis 0 is x
(index i), axis 1 is y (index j, contiguous).  Zero ghost cells surround the
grid (PI:50); boundary cells themselves evolve.  Stiffness constants use the
reference's Voigt names and order ``C11, C22, C12, C16, C26, C66``
(PI:86-98); the x-z names C11, C33, C13, C15, C35, C55 are the same six.

Discrete model (PI:54-106, PI:240-307):
  Dxx u = (u[i+1,j] - 2u[i,j] + u[i-1,j]) / h^2
  Dyy u = (u[i,j+1] - 2u[i,j] + u[i,j-1]) / h^2
  Dxy u = (u[i+1,j+1] - u[i+1,j-1] - u[i-1,j+1] + u[i-1,j-1]) / (4 h^2)
  Lx = C11 Dxx ux + C66 Dyy ux + (C12+C66) Dxy uy + 2 C16 Dxy ux
       + C16 Dxx uy + C26 Dyy uy
  Ly = C66 Dxx uy + C22 Dyy uy + (C12+C66) Dxy ux + C16 Dxx ux
       + C26 Dyy ux + 2 C26 Dxy uy
  u^{n+1} = 2 u^n - u^{n-1} + dt^2 (L u^n + f^n) / rho,  u^0 = u^-1 = 0
  frame k of the output is u^{kP+1} (recorded after step n = kP).
Per-cell stiffness (DI:151-163) multiplies the same stencils by the local
C at the output cell (non-divergence form); scalars broadcast.

Parity pinning of each function (tests/test_oracle_fp64.py):
  ricker            closed form: peak A at tau=0, zero crossings at
                    tau^2 = 1/(2 pi^2 f0^2), evenness in tau.
  gaussian_load     closed form: sum*h^2 -> 1 for a*h^2 << 1, value at the
                    centre = a/pi, FWHM = 2 sqrt(ln2/a).
  second_differences exact on polynomials (x^2, y^2, xy give 2, 2, 1 in the
                    interior), orientation i=x / j=y, symmetric matrices.
  apply_A           polynomial probes pin every coefficient position (e.g.
                    ux=x^2 gives Lx=2C11, Ly=2C16), isotropic Navier case,
                    symmetry of the dense matrix for constant C.
  apply_AT          dot-product test <A u, v> = <u, A^T v> for per-cell C.
  forward           u^1 = dt^2 f^0 / rho (recording offset), linearity in
                    the source, discrete energy conservation without source,
                    1-D plane-wave dispersion relation, time symmetry (states
                    rebuilt backward from checkpoints equal the stored ones).
  adjoint_gradient  central finite differences of forward+loss, and PyTorch
                    autograd of the plain_torch restatement in float64.
"""
from __future__ import annotations

import numpy as np

PARAM_NAMES = ("C11", "C22", "C12", "C16", "C26", "C66")


# --------------------------------------------------------------------------
# Source model (PI:28-41, PI:186-196, PI:210-217)
# --------------------------------------------------------------------------
def ricker(t, f0, t0, amplitude=1.0):
    """Ricker wavelet A (1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2), PI:28-35."""
    tau = np.asarray(t, dtype=np.float64) - t0
    q = (np.pi * f0 * tau) ** 2
    return amplitude * (1.0 - 2.0 * q) * np.exp(-q)


def gaussian_load(nx, ny, h, x0=0.0, y0=0.0, a=5e6):
    """(a/pi) exp(-a r^2) on the centred node grid of PI:186-195 (PI:38-41)."""
    x = np.linspace(-h * (nx - 1) / 2, h * (nx - 1) / 2, nx)
    y = np.linspace(-h * (ny - 1) / 2, h * (ny - 1) / 2, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    return (a / np.pi) * np.exp(-a * ((X - x0) ** 2 + (Y - y0) ** 2))


# --------------------------------------------------------------------------
# Stencils and the elastic operator (PI:44-106, DI:100-176)
# --------------------------------------------------------------------------
def _pad(u):
    return np.pad(u, 1)


def second_differences(u, h):
    """(Dxx u, Dyy u, Dxy u) with zero ghosts (PI:50, PI:56-69, PI:78-83)."""
    p = _pad(np.asarray(u, dtype=np.float64))
    ih2 = 1.0 / (h * h)
    c = p[1:-1, 1:-1]
    dxx = (p[2:, 1:-1] - 2.0 * c + p[:-2, 1:-1]) * ih2
    dyy = (p[1:-1, 2:] - 2.0 * c + p[1:-1, :-2]) * ih2
    dxy = (p[2:, 2:] - p[2:, :-2] - p[:-2, 2:] + p[:-2, :-2]) * (0.25 * ih2)
    return dxx, dyy, dxy


def _coeffs(C):
    return [np.asarray(c, dtype=np.float64) for c in C]


def apply_A(ux, uy, C, h):
    """(Lx, Ly) = A(C) u, PI:86-98 (scalars) / DI:151-163 (per-cell fields)."""
    C11, C22, C12, C16, C26, C66 = _coeffs(C)
    ux_xx, ux_yy, ux_xy = second_differences(ux, h)
    uy_xx, uy_yy, uy_xy = second_differences(uy, h)
    Lx = (C11 * ux_xx + C66 * ux_yy + (C12 + C66) * uy_xy + 2.0 * C16 * ux_xy
          + C16 * uy_xx + C26 * uy_yy)
    Ly = (C66 * uy_xx + C22 * uy_yy + (C12 + C66) * ux_xy + C16 * ux_xx
          + C26 * ux_yy + 2.0 * C26 * uy_xy)
    return Lx, Ly


def apply_AT(lx, ly, C, h):
    """Transpose of apply_A for per-cell C.

    A = sum_t diag(a_t) D_t (one term per stencil product of PI:86-98), and
    every D_t with zero ghosts is a symmetric matrix, so
    A^T l = sum_t D_t (a_t * l_row(t)).
    """
    C11, C22, C12, C16, C26, C66 = _coeffs(C)
    c1266 = C12 + C66
    # weighted fields that feed each stencil, grouped by (component, stencil)
    wx_xx = C11 * lx + C16 * ly
    wx_yy = C66 * lx + C26 * ly
    wx_xy = 2.0 * C16 * lx + c1266 * ly
    wy_xx = C16 * lx + C66 * ly
    wy_yy = C26 * lx + C22 * ly
    wy_xy = c1266 * lx + 2.0 * C26 * ly
    ATx = (second_differences(wx_xx, h)[0] + second_differences(wx_yy, h)[1]
           + second_differences(wx_xy, h)[2])
    ATy = (second_differences(wy_xx, h)[0] + second_differences(wy_yy, h)[1]
           + second_differences(wy_xy, h)[2])
    return ATx, ATy


def dA_dC(ux, uy, h):
    """Per-cell partials d(A u)/dC_p: dict name -> (x-row, y-row) fields."""
    ux_xx, ux_yy, ux_xy = second_differences(ux, h)
    uy_xx, uy_yy, uy_xy = second_differences(uy, h)
    z = np.zeros_like(ux_xx)
    return {
        "C11": (ux_xx, z),
        "C22": (z, uy_yy),
        "C12": (uy_xy, ux_xy),
        "C16": (2.0 * ux_xy + uy_xx, ux_xx),
        "C26": (uy_yy, ux_yy + 2.0 * uy_xy),
        "C66": (ux_yy + uy_xy, uy_xx + ux_xy),
    }


# --------------------------------------------------------------------------
# Forward recurrence (PI:240-307)
# --------------------------------------------------------------------------
def n_frames(nt, stride):
    """Number of recorded frames, ceil(nt / P) (PI:206, PI:299)."""
    return (nt + stride - 1) // stride


def _force(nx, ny, sources, s_n):
    fx = np.zeros((nx, ny))
    fy = np.zeros((nx, ny))
    for (i, j, w, comp) in sources:
        (fx if comp == 0 else fy)[i, j] += s_n * w
    return fx, fy


def _record(ux, uy, receivers):
    if receivers is None:
        return ux.copy(), uy.copy()
    r = np.asarray(receivers)
    return ux[r[:, 0], r[:, 1]].copy(), uy[r[:, 0], r[:, 1]].copy()


def forward(C, h, dt, rho, nt, nx, ny, sources, wavelet, stride=1,
            receivers=None, keep_states=False):
    """Run the leapfrog of PI:268-301.

    sources   list of (i, j, weight, comp) with comp 0 = x, 1 = y; the force
              at step n is wavelet[n] * weight on that component (PI:270-277).
    receivers None records whole fields (the reference's snapshot output,
              PI:294-305); otherwise an (n_rec, 2) array of (i, j).
    Returns traces (2, n_frames, ...) and, if keep_states, the list of
    states u^0 .. u^nt as (ux, uy) pairs.
    """
    ux1 = np.zeros((nx, ny)); uy1 = np.zeros((nx, ny))
    ux2 = np.zeros((nx, ny)); uy2 = np.zeros((nx, ny))
    frames_x, frames_y = [], []
    states = [(ux1.copy(), uy1.copy())] if keep_states else None
    for n in range(nt):
        fx, fy = _force(nx, ny, sources, float(wavelet[n]))
        Lx, Ly = apply_A(ux1, uy1, C, h)
        ax = (Lx + fx) / rho
        ay = (Ly + fy) / rho
        uxn = 2.0 * ux1 - ux2 + dt * dt * ax
        uyn = 2.0 * uy1 - uy2 + dt * dt * ay
        ux2, uy2, ux1, uy1 = ux1, uy1, uxn, uyn
        if keep_states:
            states.append((ux1.copy(), uy1.copy()))
        if n % stride == 0:
            rx, ry = _record(ux1, uy1, receivers)
            frames_x.append(rx)
            frames_y.append(ry)
    traces = np.stack([np.stack(frames_x), np.stack(frames_y)])
    return traces, states


# --------------------------------------------------------------------------
# Hand-derived discrete adjoint (replaces autograd at PI:447 / DI:511)
# --------------------------------------------------------------------------
def adjoint_gradient(C, h, dt, rho, nt, nx, ny, states, grad_traces,
                     stride=1, receivers=None):
    """Gradient of J w.r.t. per-cell stiffness, given dJ/d(traces).

    Adjoint recurrence (mu^{nt+1} = mu^{nt+2} = 0, m = nt .. 1):
        mu^m = r^m + (2I + (dt^2/rho) A^T) mu^{m+1} - mu^{m+2}
    with r^m the trace gradient injected at the receivers when m = kP+1.
    Gradient (u^0 = 0 removes the m = 0 term):
        dJ/dC_p(cell) = (dt^2/rho) sum_m mu^{m+1}(cell) . [dA/dC_p u^m](cell)
    Returns a (6, nx, ny) array in PARAM_NAMES order.  For scalar C the
    gradient of a constant is the sum over cells; for C_p = D * C0_p the
    chain rule gives dJ/dD = sum_p C0_p g_p and dJ/dC0_p = sum(D g_p).
    """
    G = np.asarray(grad_traces, dtype=np.float64)
    nfr = n_frames(nt, stride)
    k1 = dt * dt / rho
    mx1 = np.zeros((nx, ny)); my1 = np.zeros((nx, ny))   # mu^{m+1}
    mx2 = np.zeros((nx, ny)); my2 = np.zeros((nx, ny))   # mu^{m+2}
    grad = np.zeros((6, nx, ny))
    for m in range(nt, 0, -1):
        # gradient term pairing mu^{m+1} with u^m
        ux, uy = states[m]
        parts = dA_dC(ux, uy, h)
        for p, name in enumerate(PARAM_NAMES):
            gx, gy = parts[name]
            grad[p] += k1 * (mx1 * gx + my1 * gy)
        ATx, ATy = apply_AT(mx1, my1, C, h)
        mx = 2.0 * mx1 - mx2 + k1 * ATx
        my = 2.0 * my1 - my2 + k1 * ATy
        if (m - 1) % stride == 0 and (m - 1) // stride < nfr:
            k = (m - 1) // stride
            if receivers is None:
                mx = mx + G[0, k]
                my = my + G[1, k]
            else:
                r = np.asarray(receivers)
                np.add.at(mx, (r[:, 0], r[:, 1]), G[0, k])
                np.add.at(my, (r[:, 0], r[:, 1]), G[1, k])
        mx2, my2, mx1, my1 = mx1, my1, mx, my
    return grad


def const_gradient(grad_cells):
    """dJ/dC for six constants: sum of the per-cell gradient."""
    return grad_cells.reshape(6, -1).sum(axis=1)


def dfield_gradient(grad_cells, C0, D):
    """(dJ/dC0 (6,), dJ/dD (nx, ny)) for C_p = D * C0_p."""
    C0 = np.asarray(C0, dtype=np.float64)
    gD = np.tensordot(C0, grad_cells, axes=(0, 0))
    gC0 = (grad_cells * np.asarray(D)[None]).reshape(6, -1).sum(axis=1)
    return gC0, gD


# --------------------------------------------------------------------------
# Losses (PI:358-391, DI:179-195, DI:441-463)
# --------------------------------------------------------------------------
def mse(pred, obs):
    """mean((pred - obs)^2) and its gradient w.r.t. pred."""
    d = np.asarray(pred, dtype=np.float64) - np.asarray(obs, dtype=np.float64)
    return float(np.mean(d * d)), 2.0 * d / d.size


def total_variation(field):
    """Anisotropic TV of DI:179-195: sum |forward differences| on both axes."""
    f = np.asarray(field, dtype=np.float64)
    return float(np.abs(np.diff(f, axis=0)).sum() + np.abs(np.diff(f, axis=1)).sum())
