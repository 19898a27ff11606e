"""FP64 restatement of the reference's FEM observed-data generator.

Test infrastructure only: the checker of the HIP FEM path, imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product.

PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py

  mesh (checkerboard diagonals)   PF:117-189 (DF:263-340)
  Gaussian point load             PF:51-66, nodal array PF:298-301, PF:322
  Ricker wavelet                  PF:69-88, time factors PF:459-463
  wave speeds, CFL adjustment     PF:96-114, PF:280-296
  lumped mass                     PF:327-339
  element Jacobians, B matrices   PF:341-397
  element stiffness (damage)      DF:969-1026 (C at the node nearest the centroid)
  internal-force operator         PF:417-457 (DF:1048-1088)
  time loop, fixed boundary       PF:479-513
  output (x0.7, (nx, ny, t))      PF:517-526; file frames k = 0, s, 2s, ... PF:926-930
  damage regions, C reduction     DF:121-180, DF:183-212 (factor 0.5 at DF:877-880)

The element loop of PF:417-457 (three Gauss points, each adding
|detJ| w B^T C B u of every element into its nodes) is assembled once into a
sparse FP64 matrix K, so one step is one sparse product; the sum is the same,
taken in exact arithmetic up to FP64 rounding.  Node numbering is the
reference's, node = j * nx + i (i along x); fields are returned as (nx, ny)
arrays like PF:517-523.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

GAUSS_POINTS = np.array([[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]])  # PF:348-352
GAUSS_WEIGHTS = np.array([1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0])                                      # PF:353
DN_DXI = np.array([-1.0, 1.0, 0.0])    # PF:355-359 (P1 shape-function derivatives, constant)
DN_DETA = np.array([-1.0, 0.0, 1.0])


def gaussian_point_load(x, y, x0, y0, a):
    """(a / pi) exp(-a r^2) (PF:51-66)."""
    return (a / np.pi) * np.exp(-a * ((x - x0) ** 2 + (y - y0) ** 2))


def ricker_wavelet(t, f0, t0, amplitude=1.0):
    """A (1 - 2 (pi f0 tau)^2) exp(-(pi f0 tau)^2), tau = t - t0 (PF:69-88)."""
    s = (np.pi * f0 * (t - t0)) ** 2
    return amplitude * (1 - 2 * s) * np.exp(-s)


def wave_velocities(C11, C22, C12, C16, C26, C66, rho):
    """(vp, vs, anisotropy factor) as PF:96-114."""
    return math.sqrt(max(C11, C22) / rho), math.sqrt(C66 / rho), abs(C11 - C22) / (C11 + C22)


def time_axis(nt, t_total, dx, dy, vmax):
    """(dt, nt, t) with the CFL adjustment of PF:280-296: if vmax dt / min(dx, dy)
    exceeds 0.5, dt is cut to 0.5 min(dx, dy) / vmax and nt recomputed."""
    dt = t_total / (nt - 1)
    t = np.linspace(0, t_total, nt)
    if vmax * dt / min(dx, dy) > 0.5:
        dt_stable = 0.5 * min(dx, dy) / vmax
        nt = int(t_total / dt_stable) + 1
        dt = t_total / (nt - 1)
        t = np.linspace(0, t_total, nt)
    return dt, nt, t


def structured_mesh(nx, ny, Lx, Ly):
    """(x, y, elements, boundary_nodes, dx, dy) of PF:117-189: cell (i, j) has
    corners n0 = j nx + i, n1 = n0 + 1, n2 = n0 + nx + 1, n3 = n0 + nx; cells
    with i + j even are split along n0-n2 into [n0, n1, n2], [n0, n2, n3],
    odd cells along n1-n3 into [n0, n1, n3], [n1, n2, n3]; elements are
    ordered by j, then i, lower triangle first."""
    dx = Lx / (nx - 1)
    dy = Ly / (ny - 1)
    x = np.linspace(-Lx / 2, Lx / 2, nx)
    y = np.linspace(-Ly / 2, Ly / 2, ny)
    jj, ii = np.meshgrid(np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    i = ii.ravel()
    j = jj.ravel()
    n0 = j * nx + i
    n1, n2, n3 = n0 + 1, n0 + nx + 1, n0 + nx
    even = ((i + j) % 2 == 0)[:, None]
    elements = np.empty((2 * n0.size, 3), np.int64)
    elements[0::2] = np.where(even, np.stack([n0, n1, n2], 1), np.stack([n0, n1, n3], 1))
    elements[1::2] = np.where(even, np.stack([n0, n2, n3], 1), np.stack([n1, n2, n3], 1))
    k = np.arange(nx * ny)
    ki, kj = k % nx, k // nx
    boundary = k[(ki == 0) | (ki == nx - 1) | (kj == 0) | (kj == ny - 1)]
    return x, y, elements, boundary, dx, dy


def element_geometry(x, y, elements, nx):
    """(detJ, Jinv, dN_dx, dN_dy) per element (PF:363-397): J = [[x1 - x0,
    x2 - x0], [y1 - y0, y2 - y0]], dN/dx = Jinv[0,0] dN/dxi + Jinv[1,0] dN/deta,
    dN/dy = Jinv[0,1] dN/dxi + Jinv[1,1] dN/deta."""
    xc = x[elements % nx]
    yc = y[elements // nx]
    J = np.empty((elements.shape[0], 2, 2))
    J[:, 0, 0] = xc[:, 1] - xc[:, 0]
    J[:, 0, 1] = xc[:, 2] - xc[:, 0]
    J[:, 1, 0] = yc[:, 1] - yc[:, 0]
    J[:, 1, 1] = yc[:, 2] - yc[:, 0]
    detJ = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    dN_dx = Jinv[:, 0, 0, None] * DN_DXI + Jinv[:, 1, 0, None] * DN_DETA
    dN_dy = Jinv[:, 0, 1, None] * DN_DXI + Jinv[:, 1, 1, None] * DN_DETA
    return detJ, Jinv, dN_dx, dN_dy


def element_centre_nodes(x, y, elements, nx, chunk=4096):
    """(ci, cj): the grid node nearest each element's centroid, argmin over
    |x - mean(x_e)| and |y - mean(y_e)| (DF:1005-1010)."""
    cx = x[elements % nx].mean(axis=1)
    cy = y[elements // nx].mean(axis=1)
    ci = np.empty(elements.shape[0], np.int64)
    cj = np.empty(elements.shape[0], np.int64)
    for s in range(0, elements.shape[0], chunk):
        ci[s:s + chunk] = np.abs(x[None, :] - cx[s:s + chunk, None]).argmin(axis=1)
        cj[s:s + chunk] = np.abs(y[None, :] - cy[s:s + chunk, None]).argmin(axis=1)
    return ci, cj


def lumped_mass(elements, n_nodes, rho, dx, dy):
    """M[node] = (rho dx dy / 6) x (elements touching the node) (PF:327-339)."""
    M = np.zeros(n_nodes)
    m = rho * (dx * dy / 2.0) / 3.0
    for k in range(3):
        np.add.at(M, elements[:, k], m)
    return M


def stiffness_matrix(x, y, elements, nx, C_elem):
    """Sparse K (2N x 2N, dofs [ux nodes, uy nodes]) with f_int = K u summed
    over the three Gauss points of PF:417-457.  C_elem: (6,) constants or
    (E, 6) per element, order C11, C22, C12, C16, C26, C66."""
    N = x.size * y.size
    E = elements.shape[0]
    detJ, _, dN_dx, dN_dy = element_geometry(x, y, elements, nx)
    C_elem = np.broadcast_to(np.asarray(C_elem, np.float64), (E, 6))
    C11, C22, C12, C16, C26, C66 = (C_elem[:, q] for q in range(6))
    Cm = np.empty((E, 3, 3))
    Cm[:, 0] = np.stack([C11, C12, C16], 1)   # PF:402-406 / DF:1071-1073
    Cm[:, 1] = np.stack([C12, C22, C26], 1)
    Cm[:, 2] = np.stack([C16, C26, C66], 1)
    B = np.zeros((E, 3, 6))                    # strains (eps_xx, eps_yy, gamma_xy) from (ux_0..2, uy_0..2)
    B[:, 0, 0:3] = dN_dx
    B[:, 1, 3:6] = dN_dy
    B[:, 2, 0:3] = dN_dy
    B[:, 2, 3:6] = dN_dx
    BtCB = np.einsum("eki,ekl,elj->eij", B, Cm, B)
    Ke = np.zeros((E, 6, 6))
    for g in range(3):                         # identical integrand at each point: B is constant
        Ke += (np.abs(detJ) * GAUSS_WEIGHTS[g])[:, None, None] * BtCB
    dof = np.concatenate([elements, elements + N], axis=1)   # (E, 6)
    rows = np.repeat(dof, 6, axis=1).ravel()
    cols = np.tile(dof, (1, 6)).ravel()
    return sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(2 * N, 2 * N))


def damage_regions(X, Y, configs):
    """(mask, info) of DF:121-180: union of circles (distance <= size) and
    squares (|dx|, |dy| <= size)."""
    mask = np.zeros_like(X, dtype=bool)
    regions = []
    for k, c in enumerate(configs):
        cx, cy = c["center"]
        if c["type"] == "circle":
            r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2) <= c["size"]
        elif c["type"] == "square":
            r = (np.abs(X - cx) <= c["size"]) & (np.abs(Y - cy) <= c["size"])
        else:
            raise ValueError(f"Unknown damage type: {c['type']}")
        mask |= r
        regions.append({"id": k, "type": c["type"], "center": c["center"], "size": c["size"],
                        "reduction_factor": c.get("reduction_factor", 0.5), "n_points": int(r.sum())})
    info = {"regions": regions, "total_damaged_points": int(mask.sum()),
            "damage_percentage": 100.0 * mask.sum() / mask.size}
    return mask, info


def damaged_stiffness(C, mask, reduction_factor=0.5):
    """Six (nx, ny) fields: the base constants, times reduction_factor on the
    mask (DF:183-212)."""
    out = []
    for c in C:
        f = np.full(mask.shape, float(c))
        f[mask] *= reduction_factor
        out.append(f)
    return out


def solve(nx, ny, Lx, Ly, dt, nt, C, rho, load, wavelet, *, C_fields=None, stride=1, scale=0.7, load_x=None):
    """Central-difference FEM run of PF:479-526 in FP64.

    C: six constants (PF:402-406), or C_fields: six (nx, ny) fields sampled at
    each element's centroid node (DF:1003-1026).  load: (nx, ny) nodal
    Gaussian weights (PF:298-301), applied to the y component; wavelet: (nt,)
    time factors (PF:459-463); load_x: an optional x-component load of the
    same form (the reference loads y only).  Steps n = 1 .. nt-2 produce u^{n+1}; u^0 =
    u^1 = 0 and boundary nodes are held at zero (PF:495-497).  Returns
    (ux, uy) frames of shape (ceil(nt / stride), nx, ny) holding scale x
    u^{k stride} (PF:525-526, PF:926-930)."""
    x = np.linspace(-Lx / 2, Lx / 2, nx)
    y = np.linspace(-Ly / 2, Ly / 2, ny)
    _, _, elements, boundary, dx, dy = structured_mesh(nx, ny, Lx, Ly)
    N = nx * ny
    if C_fields is not None:
        ci, cj = element_centre_nodes(x, y, elements, nx)
        C_elem = np.stack([np.asarray(f, np.float64)[ci, cj] for f in C_fields], 1)
    else:
        C_elem = np.asarray(C, np.float64)
    K = stiffness_matrix(x, y, elements, nx, C_elem)
    M = lumped_mass(elements, N, rho, dx, dy)
    g = np.asarray(load, np.float64).T.ravel()          # node order j * nx + i (PF:322)
    gx = None if load_x is None else np.asarray(load_x, np.float64).T.ravel()
    tf = np.asarray(wavelet, np.float64)
    n_frames = (nt + stride - 1) // stride
    fx_out = np.zeros((n_frames, nx, ny))
    fy_out = np.zeros((n_frames, nx, ny))
    u_prev = np.zeros(2 * N)
    u_cur = np.zeros(2 * N)
    for n in range(1, nt - 1):
        fext_y = (tf[n] * g / rho) * M                  # PF:482-484
        f = K @ u_cur
        u_next = 2 * u_cur - u_prev
        fext_x = 0.0 if gx is None else (tf[n] * gx / rho) * M
        u_next[:N] += dt * dt * (fext_x - f[:N]) / M    # PF:489-493
        u_next[N:] += dt * dt * (fext_y - f[N:]) / M
        u_next[boundary] = 0.0                          # PF:495-497
        u_next[N + boundary] = 0.0
        u_prev, u_cur = u_cur, u_next
        if (n + 1) % stride == 0:
            k = (n + 1) // stride
            fx_out[k] = scale * u_cur[:N].reshape(ny, nx).T
            fy_out[k] = scale * u_cur[N:].reshape(ny, nx).T
    return fx_out, fy_out


def discrete_energy(K, M, u_prev, u_cur, dt, interior):
    """E^{n+1/2} = 1/2 v^T M v + 1/2 u^{n}^T K u^{n+1}, v = (u^{n+1} - u^n) / dt,
    over interior dofs: conserved exactly by the force-free central-difference
    scheme with symmetric K (a property of the recurrence, used to pin it)."""
    Mi = np.concatenate([M, M])
    v = (u_cur - u_prev) / dt
    return 0.5 * np.sum(Mi[interior] * v[interior] ** 2) + 0.5 * u_prev[interior] @ (K @ u_cur)[interior]
