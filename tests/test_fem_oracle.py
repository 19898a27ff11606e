"""Pins of the FP64 FEM restatement (oracle/fem_fp64.py) by what the
mathematics fixes: mesh counts and mirror symmetry, total mass, rigid-body
modes, the patch test, the strain-energy identity for linear fields,
closed-form stencils of single stiffness entries, the centroid rule of the
damaged element stiffness, the first nonzero step, and exact energy
conservation of the central-difference recurrence.

PF = Parameter_inversion/FEM_generate_data.py, DF = Damage_inversion/FEM_generate_data.py.
"""
import numpy as np
import pytest

from oracle import fem_fp64 as F

C_REF = (130.74e9, 11.50e9, 10.49e9, 30.67e9, 3.75e9, 14.77e9)   # PF:840-845 (C11, C22, C12, C16, C26, C66)


def _grid(nx=9, ny=8, Lx=0.04, Ly=0.035):
    x, y, el, bnd, dx, dy = F.structured_mesh(nx, ny, Lx, Ly)
    return nx, ny, Lx, Ly, x, y, el, bnd, dx, dy


def _node_field(f):
    """(nx, ny) array -> node vector in the reference order j * nx + i."""
    return np.asarray(f).T.ravel()


def _interior(nx, ny):
    k = np.arange(nx * ny)
    i, j = k % nx, k // nx
    return k[(i > 0) & (i < nx - 1) & (j > 0) & (j < ny - 1)]


def test_mesh_counts_areas_and_pattern():
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid()
    assert el.shape == (2 * (nx - 1) * (ny - 1), 3)
    assert bnd.size == 2 * nx + 2 * ny - 4
    detJ, _, _, _ = F.element_geometry(x, y, el, nx)
    np.testing.assert_allclose(detJ, dx * dy, rtol=1e-12)          # every triangle is a positively oriented half cell
    assert abs(0.5 * detJ.sum() - Lx * Ly) < 1e-12 * Lx * Ly
    # each cell's diagonal: i + j even -> n0-n2 shared by its two triangles, odd -> n1-n3 (PF:166-173)
    for c in range(0, el.shape[0], 2):
        i, j = (c // 2) % (nx - 1), (c // 2) // (nx - 1)
        n0 = j * nx + i
        shared = set(el[c]) & set(el[c + 1])
        assert shared == ({n0, n0 + nx + 1} if (i + j) % 2 == 0 else {n0 + 1, n0 + nx})


@pytest.mark.parametrize("nx,symmetric", [(9, True), (8, False)])
def test_mesh_mirror_symmetry(nx, symmetric):
    """The checkerboard makes the mesh symmetric under x -> -x when nx is odd
    (the left-right symmetry PF:120-123 asks for), and only then."""
    _, _, _, _, x, y, el, _, _, _ = _grid(nx=nx)
    i, j = el % nx, el // nx
    tri = {frozenset(zip(a, b)) for a, b in zip(i, j)}
    mirrored = {frozenset(zip(nx - 1 - a, b)) for a, b in zip(i, j)}
    assert (tri == mirrored) is symmetric


def test_lumped_mass():
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid()
    rho = 1610.0
    M = F.lumped_mass(el, nx * ny, rho, dx, dy)
    assert abs(M.sum() - rho * Lx * Ly) < 1e-12 * rho * Lx * Ly
    m = rho * dx * dy / 6
    for k in _interior(nx, ny):
        i, j = k % nx, k // nx
        assert abs(M[k] - (8 if (i + j) % 2 == 0 else 4) * m) < 1e-12 * m


def test_stiffness_symmetric_and_rigid_modes():
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid()
    K = F.stiffness_matrix(x, y, el, nx, C_REF).toarray()
    scale = np.abs(K).max()
    np.testing.assert_allclose(K, K.T, atol=1e-12 * scale)
    X, Y = np.meshgrid(x, y, indexing="ij")
    for ux, uy in ((np.ones_like(X), 0 * X), (0 * X, np.ones_like(X)), (-Y, X)):   # translations, rotation
        u = np.concatenate([_node_field(ux), _node_field(uy)])
        assert np.abs(K @ u).max() < 1e-9 * scale * np.abs(u).max()


def test_patch_test_and_energy_identity():
    """A linear field has constant strain: interior nodal forces vanish and
    u^T K u = Lx Ly eps^T C eps (engineering shear), whatever C is."""
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid()
    rng = np.random.default_rng(3)
    X, Y = np.meshgrid(x, y, indexing="ij")
    inter = _interior(nx, ny)
    for trial in range(4):
        Cv = rng.uniform(-1, 1, 6) * 1e10 + np.array([5, 5, 0, 0, 0, 5]) * 1e10
        K = F.stiffness_matrix(x, y, el, nx, Cv)
        E = rng.normal(size=(2, 2))
        u = np.concatenate([_node_field(E[0, 0] * X + E[0, 1] * Y), _node_field(E[1, 0] * X + E[1, 1] * Y)])
        f = K @ u
        N = nx * ny
        assert np.abs(f[inter]).max() < 1e-9 * np.abs(f).max()
        assert np.abs(f[N + inter]).max() < 1e-9 * np.abs(f).max()
        eps = np.array([E[0, 0], E[1, 1], E[0, 1] + E[1, 0]])
        C11, C22, C12, C16, C26, C66 = Cv
        Cm = np.array([[C11, C12, C16], [C12, C22, C26], [C16, C26, C66]])
        np.testing.assert_allclose(u @ f, Lx * Ly * eps @ Cm @ eps, rtol=1e-10)


@pytest.mark.parametrize("which", ["C11", "C22", "C66"])
def test_single_entry_stencils(which):
    """One nonzero stiffness entry gives a 1D second difference on interior
    nodes: C11 -> ux along x with weight dy/dx, C22 -> uy along y with dx/dy,
    C66 -> ux along y (dx/dy) and uy along x (dy/dx)."""
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid()
    idx = {"C11": 0, "C22": 1, "C66": 5}[which]
    Cv = np.zeros(6)
    Cv[idx] = 2.5e10
    K = F.stiffness_matrix(x, y, el, nx, Cv)
    rng = np.random.default_rng(5)
    ux, uy = rng.normal(size=(nx, ny)), rng.normal(size=(nx, ny))
    f = K @ np.concatenate([_node_field(ux), _node_field(uy)])
    fx, fy = f[:nx * ny].reshape(ny, nx).T, f[nx * ny:].reshape(ny, nx).T
    d2x = lambda u: 2 * u[1:-1, 1:-1] - u[:-2, 1:-1] - u[2:, 1:-1]
    d2y = lambda u: 2 * u[1:-1, 1:-1] - u[1:-1, :-2] - u[1:-1, 2:]
    c = Cv[idx]
    if which == "C11":
        want_x, want_y = c * dy / dx * d2x(ux), 0 * d2x(uy)
    elif which == "C22":
        want_x, want_y = 0 * d2x(ux), c * dx / dy * d2y(uy)
    else:
        want_x = c * dx / dy * d2y(ux)
        want_y = c * dy / dx * d2x(uy)
        # the cross term C66 (d_y ux)(d_x uy) couples the components; probe each alone
        f1 = K @ np.concatenate([_node_field(ux), 0 * _node_field(uy)])
        f2 = K @ np.concatenate([0 * _node_field(ux), _node_field(uy)])
        fx = f1[:nx * ny].reshape(ny, nx).T
        fy = f2[nx * ny:].reshape(ny, nx).T
    np.testing.assert_allclose(fx[1:-1, 1:-1], want_x, atol=1e-9 * c)
    np.testing.assert_allclose(fy[1:-1, 1:-1], want_y, atol=1e-9 * c)


def test_centroid_node_rule():
    """The node nearest each centroid (DF:1005-1010) is, per cell (i, j):
    even [n0,n1,n2] -> (i+1, j), [n0,n2,n3] -> (i, j+1); odd [n0,n1,n3] ->
    (i, j), [n1,n2,n3] -> (i+1, j+1).  The HIP kernel samples the damage
    field at exactly these nodes."""
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid(nx=11, ny=10)
    ci, cj = F.element_centre_nodes(x, y, el, nx)
    cell = np.arange(el.shape[0]) // 2
    i, j = cell % (nx - 1), cell // (nx - 1)
    even = (i + j) % 2 == 0
    upper = np.arange(el.shape[0]) % 2 == 1
    di = np.where(even, np.where(upper, 0, 1), np.where(upper, 1, 0))
    dj = np.where(even, np.where(upper, 1, 0), np.where(upper, 1, 0))
    np.testing.assert_array_equal(ci, i + di)
    np.testing.assert_array_equal(cj, j + dj)


def test_damaged_energy_identity():
    """Per-element C: for a linear field u^T K u = A_tri eps^T (sum_e C_e) eps."""
    nx, ny, Lx, Ly, x, y, el, bnd, dx, dy = _grid(nx=13, ny=12)
    X, Y = np.meshgrid(x, y, indexing="ij")
    mask, info = F.damage_regions(X, Y, [{"type": "circle", "center": (0.004, 0.003), "size": 0.009},
                                         {"type": "square", "center": (-0.01, -0.008), "size": 0.004}])
    assert 0 < info["total_damaged_points"] < mask.size
    fields = F.damaged_stiffness(C_REF, mask)
    ci, cj = F.element_centre_nodes(x, y, el, nx)
    C_elem = np.stack([f[ci, cj] for f in fields], 1)
    K = F.stiffness_matrix(x, y, el, nx, C_elem)
    E = np.array([[0.3, -0.2], [0.5, 0.1]])
    u = np.concatenate([_node_field(E[0, 0] * X + E[0, 1] * Y), _node_field(E[1, 0] * X + E[1, 1] * Y)])
    eps = np.array([E[0, 0], E[1, 1], E[0, 1] + E[1, 0]])
    n_dmg = int(mask[ci, cj].sum())
    C11, C22, C12, C16, C26, C66 = C_REF
    Cm = np.array([[C11, C12, C16], [C12, C22, C26], [C16, C26, C66]])
    want = dx * dy / 2 * (el.shape[0] - n_dmg + 0.5 * n_dmg) * (eps @ Cm @ eps)
    np.testing.assert_allclose(u @ (K @ u), want, rtol=1e-10)


def test_damage_region_counts():
    """Lattice counts: a square of half-width 2 dx centred on a node holds
    5 x 5 nodes; a circle of radius 2.5 dx holds the 21 nodes with
    i^2 + j^2 <= 6 (DF:146-152)."""
    n = 21
    x = np.linspace(-0.1, 0.1, n)
    d = x[1] - x[0]
    X, Y = np.meshgrid(x, x, indexing="ij")
    m_sq, _ = F.damage_regions(X, Y, [{"type": "square", "center": (x[5], x[7]), "size": 2.0 * d + 1e-12}])
    m_ci, _ = F.damage_regions(X, Y, [{"type": "circle", "center": (x[10], x[12]), "size": 2.5 * d}])
    assert m_sq.sum() == 25 and m_ci.sum() == 21
    with pytest.raises(ValueError):
        F.damage_regions(X, Y, [{"type": "hexagon", "center": (0, 0), "size": d}])


def test_first_step_and_frames():
    """u^1 = 0, so u^2 = dt^2 w[1] g / rho on interior nodes (y only) and
    frame k = 0.7 u^{k stride} (PF:482-493, PF:525-526)."""
    nx, ny, Lx, Ly = 9, 8, 0.04, 0.035
    x = np.linspace(-Lx / 2, Lx / 2, nx)
    y = np.linspace(-Ly / 2, Ly / 2, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    g = F.gaussian_point_load(X, Y, 0.0, 0.0, 4e4)
    w = np.zeros(6)
    w[1] = 3.0e3
    dt, rho = 1e-7, 1610.0
    fx, fy = F.solve(nx, ny, Lx, Ly, dt, 6, C_REF, rho, g, w, stride=1)
    assert fx.shape == (6, nx, ny) and np.all(fx[:2] == 0) and np.all(fy[:2] == 0)
    np.testing.assert_allclose(fy[2][1:-1, 1:-1], 0.7 * dt * dt * w[1] * g[1:-1, 1:-1] / rho, rtol=1e-12)
    assert np.all(fy[2][0] == 0) and np.all(fy[2][:, -1] == 0) and np.all(fx[2] == 0)
    fx5, fy5 = F.solve(nx, ny, Lx, Ly, dt, 6, C_REF, rho, g, w, stride=2)
    assert fx5.shape == (3, nx, ny)
    np.testing.assert_allclose(fy5[1], fy[2], rtol=0, atol=0)
    np.testing.assert_allclose(fy5[2], fy[4], rtol=0, atol=0)


def test_energy_conservation_after_source():
    """With the load off after a few steps the scheme conserves
    1/2 v^T M v + 1/2 u^n K u^{n+1} to FP64 rounding; a wrong sign, a dropped
    term or a non-symmetric K breaks it."""
    nx, ny, Lx, Ly = 15, 13, 0.04, 0.035
    x, y, el, bnd, dx, dy = F.structured_mesh(nx, ny, Lx, Ly)
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1610.0
    vmax = F.wave_velocities(*C_REF, rho)[0]
    dt = 0.4 * dx / vmax
    g = F.gaussian_point_load(X, Y, 0.0, 0.0, 2e5)
    nt = 80
    w = np.zeros(nt)
    w[1:6] = [1e3, 3e3, -2e3, 1e3, 5e2]
    fx, fy = F.solve(nx, ny, Lx, Ly, dt, nt, C_REF, rho, g, w, stride=1, scale=1.0)
    K = F.stiffness_matrix(x, y, el, nx, C_REF)
    M = F.lumped_mass(el, nx * ny, rho, dx, dy)
    inter = _interior(nx, ny)
    idx = np.concatenate([inter, nx * ny + inter])
    state = lambda k: np.concatenate([_node_field(fx[k]), _node_field(fy[k])])
    E = [F.discrete_energy(K, M, state(k), state(k + 1), dt, idx) for k in (10, 30, 55, 78)]
    assert E[0] > 0
    np.testing.assert_allclose(E, E[0], rtol=1e-10)


def test_time_axis_cfl():
    dt, nt, t = F.time_axis(2001, 60e-6, 0.4 / 300, 0.4 / 300, 9011.5)   # PF main: CFL 0.20, unchanged
    assert nt == 2001 and abs(dt - 3e-8) < 1e-20 and t.size == 2001
    # CFL 4.1 -> nt = floor(T / dt_stable) + 1, dt = T / (nt - 1) >= dt_stable: the recomputed
    # number lands just above 0.5 (by at most a factor nt / (nt - 1)), as in PF:289-291
    dt, nt, t = F.time_axis(101, 60e-6, 0.4 / 300, 0.4 / 300, 9011.5)
    dt_stable = 0.5 * (0.4 / 300) / 9011.5
    assert nt == int(60e-6 / dt_stable) + 1 and abs(dt - 60e-6 / (nt - 1)) < 1e-22 and t.size == nt
    assert 0.5 <= 9011.5 * dt / (0.4 / 300) <= 0.5 * nt / (nt - 1)
