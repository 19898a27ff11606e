"""Pins of the FP64 TV + Adam restatement (oracle/tv_adam_fp64.py): TV of
closed-form fields, the subgradient against central differences away from
ties, the clamp / exp chain, and Adam's first and second steps in closed form.

DI = Damage_inversion/Inversion_adam.py.
"""
import numpy as np

from oracle import tv_adam_fp64 as T


def test_tv_closed_forms():
    nx, ny = 7, 5
    assert T.tv(np.full((nx, ny), 3.0)) == 0.0
    ramp = np.add.outer(2.0 * np.arange(nx), 0.5 * np.arange(ny))   # C = 2 i + 0.5 j
    assert abs(T.tv(ramp) - ((nx - 1) * ny * 2.0 + nx * (ny - 1) * 0.5)) < 1e-12
    cb = (-1.0) ** np.add.outer(np.arange(nx), np.arange(ny))        # checkerboard +-1
    assert T.tv(cb) == 2.0 * ((nx - 1) * ny + nx * (ny - 1))


def test_subgradient_matches_central_differences():
    rng = np.random.default_rng(0)
    C = rng.normal(size=(9, 8))            # no ties: TV is differentiable here
    g = T.tv_subgradient(C)
    h = 1e-7
    fd = np.zeros_like(C)
    for idx in np.ndindex(C.shape):
        Cp, Cm = C.copy(), C.copy()
        Cp[idx] += h
        Cm[idx] -= h
        fd[idx] = (T.tv(Cp) - T.tv(Cm)) / (2 * h)
    np.testing.assert_allclose(g, fd, atol=1e-6)
    # ties give sign(0) = 0: a uniform field has zero subgradient, an isolated bump +-1 per neighbour
    bump = np.zeros((5, 5))
    bump[2, 2] = 1.0
    sb = T.tv_subgradient(bump)
    assert sb[2, 2] == 4.0 and sb[1, 2] == sb[3, 2] == sb[2, 1] == sb[2, 3] == -1.0 and sb[0, 0] == 0.0


def test_fields_clamp_and_chain():
    th = np.log(np.array([1e9, 60e9, 100e9, 300e9]))
    C, chain = T.fields(th, 50e9, 250e9)
    np.testing.assert_allclose(C, [50e9, 60e9, 100e9, 250e9], rtol=1e-12)
    np.testing.assert_allclose(chain, [0.0, 60e9, 100e9, 0.0], rtol=1e-12)
    C, chain = T.fields(np.array([-1.0, 0.5, 3.0]), 0.0, 1.0, log_params=False)
    np.testing.assert_array_equal(C, [0.0, 0.5, 1.0])
    np.testing.assert_array_equal(chain, [0.0, 1.0, 0.0])


def test_adam_first_steps_closed_form():
    """Step 1: m = (1-b1) g, v = (1-b2) g^2, so the move is lr g / (|g| + eps)
    -> lr sign(g); step 2 with g2 follows from the moment recurrences."""
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    th0 = np.array([1.0, -2.0, 0.5])
    g1 = np.array([3.0, -1e-3, 0.0])
    th1, m1, v1 = T.adam_step(th0, g1, np.zeros(3), np.zeros(3), 1, lr, b1, b2, eps)
    np.testing.assert_allclose(th1, th0 - lr * g1 / (np.abs(g1) + eps), rtol=1e-12)
    g2 = np.array([1.0, 1.0, -2.0])
    th2, m2, v2 = T.adam_step(th1, g2, m1, v1, 2, lr, b1, b2, eps)
    m_hat = (b1 * (1 - b1) * g1 + (1 - b1) * g2) / (1 - b1 ** 2)
    v_hat = (b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2) / (1 - b2 ** 2)
    np.testing.assert_allclose(th2, th1 - lr * m_hat / (np.sqrt(v_hat) + eps), rtol=1e-12)
    # weight decay adds wd theta to the gradient
    th1w, _, _ = T.adam_step(th0, 0 * g1, np.zeros(3), np.zeros(3), 1, lr, b1, b2, eps, weight_decay=0.1)
    np.testing.assert_allclose(th1w, th0 - lr * np.sign(th0) * (0.1 * np.abs(th0)) / (0.1 * np.abs(th0) + eps),
                               rtol=1e-12)
