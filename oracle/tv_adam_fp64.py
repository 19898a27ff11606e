"""FP64 restatement of the damage inversion's TV term and Adam step.

Test infrastructure only: the checker of the HIP TV + Adam update, imported by
tests/ and never by the product.

DI = Damage_inversion/Inversion_adam.py
  fields   C = clamp(exp(theta), lo, hi)                       DI:304-313
  TV       sum |C[i+1,j] - C[i,j]| + sum |C[i,j+1] - C[i,j]|   DI:179-195
  Adam     torch.optim.Adam (lr, betas, eps, weight_decay)     DI:478-483
"""
from __future__ import annotations

import numpy as np


def fields(theta, lo, hi, log_params=True):
    """(C, chain): C = clamp(exp(theta)) and dC/dtheta (zero outside [lo, hi])."""
    x = np.exp(theta) if log_params else np.asarray(theta, np.float64)
    inside = (x >= lo) & (x <= hi)
    return np.clip(x, lo, hi), np.where(inside, x if log_params else 1.0, 0.0)


def tv(C):
    """TV of one (nx, ny) field (DI:189-194)."""
    return float(np.abs(np.diff(C, axis=0)).sum() + np.abs(np.diff(C, axis=1)).sum())


def tv_subgradient(C):
    """dTV/dC with sign(0) = 0: +sign of the difference to the previous cell,
    -sign of the difference to the next, along both axes."""
    g = np.zeros_like(C, dtype=np.float64)
    sx = np.sign(np.diff(C, axis=0))
    sy = np.sign(np.diff(C, axis=1))
    g[1:, :] += sx
    g[:-1, :] -= sx
    g[:, 1:] += sy
    g[:, :-1] -= sy
    return g


def adam_step(theta, g, m, v, step, lr, beta1, beta2, eps, weight_decay=0.0):
    """One Adam update (Kingma & Ba, with torch's placement of eps after the
    bias-corrected square root): returns (theta, m, v)."""
    g = g + weight_decay * theta
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    denom = np.sqrt(v) / np.sqrt(1 - beta2 ** step) + eps
    return theta - (lr / (1 - beta1 ** step)) * m / denom, m, v
