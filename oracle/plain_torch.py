"""Op-for-op PyTorch restatement of the reference ERNN path ("plain path").

Test infrastructure only (see oracle/__init__.py): never imported by the
product.  It restates, with the same tensor ops in the same order:

  PI:44-51    conv2d_fixed (zero pad + 3x3 cross-correlation)
  PI:54-106   the constant-C elastic operator on transposed views
  DI:100-176  the per-cell-C operator (non-divergence form), including the
              reference's transposed force (DI:169-170) behind a flag
  PI:240-307  the leapfrog time loop with snapshot history
  DI:331-393  the same loop for per-cell C
  PI:358-391  the subsampled MSE loss; DI:441-463 the weighted MSE + TV loss
  DI:179-195  total variation

Autograd through the whole time loop (PI:447) gives the reference gradient.
Running it in float64 turns it into an independent check of the FP64 NumPy
oracle's hand-derived adjoint.

``hoist_weights=True`` builds the three stencil weight tensors once instead
of per step (PI:59-69 rebuilds them every step with blocking host-to-device
copies); the arithmetic is identical.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def ricker_wavelet(t, f0, t0, amplitude=1.0):
    """PI:28-35 (NumPy FP64)."""
    tau = t - t0
    q = (np.pi * f0 * tau) ** 2
    return amplitude * (1 - 2 * q) * np.exp(-q)


def source_signal(nt, dt, f0, t0, amplitude, dtype=torch.float32):
    """PI:210-217: FP64 Ricker samples stored as a (1, nt) tensor."""
    t = np.arange(nt) * dt
    return torch.tensor(ricker_wavelet(t, f0, t0, amplitude), dtype=dtype).unsqueeze(0)


def gaussian_point_load(nx, ny, h, x0=0.0, y0=0.0, a=5e6, dtype=torch.float32):
    """PI:38-41 evaluated on the grid of PI:186-195."""
    Lx = h * (nx - 1)
    Ly = h * (ny - 1)
    x = torch.linspace(-Lx / 2, Lx / 2, nx, dtype=dtype)
    y = torch.linspace(-Ly / 2, Ly / 2, ny, dtype=dtype)
    X, Y = torch.meshgrid(x, y, indexing="ij")
    r2 = (X - x0) ** 2 + (Y - y0) ** 2
    return (a / np.pi) * torch.exp(-a * r2)


def conv2d_fixed(field, kernel):
    """PI:44-51."""
    return F.conv2d(F.pad(field, (1, 1, 1, 1), mode="constant", value=0.0), kernel, padding=0)


def stencil_weights(h, device, dtype):
    """PI:56-69: h^-2 * {d2dx2, d2dy2, d2dxdy} as (1,1,3,3) tensors."""
    s = h ** (-2)
    d2dx2 = s * torch.tensor([[[[0.0, 0.0, 0.0], [1.0, -2.0, 1.0], [0.0, 0.0, 0.0]]]],
                             device=device, dtype=dtype)
    d2dy2 = s * torch.tensor([[[[0.0, 1.0, 0.0], [0.0, -2.0, 0.0], [0.0, 1.0, 0.0]]]],
                             device=device, dtype=dtype)
    d2dxdy = s * torch.tensor([[[[0.25, 0.0, -0.25], [0.0, 0.0, 0.0], [-0.25, 0.0, 0.25]]]],
                              device=device, dtype=dtype)
    return d2dx2, d2dy2, d2dxdy


def elastic_operator(ux, uy, C, h, force_x, force_y, rho, weights=None,
                     per_cell=False, transpose_force=False):
    """PI:54-106 (scalar C) / DI:100-176 (per-cell C fields of shape (B,nx,ny))."""
    C11, C22, C12, C16, C26, C66 = C
    if weights is None:
        weights = stencil_weights(h, ux.device, ux.dtype)
    d2dx2, d2dy2, d2dxdy = weights
    ux = ux.transpose(1, 2).unsqueeze(1)
    uy = uy.transpose(1, 2).unsqueeze(1)
    d2ux_dx2 = conv2d_fixed(ux, d2dx2).squeeze(1)
    d2ux_dy2 = conv2d_fixed(ux, d2dy2).squeeze(1)
    d2uy_dx2 = conv2d_fixed(uy, d2dx2).squeeze(1)
    d2uy_dy2 = conv2d_fixed(uy, d2dy2).squeeze(1)
    d2ux_dxdy = conv2d_fixed(ux, d2dxdy).squeeze(1)
    d2uy_dxdy = conv2d_fixed(uy, d2dxdy).squeeze(1)
    if per_cell:  # DI:143-148
        C11, C22, C12, C16, C26, C66 = (c.transpose(1, 2) for c in (C11, C22, C12, C16, C26, C66))
    Lux = (C11 * d2ux_dx2 + C66 * d2ux_dy2 + (C12 + C66) * d2uy_dxdy
           + 2 * C16 * d2ux_dxdy + C16 * d2uy_dx2 + C26 * d2uy_dy2)
    Luy = (C66 * d2uy_dx2 + C22 * d2uy_dy2 + (C12 + C66) * d2ux_dxdy
           + C16 * d2ux_dx2 + C26 * d2ux_dy2 + 2 * C26 * d2uy_dxdy)
    Lux = Lux.transpose(1, 2)
    Luy = Luy.transpose(1, 2)
    if transpose_force:  # DI:169-170
        force_x = force_x.transpose(1, 2)
        force_y = force_y.transpose(1, 2)
    return (Lux + force_x) / rho, (Luy + force_y) / rho


def forward(C, h, dt, nt, stride, signal, spatial, rho, component="y",
            save_all_steps=False, per_cell=False, transpose_force=False,
            hoist_weights=False):
    """PI:240-307 / DI:331-393.

    C        six 0-dim tensors (PI) or six (1,nx,ny) fields (DI), reference order
             C11, C22, C12, C16, C26, C66
    signal   (1, nt) source time function (PI:217)
    spatial  (nx, ny) spatial load (PI:195), e.g. the Gaussian or a one-hot
    Returns (ux_fields, uy_fields), each (1, n_frames, nx, ny).
    """
    device = signal.device
    dtype = signal.dtype
    nx, ny = spatial.shape
    shape = (1, nx, ny)
    ux1 = torch.zeros(shape, device=device, dtype=dtype)
    uy1 = torch.zeros(shape, device=device, dtype=dtype)
    ux2 = torch.zeros(shape, device=device, dtype=dtype)
    uy2 = torch.zeros(shape, device=device, dtype=dtype)
    weights = stencil_weights(h, device, dtype) if hoist_weights else None
    hx, hy = [], []
    g = spatial.unsqueeze(0)
    for i in range(nt):
        amp = signal[0, i]
        if component == "x":
            fx = amp * g
            fy = torch.zeros_like(ux1)
        else:
            fx = torch.zeros_like(ux1)
            fy = amp * g
        Lux1, Luy1 = elastic_operator(ux1, uy1, C, h, fx, fy, rho, weights=weights,
                                      per_cell=per_cell, transpose_force=transpose_force)
        ux_new = 2 * ux1 - ux2 + dt ** 2 * Lux1
        uy_new = 2 * uy1 - uy2 + dt ** 2 * Luy1
        ux2, uy2 = ux1, uy1
        ux1, uy1 = ux_new, uy_new
        if save_all_steps or i % stride == 0:
            hx.append(ux1)
            hy.append(uy1)
    return torch.stack(hx, dim=1), torch.stack(hy, dim=1)


def material_params_log(log_C, bounds):
    """PI:223-232: clamp(exp(log C), bounds)."""
    return tuple(torch.clamp(torch.exp(lc), *b) for lc, b in zip(log_C, bounds))


def loss_subsampled(pred_ux, pred_uy, obs_ux, obs_uy, t_stride, s_stride):
    """PI:358-391 without the shape-mismatch branch."""
    sl = (slice(None), slice(None, None, t_stride), slice(None, None, s_stride),
          slice(None, None, s_stride))
    return (torch.mean((pred_ux[sl] - obs_ux[sl]) ** 2)
            + torch.mean((pred_uy[sl] - obs_uy[sl]) ** 2))


def total_variation_2d(field):
    """DI:179-195."""
    dx = field[:, 1:, :] - field[:, :-1, :]
    dy = field[:, :, 1:] - field[:, :, :-1]
    return torch.sum(torch.abs(dx)) + torch.sum(torch.abs(dy))


def loss_damage(pred_ux, pred_uy, obs_ux, obs_uy, t_stride, fields=None, lambda_tv=0.0):
    """DI:441-463: 5*MSE_x + MSE_y + lambda * sum_p TV(C_p)."""
    px = pred_ux[:, ::t_stride]
    py = pred_uy[:, ::t_stride]
    loss = 5 * torch.mean((px - obs_ux[:, ::t_stride]) ** 2) + torch.mean((py - obs_uy[:, ::t_stride]) ** 2)
    if lambda_tv > 0 and fields is not None:
        loss = loss + lambda_tv * sum(total_variation_2d(f) for f in fields)
    return loss
