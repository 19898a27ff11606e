"""Adam waveform-misfit inversion: the reference's inverter classes and the
receiver-trace inversion of the benchmark surveys.

CFRPParameterInverterRicker         mirrors Parameter_inversion/Inversion_adam.py:314-509
SpatiallyVaryingParameterInverter   mirrors Damage_inversion/Inversion_adam.py:396-605
add_noise_snr                       mirrors Parameter_inversion/Inversion_adam.py:109-129, seeded
TraceInversion                      Adam on normalised constants / damage field with an MSE
                                    on receiver traces (the benchmark workloads)

The loops are the reference's (zero_grad -> forward -> loss -> backward ->
Adam.step -> .item() history -> early stopping); only the simulator's forward
and backward run on the HIP library.
"""
from __future__ import annotations

import os
from typing import Dict, Optional

import numpy as np
import torch
import torch.optim as optim

from .optim import FusedTVAdam
from .simulator import ERNN, total_variation_2d


def add_noise_snr(signal: torch.Tensor, snr_db: float, generator: Optional[torch.Generator] = None,
                  verbose: bool = False):
    """Additive Gaussian noise at a target SNR (PI:109-129).  The reference
    draws unseeded noise (PI:116); here a generator makes it reproducible."""
    signal_power = torch.mean(signal ** 2)
    noise_power = signal_power / (10 ** (snr_db / 10))
    noise = torch.randn(signal.shape, generator=generator, dtype=signal.dtype).to(signal.device)
    noise = noise * torch.sqrt(noise_power)
    if verbose:
        actual = 10 * torch.log10(signal_power / torch.mean(noise ** 2))
        print(f"  Target SNR: {snr_db:.2f} dB, actual {actual.item():.2f} dB")
    return signal + noise, noise


def add_noise_rms(traces: torch.Tensor, rel: float, generator: Optional[torch.Generator] = None):
    """Gaussian noise with standard deviation rel * RMS(traces)."""
    noise = torch.randn(traces.shape, generator=generator, dtype=traces.dtype).to(traces.device)
    return traces + rel * traces.pow(2).mean().sqrt() * noise


class CFRPParameterInverterRicker:
    """PI:314-509 over a CFRPWaveSimulatorRicker."""

    def __init__(self, simulator, target_ux_primary, target_uy_primary,
                 secondary_time_stride=1, secondary_space_stride=1, verbose=True):
        self.simulator = simulator
        self.target_ux_primary = target_ux_primary
        self.target_uy_primary = target_uy_primary
        self.secondary_time_stride = secondary_time_stride
        self.secondary_space_stride = secondary_space_stride
        sl = self._slice()
        self.target_ux_secondary = target_ux_primary[sl]
        self.target_uy_secondary = target_uy_primary[sl]
        self.history = {'loss': [], 'C11': [], 'C22': [], 'C12': [], 'C16': [], 'C26': [], 'C66': []}
        self.verbose = verbose

    def _slice(self):
        t, s = self.secondary_time_stride, self.secondary_space_stride
        return (slice(None), slice(None, None, t), slice(None, None, s), slice(None, None, s))

    def compute_loss(self, pred_ux_primary, pred_uy_primary):
        """PI:358-391, including the trim-to-common-shape branch (PI:371-384)."""
        sl = self._slice()
        px, py = pred_ux_primary[sl], pred_uy_primary[sl]
        tx, ty = self.target_ux_secondary, self.target_uy_secondary
        if px.shape != tx.shape:
            m = [min(a, b) for a, b in zip(px.shape[1:], tx.shape[1:])]
            cut = (slice(None), slice(0, m[0]), slice(0, m[1]), slice(0, m[2]))
            px, py, tx, ty = px[cut], py[cut], tx[cut], ty[cut]
        return torch.mean((px - tx) ** 2) + torch.mean((py - ty) ** 2)

    def invert(self, max_epochs=500, lr=1e-3, true_params=None, patience=50, min_epochs=100):
        """PI:393-483: Adam (betas 0.9/0.999, eps 1e-8) with patience early stopping."""
        optimizer = optim.Adam(self.simulator.parameters(), lr=lr, betas=(0.9, 0.999), eps=1e-8)
        best_loss, counter = float('inf'), 0
        for epoch in range(max_epochs):
            optimizer.zero_grad()
            ux, uy = self.simulator(save_all_steps=False)
            loss = self.compute_loss(ux, uy)
            loss.backward()
            optimizer.step()
            loss_val = loss.item()
            C11, C22, C12, C16, C26, C66, _ = self.simulator.get_current_params_values()
            self.history['loss'].append(loss_val)
            for k, v in zip(('C11', 'C22', 'C12', 'C16', 'C26', 'C66'), (C11, C22, C12, C16, C26, C66)):
                self.history[k].append(v)
            if loss_val < best_loss:
                best_loss, counter = loss_val, 0
            else:
                counter += 1
            if self.verbose:
                print(f"Epoch {epoch + 1:4d}: Loss={loss_val:.6e} | C11={C11:.3f}, C22={C22:.3f}, "
                      f"C12={C12:.3f} GPa | C16={C16:.3f}, C26={C26:.3f}, C66={C66:.3f} GPa")
            if counter >= patience and epoch > min_epochs:
                break
        return self.get_results(true_params)

    def get_results(self, true_params=None):
        """PI:485-509 (relative error 0 for a true value of 0, as the reference)."""
        C11, C22, C12, C16, C26, C66, _ = self.simulator.get_current_params_values()
        inv = {'C11': C11, 'C22': C22, 'C12': C12, 'C16': C16, 'C26': C26, 'C66': C66}
        res = {'inverted': inv, 'known': {'rho': self.simulator.rho_known}, 'history': self.history}
        if true_params:
            res['true'] = true_params
            res['error'] = {k: (abs(inv[k] - true_params[k]) / true_params[k] if true_params[k] != 0 else 0)
                            for k in inv}
        return res


class SpatiallyVaryingParameterInverter:
    """DI:396-605 over a SpatiallyVaryingAnisotropicSimulator."""

    def __init__(self, simulator, target_ux_primary, target_uy_primary, secondary_time_stride=1,
                 lambda_tv=0.0, verbose=True):
        self.simulator = simulator
        self.target_ux_primary = target_ux_primary
        self.target_uy_primary = target_uy_primary
        self.secondary_time_stride = secondary_time_stride
        self.lambda_tv = lambda_tv
        self.target_ux_secondary = target_ux_primary[:, ::secondary_time_stride, :, :]
        self.target_uy_secondary = target_uy_primary[:, ::secondary_time_stride, :, :]
        self.history = {'loss': [], 'loss_data': [], 'loss_tv': [], 'C11_mean': [], 'C22_mean': [],
                        'C12_mean': [], 'C16_mean': [], 'C26_mean': [], 'C66_mean': []}
        self.verbose = verbose

    def data_loss(self, pred_ux_primary, pred_uy_primary):
        """DI:445-451: 5 MSE_x + MSE_y on the secondary time samples."""
        s = self.secondary_time_stride
        px = pred_ux_primary[:, ::s, :, :]
        py = pred_uy_primary[:, ::s, :, :]
        return 5 * torch.mean((px - self.target_ux_secondary) ** 2) + \
            torch.mean((py - self.target_uy_secondary) ** 2)

    def compute_loss(self, pred_ux_primary, pred_uy_primary):
        """DI:441-463: 5 MSE_x + MSE_y + lambda sum_p TV(C_p)."""
        loss_data = self.data_loss(pred_ux_primary, pred_uy_primary)
        loss_tv = 0.0
        if self.lambda_tv > 0:
            C = self.simulator.get_material_params()[:6]
            loss_tv = self.lambda_tv * sum(total_variation_2d(c) for c in C)
        return loss_data + loss_tv, loss_data, loss_tv

    def fused_optimizer(self, lr=1e-3, betas=(0.9, 0.999), weight_decay=0.0):
        """FusedTVAdam over the simulator's six fields with its bounds and
        parameterisation (DI:245-270, DI:304-322) and this inverter's lambda_tv."""
        sim = self.simulator
        names = ('C11', 'C22', 'C12', 'C16', 'C26', 'C66')
        params = [getattr(sim, ('log_' if sim.use_log_params else '') + n) for n in names]
        return FusedTVAdam(params, lr=lr, betas=betas, weight_decay=weight_decay, lambda_tv=self.lambda_tv,
                           bounds=[getattr(sim, n + '_bounds') for n in names], log_params=sim.use_log_params)

    def invert(self, max_epochs=1000, lr=1e-3, betas=(0.9, 0.999), weight_decay=0.0,
               checkpoint_interval=50, output_dir="./checkpoints_fem_damage_adam_fixed_bc",
               patience=100, min_epochs=50, fused_update=False):
        """DI:465-560.  fused_update: the TV term's gradient and the Adam step
        run in one HIP call (FusedTVAdam) after the data loss's backward,
        instead of autograd through the TV graph and torch.optim.Adam."""
        if fused_update:
            optimizer = self.fused_optimizer(lr, betas, weight_decay)
        else:
            optimizer = optim.Adam(self.simulator.parameters(), lr=lr, betas=betas, weight_decay=weight_decay)
        best, counter = float('inf'), 0
        for epoch in range(max_epochs):
            optimizer.zero_grad()
            ux, uy = self.simulator(save_all_steps=False)
            if fused_update:
                data = self.data_loss(ux, uy)
                data.backward()
                optimizer.step()
                tv = self.lambda_tv * float(optimizer.last_tv.sum()) if self.lambda_tv > 0 else 0.0
                total = data.detach() + tv
            else:
                total, data, tv = self.compute_loss(ux, uy)
                total.backward()
                optimizer.step()
            cur = total.item()
            self.history['loss'].append(cur)
            self.history['loss_data'].append(data.item())
            self.history['loss_tv'].append(tv if isinstance(tv, float) else tv.item())
            means = self.simulator.get_mean_params()
            for k, v in zip(('C11_mean', 'C22_mean', 'C12_mean', 'C16_mean', 'C26_mean', 'C66_mean'), means[:6]):
                self.history[k].append(v)
            if cur < best:
                best, counter = cur, 0
            else:
                counter += 1
            if self.verbose:
                print(f"Epoch {epoch + 1:4d}: Loss={cur:.6e} (Data={self.history['loss_data'][-1]:.6e}, "
                      f"TV={self.history['loss_tv'][-1]:.6e}) | Mean: C11={means[0]:.2f}, "
                      f"C22={means[1]:.2f}, C66={means[5]:.2f} GPa")
            if checkpoint_interval and (epoch + 1) % checkpoint_interval == 0:
                self.save_checkpoint(epoch + 1, output_dir)
            if counter >= patience and epoch > min_epochs:
                break
            if epoch > 0 and abs(self.history['loss'][-1] - self.history['loss'][-2]) < 1e-40:
                break
        return self.get_results()

    def save_checkpoint(self, epoch, output_dir="./checkpoints_fem_damage_adam_fixed_bc"):
        """DI:562-581: parameter fields as compressed npz."""
        os.makedirs(output_dir, exist_ok=True)
        C = self.simulator.get_material_params()[:6]
        names = ('C11', 'C22', 'C12', 'C16', 'C26', 'C66')
        ck = {'epoch': epoch, **{n: c.detach().cpu().numpy() for n, c in zip(names, C)}}
        path = os.path.join(output_dir, f'checkpoint_epoch_{epoch:04d}.npz')
        np.savez_compressed(path, **ck)
        return path

    def get_results(self):
        """DI:583-605."""
        C = self.simulator.get_material_params()[:6]
        names = ('C11', 'C22', 'C12', 'C16', 'C26', 'C66')
        return {'inverted_fields': {n: c.detach().cpu().numpy() for n, c in zip(names, C)},
                'mean_values': {n: c.mean().item() / 1e9 for n, c in zip(names, C)},
                'history': self.history}


class TraceInversion:
    """Adam on an ERNN (normalised constants and/or damage field) against
    observed receiver traces: loss = mean((traces - observed)^2) [+ lambda TV(D)]."""

    def __init__(self, model: ERNN, observed: torch.Tensor, lr: float = 1e-2, lambda_tv: float = 0.0,
                 lr_damage: Optional[float] = None):
        self.model = model
        self.observed = observed
        self.lambda_tv = lambda_tv
        groups = [{"params": [model.theta], "lr": lr}]
        if hasattr(model, "damage") and model.damage.requires_grad:
            groups.append({"params": [model.damage], "lr": lr_damage or lr})
        self.opt = optim.Adam(groups, betas=(0.9, 0.999), eps=1e-8)
        self.history: Dict[str, list] = {"loss": [], "C": []}

    def loss(self, traces):
        l = torch.mean((traces - self.observed) ** 2)
        if self.lambda_tv > 0 and hasattr(self.model, "damage"):
            nx, ny = self.model.survey.set_grid(0)
            l = l + self.lambda_tv * total_variation_2d(self.model.damage.view(1, nx, ny))
        return l

    def step(self) -> float:
        self.opt.zero_grad()
        l = self.loss(self.model())
        l.backward()
        self.opt.step()
        with torch.no_grad():
            if hasattr(self.model, "damage"):
                self.model.damage.clamp_(0.0, 1.0)
        self.history["loss"].append(l.item())
        self.history["C"].append(self.model.stiffness().detach().cpu().numpy().copy())
        return self.history["loss"][-1]

    def run(self, iters: int, verbose: bool = False):
        for k in range(iters):
            l = self.step()
            if verbose:
                print(f"iter {k + 1:4d} loss {l:.6e} C(GPa) {np.round(self.history['C'][-1] / 1e9, 3)}")
        return self.history
