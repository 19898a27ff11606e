"""The drop-in: the reference's parameter-inversion main() on this package.

Parameter_inversion/Inversion_adam.py:669-891 reads the FEM observed data
(PI:714-745, noise at 30 dB), derives the grid and time axis from it
(PI:748-766), builds CFRPWaveSimulatorRicker with its initial guess
(PI:810-831), wraps it in CFRPParameterInverterRicker with secondary strides
2 and 8 (PI:841-847) and runs invert(lr=0.04) (PI:851).  This test does the
same with this package's FEM generator, data loader, simulator and inverter,
and runs the same loop on the reference's own method (the op-for-op plain
path, oracle/plain_torch.py, with autograd and torch.optim.Adam, PI:403-450).
The loss and the six constants must follow the same trajectory epoch by
epoch within 1e-4 relative.
"""
import numpy as np
import pytest
import torch

import ernn
from ernn import data, fem
from oracle import plain_torch as pt
from gpu_helpers import DEV

pytestmark = pytest.mark.gpu

PF_C = dict(C11=130.74e9, C22=11.50e9, C12=10.49e9, C16=30.67e9, C26=3.75e9, C66=14.77e9)   # PF:840-845
INIT = dict(C11_init=90e9, C22_init=20e9, C12_init=30e9, C16_init=20e9, C26_init=10e9, C66_init=10e9)   # PI:822-827
EPOCHS = 8
LR = 0.04


def test_parameter_inversion_main_matches_plain_path(tmp_path):
    # observed data: the FEM generator's file (PF:823-945) at the reference's size
    # the reference's own file name and format (PF:905-945): HDF5, ux / uy in gzip frames
    path = fem.generate_reference_data(str(tmp_path / "anisotropic_ricker_data_ds5_FEM_fixed_bc.mat"), stride=5,
                                       nx=301, ny=301, Lx=0.4, Ly=0.4, nt=2001, t_total=60e-6,
                                       source_amplitude=1e10, source_f0=100e3, source_t0=10e-6,
                                       gaussian_param=5e6, rho=1610.0, device=DEV, **PF_C)
    ux_p, uy_p, meta = data.load_reference_data(path, add_noise=True, snr_db=30.0, device=DEV,
                                                generator=torch.Generator().manual_seed(0))
    x, y, t = meta["x_original"], meta["y_original"], meta["t_primary"]
    si, P = meta["source_info"], meta["primary_stride"]
    nx, ny = len(x), len(y)
    h = (x[-1] - x[0]) / (nx - 1)
    nt = (len(t) - 1) * P + 1
    dt = t[-1] / (nt - 1)
    assert (nx, ny, nt, P) == (301, 301, 2001, 5)
    kw = dict(domain_shape=(nx, ny), h=h, dt=dt, nt=nt, primary_time_stride=P, source_component='y',
              source_f0=float(si["source_f0"]), source_t0=float(si["source_t0"]),
              source_amplitude=float(si["source_amplitude"]), rho_known=float(si["rho"]),
              gaussian_param=float(si["gaussian_param"]), source_position=(0.0, 0.0))
    true = {k: float(si[k]) / 1e9 for k in ("C11", "C22", "C12", "C16", "C26", "C66")}

    # this package: the reference's main() from the simulator on
    sim = ernn.CFRPWaveSimulatorRicker(**kw, **INIT).to(DEV)
    inv = ernn.CFRPParameterInverterRicker(sim, ux_p, uy_p, secondary_time_stride=2, secondary_space_stride=8,
                                           verbose=False)
    res = inv.invert(max_epochs=EPOCHS, lr=LR, true_params=true)
    hist = res["history"]
    assert len(hist["loss"]) == EPOCHS

    # the reference's method: plain path + autograd + torch.optim.Adam, same loop
    logs = [torch.nn.Parameter(torch.log(torch.tensor(v, dtype=torch.float32)).to(DEV)) for v in INIT.values()]
    opt = torch.optim.Adam(logs, lr=LR, betas=(0.9, 0.999), eps=1e-8)
    bounds = [(1e9, 10000e9)] * 6
    sig, g = sim.source_signal, sim.gaussian_dist
    rho = torch.tensor(kw["rho_known"], device=DEV)
    names = ("C11", "C22", "C12", "C16", "C26", "C66")
    ref = {"loss": [], **{n: [] for n in names}}
    for epoch in range(EPOCHS):
        opt.zero_grad()
        C = pt.material_params_log(logs, bounds)
        px, py = pt.forward(C, h, dt, nt, P, sig, g, rho, hoist_weights=True)
        loss = pt.loss_subsampled(px, py, ux_p, uy_p, 2, 8)
        loss.backward()
        opt.step()
        ref["loss"].append(loss.item())
        with torch.no_grad():
            for n, c in zip(names, pt.material_params_log(logs, bounds)):
                ref[n].append(c.item() / 1e9)
        del px, py, loss
        torch.cuda.empty_cache()

    err = {k: float(np.max(np.abs(np.array(hist[k]) - np.array(ref[k])) / np.abs(np.array(ref[k]))))
           for k in ref}
    print("PI main(): loss " + " ".join(f"{v:.4e}" for v in hist["loss"]))
    print("PI main(): max relative difference to the plain path per history " +
          ", ".join(f"{k} {v:.1e}" for k, v in err.items()))
    print("PI main(): after %d epochs (GPa) " % EPOCHS + ", ".join(f"{k} {res['inverted'][k]:.3f}" for k in names))
    assert hist["loss"][-1] < hist["loss"][0]
    assert max(err.values()) < 1e-4
