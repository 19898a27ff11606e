"""The Adam waveform-misfit loop on the HIP path recovers the constants
(plain-128: 128^2, 600 steps, 64 top-row receivers, 1 % RMS seeded noise,
+/-20 % seeded start, 50 iterations, lr 1e-2 on normalised constants,
C16 = C26 = 0 known).  The loop is the reference's (Parameter_inversion/
Inversion_adam.py:436-450, Adam at :403-408)."""
import numpy as np
import pytest
import torch

import ernn
from ernn import configs
from ernn.inversion import TraceInversion, add_noise_rms
from gpu_helpers import DEV

pytestmark = pytest.mark.gpu
ORTHO = (True, True, True, False, False, True)


def _observed(wl):
    eng = ernn.Engine(wl.survey, device=DEV)
    wav = torch.tensor(wl.survey.wavelets(), device=DEV)
    with torch.no_grad():
        obs = ernn.ernn_traces(eng, params=torch.tensor(wl.C_true[0], dtype=torch.float32, device=DEV),
                               wavelets=wav)
    return add_noise_rms(obs.view(2, -1), 0.01, torch.Generator().manual_seed(1)).reshape(-1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_plain_128_inversion_recovers_constants_in_50_iterations(seed):
    """The configuration's own budget: 50 Adam iterations at lr 1e-2 on the
    normalised constants (ERNN's default scale) bring C11, C22 (C33), C12
    (C13) and C66 (C55) within 2 % of the true values."""
    wl = configs.plain_128()
    obs = _observed(wl)
    Ct = wl.C_true[0]
    rng = np.random.default_rng(seed)
    Ci = Ct * (1 + rng.uniform(-0.2, 0.2, 6))
    Ci[3:5] = 0.0
    m = ernn.ERNN(wl.survey, Ci, device=DEV, trainable=ORTHO)
    inv = TraceInversion(m, obs, lr=1e-2)
    inv.run(50)
    H = np.array(inv.history["C"])[:, 0]
    nz = [0, 1, 2, 5]
    err0 = np.abs(Ci - Ct)[nz] / Ct[nz]
    err50 = np.abs(H[49] - Ct)[nz] / Ct[nz]
    loss = inv.history["loss"]
    print(f"seed {seed}: loss {loss[0]:.3e} -> {loss[49]:.3e}; rel err % of C11, C33, C13, C55 at start "
          f"{np.round(100 * err0, 2)}, after 50 iterations {np.round(100 * err50, 2)}")
    assert loss[49] < 0.1 * loss[0]
    assert np.all(err50 < 0.02)
    assert np.all(H[:, 3:5] == 0.0)
