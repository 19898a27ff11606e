"""Observed-data files in the reference's layout (ernn/data.py), CPU only:
round trips through .npz, the stride parsed from the file name as the
reference does (PI:578-580), seeded noise at a target SNR (PI:561-576), and
the clear error for HDF5 without h5py."""
import importlib.util

import numpy as np
import pytest
import torch

from ernn import data


def _fields(nt=6, nx=5, ny=7, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nt, nx, ny)).astype(np.float32), rng.standard_normal((nt, nx, ny)).astype(np.float32)


def test_reference_layout_round_trip(tmp_path):
    ux, uy = _fields()
    x, y, t = np.linspace(0, 1, 5, dtype=np.float32), np.linspace(0, 2, 7, dtype=np.float32), np.arange(6, dtype=np.float32)
    p = data.save_reference_data(str(tmp_path / "anisotropic_ricker_data_ds7_FEM_fixed_bc.npz"), ux, uy, x, y, t,
                                 source_info={"f0": 1e6, "position": np.array([0.0, 1.0]), "comp": "y"})
    gx, gy, meta = data.load_reference_data(p)
    assert gx.shape == (1, 6, 5, 7) and gx.dtype == torch.float32
    assert np.array_equal(gx[0].numpy(), ux) and np.array_equal(gy[0].numpy(), uy)
    assert meta["primary_stride"] == 7 and meta["noise_added"] is False
    assert np.array_equal(meta["t_primary"], t) and np.array_equal(meta["x_original"], x)
    assert meta["source_info"]["f0"] == 1e6 and str(meta["source_info"]["comp"]) == "y"
    assert np.array_equal(meta["source_info"]["position"], [0.0, 1.0])


def test_damage_layout_round_trip_and_default_stride(tmp_path):
    ux, uy = _fields(seed=1)
    mask = (np.arange(35).reshape(5, 7) % 3 == 0)
    p = data.save_reference_data(str(tmp_path / "damage.npz"), ux, uy, np.zeros(5), np.zeros(7), np.zeros(6),
                                 damage_mask=mask, damage_info={"total_damaged_points": 12, "damage_percentage": 34.3})
    gx, gy, meta = data.load_fem_damage_data(p)
    assert meta["primary_stride"] == 5                       # no "_dsN" in the name: the reference's default
    assert meta["damage_mask"].dtype == np.uint8 and np.array_equal(meta["damage_mask"], mask.astype(np.uint8))
    assert meta["damage_info"]["total_damaged_points"] == 12
    assert meta["damage_info"]["damage_percentage"] == pytest.approx(34.3)
    p2 = data.save_reference_data(str(tmp_path / "plain_ds3.npz"), ux, uy, np.zeros(5), np.zeros(7), np.zeros(6))
    assert data.load_fem_damage_data(p2)[2]["damage_mask"] is None


def test_noise_is_seeded_at_the_target_snr(tmp_path):
    ux, uy = _fields(nt=40, nx=30, ny=30, seed=2)
    p = data.save_reference_data(str(tmp_path / "d_ds5.npz"), ux, uy, np.zeros(30), np.zeros(30), np.zeros(40))
    a = data.load_reference_data(p, add_noise=True, snr_db=20.0, generator=torch.Generator().manual_seed(3))
    b = data.load_reference_data(p, add_noise=True, snr_db=20.0, generator=torch.Generator().manual_seed(3))
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    noise = a[2]["noise_ux"]
    snr = 10 * np.log10(np.mean(ux.astype(np.float64) ** 2) / np.mean(noise.astype(np.float64) ** 2))
    assert snr == pytest.approx(20.0, abs=0.1)
    assert np.allclose(a[0][0].numpy(), ux + noise[0], atol=1e-6)


def test_bad_inputs(tmp_path):
    ux, uy = _fields()
    with pytest.raises(ValueError):
        data.save_reference_data(str(tmp_path / "x.npz"), ux, uy[:, :4], 0, 0, 0)
    with pytest.raises(ValueError):
        data.save_reference_data(str(tmp_path / "x.h5"), ux, uy, 0, 0, 0)
    with pytest.raises(FileNotFoundError):
        data.load_reference_data(str(tmp_path / "missing_ds5.npz"))


@pytest.mark.skipif(importlib.util.find_spec("h5py") is not None, reason="h5py is installed")
def test_hdf5_without_h5py_names_the_conversion(tmp_path):
    p = tmp_path / "anisotropic_ricker_data_ds5_FEM_fixed_bc.mat"
    p.write_bytes(b"\x89HDF\r\n\x1a\n")
    with pytest.raises(ImportError, match="h5py"):
        data.load_reference_data(str(p))
