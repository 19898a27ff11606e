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
        data.save_reference_data(str(tmp_path / "x.txt"), ux, uy, 0, 0, 0)
    with pytest.raises(FileNotFoundError):
        data.load_reference_data(str(tmp_path / "missing_ds5.npz"))


@pytest.mark.skipif(importlib.util.find_spec("h5py") is not None, reason="h5py is installed")
def test_truncated_hdf5_is_an_error_without_h5py(tmp_path):
    """Without h5py the package's own reader takes HDF5 files; a file that is
    only a signature is refused with its error."""
    from ernn import hdf5
    p = tmp_path / "anisotropic_ricker_data_ds5_FEM_fixed_bc.mat"
    p.write_bytes(b"\x89HDF\r\n\x1a\n")
    with pytest.raises((hdf5.HDF5Error, IndexError, ValueError)):
        data.load_reference_data(str(p))
    q = tmp_path / "not_hdf5_ds5.mat"
    q.write_bytes(b"plain text, no signature" * 40)
    with pytest.raises(hdf5.HDF5Error, match="signature"):
        data.load_reference_data(str(q))


def test_reference_layout_round_trip_through_hdf5(tmp_path):
    """save_reference_data to an HDF5 name writes the reference's layout (ux,
    uy in gzip frames, region subgroups) and the loaders read it back."""
    ux, uy = _fields(nt=9, nx=6, ny=5, seed=4)
    x, y, t = np.linspace(0, 1, 6, dtype=np.float32), np.linspace(0, 2, 5, dtype=np.float32), np.arange(9, dtype=np.float32)
    mask = (np.arange(30).reshape(6, 5) % 4 == 0)
    p = data.save_reference_data(str(tmp_path / "anisotropic_ricker_damage_data_ds5_FEM_fixed_bc.mat"), ux, uy, x, y, t,
                                 source_info={"source_f0": 100e3, "component": "y"}, damage_mask=mask,
                                 damage_info={"total_damaged_points": 8, "damage_percentage": 26.7,
                                              "region_0/type": "ellipse", "region_0/center": np.array([0.1, 0.2])})
    gx, gy, meta = data.load_fem_damage_data(p)
    assert np.array_equal(gx[0].numpy(), ux) and np.array_equal(gy[0].numpy(), uy)
    assert meta["primary_stride"] == 5 and np.array_equal(meta["t_primary"], t)
    assert np.array_equal(meta["damage_mask"], mask.astype(np.uint8))
    assert meta["damage_info"]["total_damaged_points"] == 8
    assert meta["source_info"]["source_f0"] == 100e3 and meta["source_info"]["component"] == b"y"
    from ernn import hdf5
    with hdf5.File(p) as f:
        assert f["damage_info/region_0/type"][()] == b"ellipse"
        assert np.array_equal(f["damage_info"]["region_0"]["center"][:], [0.1, 0.2])
