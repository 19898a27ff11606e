"""Observed-data files: the reference's reference-data layout, without h5py.

load_reference_data    mirrors Parameter_inversion/Inversion_adam.py:516-591
load_fem_damage_data   mirrors Damage_inversion/Inversion_adam.py:608-674
save_reference_data    writes the same layout (the reference's writers,
                       Parameter_inversion/FEM_generate_data.py:900-945 and
                       Damage_inversion/FEM_generate_data.py:1619-1684, write
                       it as HDF5)

The reference keeps its data in HDF5 files: datasets ux, uy (nt_ds, nx, ny)
float32 in chunks of one frame with gzip 4, x (nx,), y (ny,), t (nt_ds,),
optional damage_mask (nx, ny) uint8, groups source_info and damage_info of
scalars (damage_info with region_<i> subgroups), and the primary time stride
in the file name ("_ds5").  HDF5 files (.mat, .h5, .hdf5) are read through
h5py where it is installed and through this package's own reader
(ernn.hdf5) where it is not, and written in that layout by ernn.hdf5.write.
The same layout also goes to NumPy .npz archives (group members as
"source_info/<key>" entries, loaded with allow_pickle=False).
"""
from __future__ import annotations

import os
import re
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from .inversion import add_noise_snr

_GROUPS = ("source_info", "damage_info")


def primary_stride(path: str, default: int = 5) -> int:
    """The primary time stride from a file name ("..._ds5...") as the
    reference parses it (PI:578-580, DI:660-662); `default` when absent."""
    m = re.search(r"_ds(\d+)", path)
    return int(m.group(1)) if m else default


def _read(path: str) -> Dict[str, object]:
    """Datasets and group scalars of a reference-layout file as a flat dict."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"data file not found: {path}")
    if path.endswith(".npz"):
        out: Dict[str, object] = {}
        with np.load(path, allow_pickle=False) as z:
            for k in z.files:
                v = z[k]
                out[k] = v[()] if v.shape == () else v
        return out
    try:
        import h5py
        opener = lambda p: h5py.File(p, "r")
        is_group = lambda o: isinstance(o, h5py.Group)
    except ImportError:   # the reference reads with h5py; without it, this package's reader
        from . import hdf5
        opener = hdf5.File
        is_group = lambda o: isinstance(o, hdf5.Group)
    out = {}
    with opener(path) as f:
        for k in ("ux", "uy", "x", "y", "t", "damage_mask"):
            if k in f:
                out[k] = f[k][:]
        for g in _GROUPS:   # the group's datasets, as DI:631-645 reads them (region_<i> subgroups aside)
            if g in f:
                grp = f[g]
                for k in grp.keys():
                    obj = grp[k]
                    if not is_group(obj):
                        out[f"{g}/{k}"] = obj[()]
    return out


def _group(raw: Dict[str, object], name: str) -> Dict[str, object]:
    p = name + "/"
    return {k[len(p):]: v for k, v in raw.items() if k.startswith(p)}


def _fields(raw: Dict[str, object], device) -> Tuple[torch.Tensor, torch.Tensor]:
    ux = np.asarray(raw["ux"])
    uy = np.asarray(raw["uy"])
    if ux.ndim != 3 or ux.shape != uy.shape:
        raise ValueError(f"ux, uy must be (nt, nx, ny) arrays of one shape, got {ux.shape} and {uy.shape}")
    to = lambda a: torch.tensor(a, dtype=torch.float32).unsqueeze(0).to(device)   # (1, nt, nx, ny), PI:557-558
    return to(ux), to(uy)


def load_reference_data(path: str, add_noise: bool = False, snr_db: float = 20.0, device=None,
                        generator: Optional[torch.Generator] = None):
    """(ux, uy, metadata) with ux, uy of shape (1, nt_ds, nx, ny) float32 on
    `device` (PI:516-591).  With add_noise, Gaussian noise at snr_db per
    component (add_noise_snr, seeded by `generator`), the noise kept in the
    metadata like PI:561-576."""
    raw = _read(path)
    ux, uy = _fields(raw, device)
    meta: Dict[str, object] = {"noise_added": False}
    if add_noise:
        ux, nx_ = add_noise_snr(ux, snr_db, generator)
        uy, ny_ = add_noise_snr(uy, snr_db, generator)
        meta = {"noise_added": True, "snr_db": snr_db, "noise_ux": nx_.cpu().numpy(), "noise_uy": ny_.cpu().numpy()}
    meta.update({"x_original": np.asarray(raw["x"]), "y_original": np.asarray(raw["y"]),
                 "t_primary": np.asarray(raw["t"]), "source_info": _group(raw, "source_info"),
                 "primary_stride": primary_stride(path)})
    return ux, uy, meta


def load_fem_damage_data(path: str, device=None):
    """(ux, uy, metadata) of a damage data file (DI:608-674): as
    load_reference_data, plus damage_mask (or None) and damage_info."""
    raw = _read(path)
    ux, uy = _fields(raw, device)
    meta = {"x_original": np.asarray(raw["x"]), "y_original": np.asarray(raw["y"]),
            "t_primary": np.asarray(raw["t"]), "source_info": _group(raw, "source_info"),
            "primary_stride": primary_stride(path),
            "damage_mask": np.asarray(raw["damage_mask"]) if "damage_mask" in raw else None,
            "damage_info": _group(raw, "damage_info")}
    return ux, uy, meta


HDF5_EXTENSIONS = (".mat", ".h5", ".hdf5")


def save_reference_data(path: str, ux, uy, x, y, t, source_info: Optional[Dict[str, object]] = None,
                        damage_mask=None, damage_info: Optional[Dict[str, object]] = None) -> str:
    """Write ux, uy (nt, nx, ny) float32 and the coordinate / metadata entries
    of the reference layout; returns the path.  .mat / .h5 / .hdf5: an HDF5
    file as the reference's writers make it (PF:908-945, DF:1627-1684: ux, uy
    in chunks of one frame with gzip 4; damage_info keys "region_<i>/<name>"
    become region subgroups); .npz: a compressed NumPy archive.  Put the
    primary stride in the name ("_ds5") as the reference does."""
    if not path.endswith(".npz") and not path.endswith(HDF5_EXTENSIONS):
        raise ValueError(f"save_reference_data writes .npz archives or HDF5 files ({', '.join(HDF5_EXTENSIONS)})")
    arrays = {"ux": np.asarray(ux, np.float32), "uy": np.asarray(uy, np.float32), "x": np.asarray(x),
              "y": np.asarray(y), "t": np.asarray(t)}
    if arrays["ux"].ndim != 3 or arrays["ux"].shape != arrays["uy"].shape:
        raise ValueError("ux, uy must be (nt, nx, ny) arrays of one shape")
    if damage_mask is not None:
        arrays["damage_mask"] = np.asarray(damage_mask, np.uint8)
    for g, d in (("source_info", source_info), ("damage_info", damage_info)):
        for k, v in (d or {}).items():
            a = np.asarray(v)
            if a.dtype == object:
                raise ValueError(f"{g}/{k}: only numeric or string values are stored")
            arrays[f"{g}/{k}"] = a
    if path.endswith(".npz"):
        np.savez_compressed(path, **arrays)
        return path
    from . import hdf5
    tree: Dict[str, object] = {}
    for key, a in arrays.items():
        node = tree
        parts = key.split("/")
        for part in parts[:-1]:
            node = node.setdefault(part, {})
        node[parts[-1]] = a.item().decode("utf-8") if a.dtype.kind == "S" and a.ndim == 0 else \
            (str(a) if a.dtype.kind == "U" and a.ndim == 0 else a)
    nt, nx, ny = arrays["ux"].shape
    hdf5.write(path, tree, chunked={"ux": (1, nx, ny), "uy": (1, nx, ny)}, level=4)
    return path
