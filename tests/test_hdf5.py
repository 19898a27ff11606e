"""The HDF5 reader and writer (ernn/hdf5.py), CPU only.

Reader: files written by the HDF5 C library itself (the library h5py wraps),
in the reference's observed-data layout (PF:902-945, DF:1619-1684), at both
format bounds the library writes (version-0 files with B-tree chunk indexes,
as h5py writes by default; libver latest with fixed-array indexes, paged and
not), with deflate and shuffle: every dataset equals the values the fixture
generator wrote (tests/golden/make_hdf5_fixtures.py).
Writer: its files read back through the same C library where it is
installed, and through the reader everywhere; a chunk count past one B-tree
node; scalars, strings, nested groups.
"""
import os
import sys

import numpy as np
import pytest

from ernn import hdf5

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_hdf5_fixtures as fixtures   # noqa: E402

FILES = ["hdf5_earliest", "hdf5_latest", "hdf5_shuffle", "hdf5_latest_paged"]


def _equal(got, want):
    if want.dtype.kind == "S":
        return got == want[()]
    return np.asarray(got).dtype == want.dtype and np.array_equal(np.asarray(got), want)


@pytest.mark.parametrize("name", FILES)
def test_reads_files_written_by_the_hdf5_library(name):
    want = np.load(os.path.join(GOLDEN, name + ".expected.npz"))
    with hdf5.File(os.path.join(GOLDEN, name + ".h5")) as f:
        assert sorted(f.keys()) == ["damage_info", "damage_mask", "source_info", "t", "ux", "uy", "x", "y"]
        assert sorted(f["damage_info"].keys()) == ["damage_percentage", "region_0", "region_1", "total_damaged_points"]
        assert "source_info" in f and "damage_info/region_1/type" in f and "nope" not in f
        for k in want.files:
            assert _equal(f[k][()], want[k]), k
        ux = f["ux"]
        assert ux.shape == want["ux"].shape and ux.dtype == np.float32
        assert np.array_equal(ux[2:4, 1], want["ux"][2:4, 1])


def test_reference_loader_reads_the_library_file():
    """The reference's damage loader path (DI:617-645) on a file the library wrote."""
    from ernn import data
    import shutil
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "anisotropic_ricker_damage_data_ds5_FEM_fixed_bc.mat")
        shutil.copy(os.path.join(GOLDEN, "hdf5_earliest.h5"), p)
        ux, uy, meta = data.load_fem_damage_data(p)
    want = np.load(os.path.join(GOLDEN, "hdf5_earliest.expected.npz"))
    assert np.array_equal(ux[0].numpy(), want["ux"]) and np.array_equal(uy[0].numpy(), want["uy"])
    assert meta["primary_stride"] == 5
    assert meta["damage_info"]["damage_percentage"] == want["damage_info/damage_percentage"]
    assert meta["source_info"]["component"] == b"y"
    assert np.array_equal(meta["damage_mask"], want["damage_mask"])


def _tree(seed=7, nt=75):
    return fixtures.layout(seed, nt, 9, 8)   # 75 frames: more chunks than one B-tree node holds


def test_writer_round_trip_through_the_reader(tmp_path):
    tree = _tree()
    p = str(tmp_path / "w.h5")
    hdf5.write(p, tree, chunked={"ux": (1, 9, 8), "uy": (1, 9, 8)})
    flat = fixtures.flatten(tree)
    with hdf5.File(p) as f:
        for k, v in flat.items():
            assert _equal(f[k][()], v), k
        assert f["source_info/nx"][()] == 9 and f["source_info/nx"].shape == ()


def test_writer_output_reads_back_in_the_hdf5_library(tmp_path):
    lib = fixtures.load_libhdf5()
    if lib is None:
        pytest.skip("the HDF5 C library is not installed")
    h5 = fixtures.H5(lib)
    tree = _tree(seed=8)
    p = str(tmp_path / "w.h5")
    hdf5.write(p, tree, chunked={"ux": (1, 9, 8), "uy": (1, 9, 8)})
    for k, v in fixtures.flatten(tree).items():
        if v.dtype.kind == "S":
            assert h5.read_string(p, k) == v[()], k
        else:
            assert np.array_equal(h5.read(p, k, v.dtype, v.shape), v), k
    assert h5.exists(p, "damage_info") and not h5.exists(p, "missing")


def test_partial_edge_chunks_and_multi_dimensional_chunking(tmp_path):
    a = np.arange(7 * 10 * 9, dtype=np.float64).reshape(7, 10, 9)
    p = str(tmp_path / "c.h5")
    hdf5.write(p, {"a": a, "b": np.arange(5, dtype=np.int32)}, chunked={"a": (3, 4, 5)})
    with hdf5.File(p) as f:
        assert np.array_equal(f["a"][:], a)
        assert np.array_equal(f["b"][:], np.arange(5))


def test_not_hdf5_is_refused(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"\0" * 4096)
    with pytest.raises(hdf5.HDF5Error):
        hdf5.File(str(p))
    with pytest.raises(ValueError):
        hdf5.File(str(p), "w")
