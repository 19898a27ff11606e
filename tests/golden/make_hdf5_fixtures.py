"""Writes the HDF5 fixtures of tests/test_hdf5.py with the HDF5 C library
(the library h5py wraps; h5py itself is not installed here), through ctypes.

Each file has the layout of the reference's FEM writers
(Parameter_inversion/FEM_generate_data.py:902-945,
Damage_inversion/FEM_generate_data.py:1619-1684): ux, uy (nt_ds, nx, ny)
float32 in chunks of one frame with gzip 4, x, y, t float32, damage_mask
uint8, a source_info group of scalars, arrays and strings, a damage_info
group with region_<i> subgroups.  Small sizes; the values are seeded and
saved beside each file as <name>.expected.npz (group members as "group/name"
keys), which is what the tests compare against.

  hdf5_earliest.h5  HDF5's default format bounds (what h5py writes by default
                    and what libver='earliest' asks for): version-0
                    superblock, symbol-table groups, version-1 B-tree chunks
  hdf5_latest.h5    libver latest: version-2 superblock, version-2 object
                    headers, link messages, fixed-array chunk index
  hdf5_shuffle.h5   earliest bounds with the shuffle filter before deflate and
                    more chunks than one B-tree node holds
  hdf5_latest_paged.h5  libver latest with more chunks than one fixed-array page

Run from the repository root:  python tests/golden/make_hdf5_fixtures.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_CANDIDATES = ["libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"]


def load_libhdf5():
    for name in LIB_CANDIDATES:
        try:
            lib = C.CDLL(name)
        except OSError:
            continue
        lib.H5open()
        return lib
    return None


class H5:
    """The few HDF5 C calls the fixtures need."""

    def __init__(self, lib):
        self.lib = lib
        hid = C.c_int64
        for fn, res, args in [
            ("H5Fcreate", hid, [C.c_char_p, C.c_uint, hid, hid]),
            ("H5Fopen", hid, [C.c_char_p, C.c_uint, hid]),
            ("H5Fclose", C.c_int, [hid]),
            ("H5Pcreate", hid, [hid]),
            ("H5Pclose", C.c_int, [hid]),
            ("H5Pset_chunk", C.c_int, [hid, C.c_int, C.POINTER(C.c_ulonglong)]),
            ("H5Pset_deflate", C.c_int, [hid, C.c_uint]),
            ("H5Pset_shuffle", C.c_int, [hid]),
            ("H5Pset_libver_bounds", C.c_int, [hid, C.c_int, C.c_int]),
            ("H5Screate", hid, [C.c_int]),
            ("H5Screate_simple", hid, [C.c_int, C.POINTER(C.c_ulonglong), C.c_void_p]),
            ("H5Sclose", C.c_int, [hid]),
            ("H5Sget_simple_extent_ndims", C.c_int, [hid]),
            ("H5Sget_simple_extent_dims", C.c_int, [hid, C.POINTER(C.c_ulonglong), C.c_void_p]),
            ("H5Tcopy", hid, [hid]),
            ("H5Tset_size", C.c_int, [hid, C.c_size_t]),
            ("H5Tset_cset", C.c_int, [hid, C.c_int]),
            ("H5Tclose", C.c_int, [hid]),
            ("H5Dcreate2", hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
            ("H5Dopen2", hid, [hid, C.c_char_p, hid]),
            ("H5Dwrite", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            ("H5Dread", C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            ("H5Dget_space", hid, [hid]),
            ("H5Dclose", C.c_int, [hid]),
            ("H5Gcreate2", hid, [hid, C.c_char_p, hid, hid, hid]),
            ("H5Gclose", C.c_int, [hid]),
            ("H5Lexists", C.c_int, [hid, C.c_char_p, hid]),
            ("H5Dvlen_reclaim", C.c_int, [hid, hid, hid, C.c_void_p]),
        ]:
            f = getattr(lib, fn)
            f.restype, f.argtypes = res, args
        g = lambda n: C.c_int64.in_dll(lib, n).value
        self.FLOAT, self.DOUBLE = g("H5T_NATIVE_FLOAT_g"), g("H5T_NATIVE_DOUBLE_g")
        self.INT64, self.UINT8 = g("H5T_NATIVE_LLONG_g"), g("H5T_NATIVE_UCHAR_g")
        self.INT32 = g("H5T_NATIVE_INT_g")
        self.C_S1 = g("H5T_C_S1_g")
        self.DCPL, self.FAPL = g("H5P_CLS_DATASET_CREATE_ID_g"), g("H5P_CLS_FILE_ACCESS_ID_g")
        self.NATIVE = {np.dtype(np.float32): self.FLOAT, np.dtype(np.float64): self.DOUBLE,
                       np.dtype(np.int64): self.INT64, np.dtype(np.uint8): self.UINT8, np.dtype(np.int32): self.INT32}

    @staticmethod
    def dims(shape):
        return (C.c_ulonglong * max(len(shape), 1))(*shape)

    def check(self, v, what):
        if v < 0:
            raise RuntimeError(f"HDF5 call failed: {what}")
        return v

    def create(self, path, latest=False):
        fapl = 0
        if latest:
            fapl = self.check(self.lib.H5Pcreate(self.FAPL), "H5Pcreate fapl")
            self.check(self.lib.H5Pset_libver_bounds(fapl, 2, 2), "H5Pset_libver_bounds")   # H5F_LIBVER_LATEST
        f = self.check(self.lib.H5Fcreate(path.encode(), 0x0002, 0, fapl), "H5Fcreate")   # H5F_ACC_TRUNC
        if fapl:
            self.lib.H5Pclose(fapl)
        return f

    def dataset(self, loc, name, value, chunks=None, level=4, shuffle=False):
        if isinstance(value, str):
            t = self.check(self.lib.H5Tcopy(self.C_S1), "H5Tcopy")
            self.lib.H5Tset_size(t, C.c_size_t(-1).value)   # H5T_VARIABLE
            self.lib.H5Tset_cset(t, 1)                      # UTF-8
            sp = self.lib.H5Screate(0)                      # H5S_SCALAR
            d = self.check(self.lib.H5Dcreate2(loc, name.encode(), t, sp, 0, 0, 0), f"H5Dcreate2 {name}")
            ptr = C.c_char_p(value.encode())
            self.check(self.lib.H5Dwrite(d, t, 0, 0, 0, C.addressof(ptr)), f"H5Dwrite {name}")
            self.lib.H5Dclose(d); self.lib.H5Sclose(sp); self.lib.H5Tclose(t)
            return
        a = np.ascontiguousarray(value) if np.ndim(value) else np.asarray(value)
        sp = self.lib.H5Screate(0) if a.ndim == 0 else self.lib.H5Screate_simple(a.ndim, self.dims(a.shape), None)
        dcpl = 0
        if chunks is not None:
            dcpl = self.lib.H5Pcreate(self.DCPL)
            self.check(self.lib.H5Pset_chunk(dcpl, len(chunks), self.dims(chunks)), "H5Pset_chunk")
            if shuffle:
                self.check(self.lib.H5Pset_shuffle(dcpl), "H5Pset_shuffle")
            self.check(self.lib.H5Pset_deflate(dcpl, level), "H5Pset_deflate")
        t = self.NATIVE[a.dtype]
        d = self.check(self.lib.H5Dcreate2(loc, name.encode(), t, sp, 0, dcpl, 0), f"H5Dcreate2 {name}")
        buf = np.ascontiguousarray(a).reshape(-1) if a.ndim else np.array([a], a.dtype)
        self.check(self.lib.H5Dwrite(d, t, 0, 0, 0, buf.ctypes.data), f"H5Dwrite {name}")
        self.lib.H5Dclose(d); self.lib.H5Sclose(sp)
        if dcpl:
            self.lib.H5Pclose(dcpl)

    def group(self, loc, name):
        return self.check(self.lib.H5Gcreate2(loc, name.encode(), 0, 0, 0), f"H5Gcreate2 {name}")

    def read(self, path, name, dtype, shape):
        """One numeric dataset through the library (used by the writer test)."""
        f = self.check(self.lib.H5Fopen(path.encode(), 0, 0), "H5Fopen")
        d = self.check(self.lib.H5Dopen2(f, name.encode(), 0), f"H5Dopen2 {name}")
        out = np.empty(shape, dtype)
        buf = out.reshape(-1) if out.ndim else out.reshape(1)
        self.check(self.lib.H5Dread(d, self.NATIVE[np.dtype(dtype)], 0, 0, 0, buf.ctypes.data), f"H5Dread {name}")
        self.lib.H5Dclose(d); self.lib.H5Fclose(f)
        return out

    def read_string(self, path, name):
        f = self.check(self.lib.H5Fopen(path.encode(), 0, 0), "H5Fopen")
        d = self.check(self.lib.H5Dopen2(f, name.encode(), 0), f"H5Dopen2 {name}")
        t = self.lib.H5Tcopy(self.C_S1)
        self.lib.H5Tset_size(t, C.c_size_t(-1).value)
        self.lib.H5Tset_cset(t, 1)   # UTF-8, as h5py stores str
        ptr = C.c_char_p()
        self.check(self.lib.H5Dread(d, t, 0, 0, 0, C.addressof(ptr)), f"H5Dread {name}")
        v = ptr.value
        self.lib.H5Tclose(t); self.lib.H5Dclose(d); self.lib.H5Fclose(f)
        return v

    def exists(self, path, name):
        f = self.check(self.lib.H5Fopen(path.encode(), 0, 0), "H5Fopen")
        r = self.lib.H5Lexists(f, name.encode(), 0)
        self.lib.H5Fclose(f)
        return r > 0


def layout(seed, nt, nx, ny):
    """The reference's file layout as a nested dict of seeded values."""
    rng = np.random.default_rng(seed)
    ux = (rng.standard_normal((nt, nx, ny)) * 1e-9).astype(np.float32)
    uy = (rng.standard_normal((nt, nx, ny)) * 1e-9).astype(np.float32)
    return {
        "ux": ux, "uy": uy,
        "x": np.linspace(-0.2, 0.2, nx).astype(np.float32),
        "y": np.linspace(-0.2, 0.2, ny).astype(np.float32),
        "t": (np.arange(nt) * 5 * 3e-8).astype(np.float32),
        "damage_mask": (rng.random((nx, ny)) < 0.1).astype(np.uint8),
        "damage_info": {
            "total_damaged_points": np.int64(37),
            "damage_percentage": np.float64(4.25),
            "region_0": {"type": "ellipse", "center": np.array([0.05, -0.02]), "size": np.array([0.03, 0.02]),
                         "reduction_factor": np.float64(0.5), "n_points": np.int64(21)},
            "region_1": {"type": "circle", "center": np.array([-0.04, 0.06]), "size": np.float64(0.015),
                         "reduction_factor": np.float64(0.5), "n_points": np.int64(16)},
        },
        "source_info": {"source_f0": np.float64(100e3), "source_t0": np.float64(10e-6),
                        "source_amplitude": np.float64(1e10), "gaussian_param": np.float64(5e6),
                        "source_x": np.float64(0.0), "component": "y", "nx": np.int64(nx),
                        "indices": np.array([[3, 4], [5, 6]], np.int64)},
    }


def flatten(tree, prefix=""):
    out = {}
    for k, v in tree.items():
        p = f"{prefix}/{k}" if prefix else k
        if isinstance(v, dict):
            out.update(flatten(v, p))
        else:
            out[p] = np.asarray(v.encode() if isinstance(v, str) else v)
    return out


def write_with_library(h5, path, tree, latest=False, shuffle=False, frame_chunks=True):
    f = h5.create(path, latest)

    def put(loc, node, prefix):
        for k, v in node.items():
            p = f"{prefix}/{k}" if prefix else k
            if isinstance(v, dict):
                g = h5.group(loc, k)
                put(g, v, p)
                h5.lib.H5Gclose(g)
            else:
                chunks = (1,) + tuple(np.shape(v)[1:]) if (p in ("ux", "uy") and frame_chunks) else None
                h5.dataset(loc, k, v, chunks=chunks, shuffle=shuffle and chunks is not None)
    put(f, tree, "")
    h5.lib.H5Fclose(f)


def main():
    lib = load_libhdf5()
    if lib is None:
        raise SystemExit("the HDF5 C library is not available")
    h5 = H5(lib)
    for name, seed, (nt, nx, ny), latest, shuffle in [("hdf5_earliest", 1, (9, 13, 11), False, False),
                                                      ("hdf5_latest", 2, (7, 12, 10), True, False),
                                                      ("hdf5_shuffle", 3, (70, 6, 5), False, True),
                                                      ("hdf5_latest_paged", 4, (1100, 3, 2), True, False)]:
        tree = layout(seed, nt, nx, ny)
        path = os.path.join(HERE, name + ".h5")
        write_with_library(h5, path, tree, latest=latest, shuffle=shuffle)
        np.savez(os.path.join(HERE, name + ".expected.npz"), **flatten(tree))
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
