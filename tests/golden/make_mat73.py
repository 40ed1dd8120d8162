"""Write the MATLAB -v7.3 fixtures of tests/test_mat73.py with the HDF5 library.

The files are written by libhdf5 itself (the conda build in this image,
/opt/conda/lib/libhdf5.so.103, HDF5 1.10.6), called through ctypes, so the
reader under test (mesolf_amd/csrc/mat73.cpp) is checked against an
independent writer.  They follow what MATLAB's save -v7.3 writes: a 512-byte
user block holding MATLAB's 128-byte text header, one dataset per variable
with a MATLAB_class string attribute, dims in reversed (C) order, structs as
groups, empty arrays as their dims with MATLAB_empty = 1.

mat73_psf.mat (default library format: superblock 0, v1 object headers,
symbol-table groups, v1 B-tree chunk index -- what MATLAB writes):
  H         single [45,45,15,15,2]  chunked [45,45,15,3,1] (C order), deflate 3
  Ht        double [45,45,15,15,2]  chunked [45,45,15,5,2], shuffle + deflate
  Nnum      double 1x1               contiguous
  x1objspace double 1x5, x3objspace double 1x2  (metres)
  CAindex   int32 2x2
  mask      logical 7x3, big_i16 int16 4x5 stored big-endian, f64be 3x3
            big-endian double, fl32 single 6x4 chunked with fletcher32
  holes     double 10x9 chunked 4x4, fill value 7, only rows 0-3 written
  empty     double 0x3 (MATLAB_empty)
  meta      struct with field scale (double 1x1)
  cplx      complex double 2x1 (compound; refused by the reader)
mat73_latest.mat (libver latest: superblock 3, v2 object headers, link
messages, layout v4 with fixed-array, single-chunk and implicit indexes):
  H, Ht as above (fixed array), one (single 5x4 single chunk, deflate),
  imp (double 6x6 chunks 3x3, early allocation: implicit index), cont
  (contiguous uint16 3x4), small (compact int8 2x2).

The expected arrays other than H and Ht (recomputed by psf_arrays()) go to
mat73_expected.npz (MATLAB shapes; struct fields as "meta__scale").
Run: python tests/golden/make_mat73.py   (needs /opt/conda/lib/libhdf5.so.103)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBHDF5 = "/opt/conda/lib/libhdf5.so.103"

H5F_ACC_TRUNC = 0x0002
H5S_SCALAR = 0
H5S_SELECT_SET = 0
H5D_COMPACT, H5D_CONTIGUOUS, H5D_CHUNKED = 0, 1, 2
H5D_ALLOC_TIME_EARLY = 1
H5F_LIBVER_LATEST = 2  # H5F_LIBVER_V110 in 1.10
H5T_STR_NULLTERM = 0
H5T_COMPOUND = 6


class H5:
    def __init__(self):
        self.l = C.CDLL(LIBHDF5)
        self.l.H5open()
        for fn in ("H5Fcreate", "H5Pcreate", "H5Screate_simple", "H5Screate", "H5Dcreate2", "H5Tcopy",
                   "H5Acreate2", "H5Gcreate2", "H5Dget_space", "H5Tcreate"):
            getattr(self.l, fn).restype = C.c_int64
        self.l.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, C.c_int64, C.c_int64]
        self.l.H5Pcreate.argtypes = [C.c_int64]
        self.l.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        self.l.H5Screate.argtypes = [C.c_int]
        self.l.H5Dcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int64]
        self.l.H5Dwrite.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        self.l.H5Tcopy.argtypes = [C.c_int64]
        self.l.H5Tset_size.argtypes = [C.c_int64, C.c_size_t]
        self.l.H5Tset_strpad.argtypes = [C.c_int64, C.c_int]
        self.l.H5Tcreate.argtypes = [C.c_int, C.c_size_t]
        self.l.H5Tinsert.argtypes = [C.c_int64, C.c_char_p, C.c_size_t, C.c_int64]
        self.l.H5Acreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
        self.l.H5Awrite.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
        self.l.H5Gcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64]
        self.l.H5Pset_userblock.argtypes = [C.c_int64, C.c_uint64]
        self.l.H5Pset_libver_bounds.argtypes = [C.c_int64, C.c_int, C.c_int]
        self.l.H5Pset_chunk.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_uint64)]
        self.l.H5Pset_deflate.argtypes = [C.c_int64, C.c_uint]
        self.l.H5Pset_shuffle.argtypes = [C.c_int64]
        self.l.H5Pset_fletcher32.argtypes = [C.c_int64]
        self.l.H5Pset_layout.argtypes = [C.c_int64, C.c_int]
        self.l.H5Pset_alloc_time.argtypes = [C.c_int64, C.c_int]
        self.l.H5Pset_fill_value.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
        self.l.H5Dget_space.argtypes = [C.c_int64]
        self.l.H5Sselect_hyperslab.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64)]
        for fn in ("H5Fclose", "H5Dclose", "H5Sclose", "H5Pclose", "H5Tclose", "H5Aclose", "H5Gclose"):
            getattr(self.l, fn).argtypes = [C.c_int64]

    def g(self, name):
        return C.c_int64.in_dll(self.l, name).value

    @staticmethod
    def dims(shape):
        return (C.c_uint64 * max(1, len(shape)))(*shape)

    def ok(self, v, what):
        if v < 0:
            raise RuntimeError(f"HDF5 call failed: {what}")
        return v


H5T = {
    np.dtype("<f4"): "H5T_IEEE_F32LE_g", np.dtype("<f8"): "H5T_IEEE_F64LE_g",
    np.dtype(">f8"): "H5T_IEEE_F64BE_g", np.dtype("<i4"): "H5T_STD_I32LE_g",
    np.dtype(">i2"): "H5T_STD_I16BE_g", np.dtype("<u1"): "H5T_STD_U8LE_g",
    np.dtype("<u2"): "H5T_STD_U16LE_g", np.dtype("<u8"): "H5T_STD_U64LE_g",
    np.dtype("<i1"): "H5T_STD_I8LE_g",
}


def _matlab_header():
    txt = b"MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: Thu Jan  1 00:00:00 1970 HDF5 schema 1.00 ."
    return txt.ljust(116, b" ") + b"\x00" * 8 + b"\x00\x02" + b"IM"


class Writer:
    def __init__(self, h5: H5, path: str, latest: bool):
        self.h5, self.path = h5, path
        l = h5.l
        fcpl = l.H5Pcreate(h5.g("H5P_CLS_FILE_CREATE_ID_g"))
        h5.ok(l.H5Pset_userblock(fcpl, 512), "userblock")
        fapl = l.H5Pcreate(h5.g("H5P_CLS_FILE_ACCESS_ID_g"))
        if latest:
            h5.ok(l.H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST), "libver")
        self.f = h5.ok(l.H5Fcreate(path.encode(), H5F_ACC_TRUNC, fcpl, fapl), "create")
        l.H5Pclose(fcpl)
        l.H5Pclose(fapl)

    def _attr_str(self, obj, name, value: str):
        l, h5 = self.h5.l, self.h5
        t = l.H5Tcopy(h5.g("H5T_C_S1_g"))
        l.H5Tset_size(t, len(value))
        l.H5Tset_strpad(t, H5T_STR_NULLTERM)
        sp = l.H5Screate(H5S_SCALAR)
        a = h5.ok(l.H5Acreate2(obj, name.encode(), t, sp, 0, 0), "attr")
        buf = C.create_string_buffer(value.encode(), len(value))
        h5.ok(l.H5Awrite(a, t, buf), "attr write")
        l.H5Aclose(a)
        l.H5Sclose(sp)
        l.H5Tclose(t)

    def _attr_u8(self, obj, name, value: int):
        l, h5 = self.h5.l, self.h5
        t = h5.g("H5T_STD_U8LE_g")
        sp = l.H5Screate(H5S_SCALAR)
        a = h5.ok(l.H5Acreate2(obj, name.encode(), t, sp, 0, 0), "attr")
        v = C.c_uint8(value)
        h5.ok(l.H5Awrite(a, t, C.byref(v)), "attr write")
        l.H5Aclose(a)
        l.H5Sclose(sp)

    def dataset(self, loc, name, arr, mclass, layout="contiguous", chunk=None, deflate=0, shuffle=False,
                fletcher=False, fill=None, rows_written=None, early=False, empty=False):
        """arr in MATLAB shape; stored with reversed dims (C order of arr.T)."""
        l, h5 = self.h5.l, self.h5
        a = np.asarray(arr)
        data = np.ascontiguousarray(a.T)  # HDF5 row-major of the reversed dims
        shape = data.shape
        sp = h5.ok(l.H5Screate_simple(len(shape), H5.dims(shape), None), "space")
        dcpl = l.H5Pcreate(h5.g("H5P_CLS_DATASET_CREATE_ID_g"))
        t = h5.g(H5T[data.dtype])
        if layout == "compact":
            l.H5Pset_layout(dcpl, H5D_COMPACT)
        elif layout == "chunked":
            h5.ok(l.H5Pset_chunk(dcpl, len(chunk), H5.dims(chunk)), "chunk")
            if fletcher:
                l.H5Pset_fletcher32(dcpl)
            if shuffle:
                l.H5Pset_shuffle(dcpl)
            if deflate:
                l.H5Pset_deflate(dcpl, deflate)
            if fill is not None:
                fv = np.array(fill, dtype=data.dtype)
                l.H5Pset_fill_value(dcpl, t, fv.ctypes.data)
            if early:
                l.H5Pset_alloc_time(dcpl, H5D_ALLOC_TIME_EARLY)
        d = h5.ok(l.H5Dcreate2(loc, name.encode(), t, sp, 0, dcpl, 0), f"dataset {name}")
        if rows_written is None:
            h5.ok(l.H5Dwrite(d, t, 0, 0, 0, data.ctypes.data), f"write {name}")
        else:
            # write only the first rows_written rows of the HDF5 array
            part = np.ascontiguousarray(data[:rows_written])
            ms = l.H5Screate_simple(len(shape), H5.dims(part.shape), None)
            fs = l.H5Dget_space(d)
            start = H5.dims([0] * len(shape))
            count = H5.dims(part.shape)
            h5.ok(l.H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, None, count, None), "hyperslab")
            h5.ok(l.H5Dwrite(d, t, ms, fs, 0, part.ctypes.data), f"write {name}")
            l.H5Sclose(ms)
            l.H5Sclose(fs)
        self._attr_str(d, "MATLAB_class", mclass)
        if empty:
            self._attr_u8(d, "MATLAB_empty", 1)
        l.H5Dclose(d)
        l.H5Pclose(dcpl)
        l.H5Sclose(sp)

    def complex_dataset(self, name, z):
        l, h5 = self.h5.l, self.h5
        t = l.H5Tcreate(H5T_COMPOUND, 16)
        l.H5Tinsert(t, b"real", 0, h5.g("H5T_IEEE_F64LE_g"))
        l.H5Tinsert(t, b"imag", 8, h5.g("H5T_IEEE_F64LE_g"))
        data = np.ascontiguousarray(np.asarray(z, np.complex128).T)
        sp = l.H5Screate_simple(data.ndim, H5.dims(data.shape), None)
        d = h5.ok(l.H5Dcreate2(self.f, name.encode(), t, sp, 0, 0, 0), "complex")
        h5.ok(l.H5Dwrite(d, t, 0, 0, 0, data.ctypes.data), "complex write")
        self._attr_str(d, "MATLAB_class", "double")
        l.H5Dclose(d)
        l.H5Sclose(sp)
        l.H5Tclose(t)

    def group(self, name, mclass="struct"):
        g = self.h5.ok(self.h5.l.H5Gcreate2(self.f, name.encode(), 0, 0, 0), "group")
        self._attr_str(g, "MATLAB_class", mclass)
        return g

    def close(self):
        self.h5.l.H5Fclose(self.f)
        with open(self.path, "r+b") as fh:
            fh.write(_matlab_header())


def psf_arrays():
    """Deterministic [45,45,15,15,2] H, Ht (one lenslet ring around the centre
    lenslet, R = 1).  Values are k/64 for small integers k, exact in FP32;
    the per-view factor makes the view weights of recon:30-38 differ and view
    (7, 3) falls under the 0.2 cut."""
    S, N, Z = 45, 15, 2
    x1, x2, a, b, z = np.meshgrid(*(np.arange(n) for n in (S, S, N, N, Z)), indexing="ij")
    d2 = (x1 - 22 - (a - 7)) ** 2 + (x2 - 22 - (b - 7)) ** 2
    view = ((a * 3 + b) % 5 + 2) * np.where((a == 7) & (b == 3), 0, 1) + np.where((a == 7) & (b == 3), 1, 0)
    H = np.where(d2 <= 90 + 60 * z, ((x1 * 7 + x2 * 3 + a + 2 * b + 5 * z) % 13 + 1) * view / 64.0, 0.0)
    Ht = np.where(d2 <= 100 + 30 * z, ((x1 * 5 + x2 * 11 + 3 * a + b + z) % 11 + 1) / 64.0, 0.0)
    return np.asfortranarray(H.astype(np.float32)), np.asfortranarray(Ht.astype(np.float64))


def build():
    h5 = H5()
    rng = np.random.default_rng(73)
    H, Ht = psf_arrays()
    exp = {"H": H, "Ht": Ht}

    w = Writer(h5, os.path.join(HERE, "mat73_psf.mat"), latest=False)
    w.dataset(w.f, "H", H, "single", "chunked", chunk=(1, 3, 15, 45, 45), deflate=3)
    w.dataset(w.f, "Ht", Ht, "double", "chunked", chunk=(2, 5, 15, 45, 45), deflate=3, shuffle=True)
    exp["Nnum"] = np.array([[15.0]])
    w.dataset(w.f, "Nnum", exp["Nnum"], "double")
    exp["x1objspace"] = np.array([[-2.0e-6, 0.0, 2.0e-6, 4.0e-6, 6.0e-6]])
    exp["x3objspace"] = np.array([[-1.0e-5, -6.0e-6]])
    w.dataset(w.f, "x1objspace", exp["x1objspace"], "double")
    w.dataset(w.f, "x3objspace", exp["x3objspace"], "double")
    exp["CAindex"] = np.array([[1, 15], [3, 13]], dtype=np.int32)
    w.dataset(w.f, "CAindex", exp["CAindex"], "int32")
    exp["mask"] = (rng.random((7, 3)) > 0.5).astype(np.uint8)
    w.dataset(w.f, "mask", exp["mask"], "logical")
    exp["big_i16"] = rng.integers(-30000, 30000, (4, 5)).astype(">i2")
    w.dataset(w.f, "big_i16", exp["big_i16"], "int16")
    exp["f64be"] = rng.standard_normal((3, 3)).astype(">f8")
    w.dataset(w.f, "f64be", exp["f64be"], "double", "chunked", chunk=(2, 2), deflate=1)
    exp["fl32"] = rng.standard_normal((6, 4)).astype("<f4")
    w.dataset(w.f, "fl32", exp["fl32"], "single", "chunked", chunk=(3, 4), fletcher=True)
    holes = rng.standard_normal((10, 9))
    # HDF5 dims (9, 10): rows 0-3 of the HDF5 array = MATLAB columns 0-3
    w.dataset(w.f, "holes", holes, "double", "chunked", chunk=(4, 4), fill=7.0, rows_written=4)
    hx = np.full((10, 9), 7.0)
    hx[:, :4] = holes[:, :4]
    exp["holes"] = hx
    w.dataset(w.f, "empty", np.array([[0, 3]], dtype="<u8"), "double", empty=True)
    g = w.group("meta")
    exp["meta/scale"] = np.array([[0.25]])
    w.dataset(g, "scale", exp["meta/scale"], "double")
    h5.l.H5Gclose(g)
    w.complex_dataset("cplx", np.array([[1 + 2j], [3 - 1j]]).T)
    w.close()

    w = Writer(h5, os.path.join(HERE, "mat73_latest.mat"), latest=True)
    w.dataset(w.f, "H", H, "single", "chunked", chunk=(1, 3, 15, 45, 45), deflate=3)
    w.dataset(w.f, "Ht", Ht, "double", "chunked", chunk=(2, 5, 15, 45, 45), deflate=3, shuffle=True)
    w.dataset(w.f, "Nnum", exp["Nnum"], "double")
    exp["one"] = rng.standard_normal((5, 4)).astype("<f4")
    w.dataset(w.f, "one", exp["one"], "single", "chunked", chunk=(4, 5), deflate=2)
    exp["imp"] = rng.standard_normal((6, 6))
    w.dataset(w.f, "imp", exp["imp"], "double", "chunked", chunk=(3, 3), early=True)
    exp["cont"] = rng.integers(0, 65535, (3, 4)).astype("<u2")
    w.dataset(w.f, "cont", exp["cont"], "uint16")
    exp["small"] = rng.integers(-128, 127, (2, 2)).astype("<i1")
    w.dataset(w.f, "small", exp["small"], "int8", "compact")
    w.close()

    # H and Ht are not stored: psf_arrays() recomputes them exactly
    np.savez_compressed(os.path.join(HERE, "mat73_expected.npz"),
                        **{k.replace("/", "__"): np.asarray(v, dtype=np.float64) for k, v in exp.items()
                           if k not in ("H", "Ht")})


if __name__ == "__main__":
    build()
    for f in ("mat73_psf.mat", "mat73_latest.mat", "mat73_expected.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))
