"""The native MATLAB -v7.3 reader (mesolf_amd/csrc/mat73.cpp) against files it
did not write.

* scipy's testhdf5_7.4_GLNX86.mat is a file MATLAB itself saved with -v7.3;
  its variable must equal testdouble_7.4_GLNX86.mat (a v5 file of the same
  variable) as scipy.io.loadmat reads it.
* PyTables' HDF5 compatibility files hold the 6x5 array i+j in f64/i64/i32,
  little- and big-endian, and a chunked, extended int32 array
  (tables/tests/test_hdf5compat.py states the expected values).
* tests/golden/mat73_psf.mat and mat73_latest.mat were written by libhdf5
  1.10.6 (tests/golden/make_mat73.py) in MATLAB's layout, in the library's
  default format and in its newest format.

No MATLAB PSF file ships with MesoLF, so the PSF loader's output is pinned
through these files and through mlf_psf_create_dense on the same arrays
(tests/test_gpu_mat73.py).  All reads here are host code: no GPU is used.
"""
import os

import numpy as np
import pytest

from mesolf_amd import MatFile, MatFileError
from mesolf_amd.lib import Unsupported, InvalidInput
from tests.golden import make_mat73

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PSF_FILE = os.path.join(GOLDEN, "mat73_psf.mat")
LATEST_FILE = os.path.join(GOLDEN, "mat73_latest.mat")
EXPECTED = np.load(os.path.join(GOLDEN, "mat73_expected.npz"))


def _scipy_data():
    import scipy
    return os.path.join(os.path.dirname(scipy.__file__), "io", "matlab", "tests", "data")


PYTABLES = "/opt/conda/lib/python3.9/site-packages/tables/tests"


def test_matlab_written_v73_file_equals_v5_copy():
    d = _scipy_data()
    path = os.path.join(d, "testhdf5_7.4_GLNX86.mat")
    if not os.path.exists(path):
        pytest.skip("scipy test data not installed")
    import scipy.io as sio
    want = sio.loadmat(os.path.join(d, "testdouble_7.4_GLNX86.mat"))["testdouble"]
    with MatFile(path) as m:
        assert m.variables() == ["testdouble"]
        info = m.info("testdouble")
        assert info["class"] == "double" and info["shape"] == (1, 9)
        got = m.read("testdouble")
    assert got.shape == want.shape and np.array_equal(got, want)
    # and it is pi/4 * (0:8), as scipy's own tests state
    assert np.allclose(got.ravel(), np.arange(9) * np.pi / 4, rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", ["smpl_f64le.h5", "smpl_f64be.h5", "smpl_i64le.h5", "smpl_i64be.h5",
                                  "smpl_i32le.h5", "smpl_i32be.h5"])
def test_pytables_numeric_files(name):
    path = os.path.join(PYTABLES, name)
    if not os.path.exists(path):
        pytest.skip("PyTables test files not installed")
    with MatFile(path) as m:
        a = m.read("TestArray")
    # HDF5 shape (6, 5) row-major = MATLAB shape (5, 6); element (i, j) = i + j
    i, j = np.meshgrid(np.arange(6), np.arange(5), indexing="ij")
    assert np.array_equal(a.T, i + j)


def test_pytables_extendible_chunked_file():
    path = os.path.join(PYTABLES, "smpl_SDSextendible.h5")
    if not os.path.exists(path):
        pytest.skip("PyTables test files not installed")
    want = np.array([[1, 1, 1, 3, 3], [1, 1, 1, 3, 3], [1, 1, 1, 0, 0]] + [[2, 0, 0, 0, 0]] * 7)
    with MatFile(path) as m:
        assert np.array_equal(m.read("ExtendibleArray").T, want)


def test_psf_file_variables_and_classes():
    with MatFile(PSF_FILE) as m:
        names = m.variables()
        assert names == sorted(["CAindex", "H", "Ht", "Nnum", "big_i16", "cplx", "empty", "f64be", "fl32",
                                "holes", "mask", "meta", "x1objspace", "x3objspace"])
        assert m.info("H") == {"class": "single", "shape": (45, 45, 15, 15, 2), "numel": 45 * 45 * 225 * 2,
                               "numeric": True, "stored_bytes": 4}
        assert m.info("Ht")["class"] == "double" and m.info("Ht")["stored_bytes"] == 8
        assert m.info("mask")["class"] == "logical"
        assert m.info("meta")["class"] == "struct" and not m.info("meta")["numeric"]
        assert m.info("cplx")["class"] == "double" and not m.info("cplx")["numeric"]
        e = m.info("empty")
        assert e["numel"] == 0 and e["shape"] == (0, 3)


@pytest.mark.parametrize("path", [PSF_FILE, LATEST_FILE])
def test_psf_arrays_round_trip(path):
    H, Ht = make_mat73.psf_arrays()
    with MatFile(path) as m:
        h32 = m.read("H", np.float32)
        ht64 = m.read("Ht")
        assert np.array_equal(h32, H) and h32.flags.f_contiguous
        assert np.array_equal(ht64, Ht)
        # float64 of single data and float32 of double data convert per element
        assert np.array_equal(m.read("H"), H.astype(np.float64))
        assert np.array_equal(m.read("Ht", np.float32), Ht.astype(np.float32))
        assert m.read("Nnum")[0, 0] == 15.0


def test_every_layout_and_filter():
    with MatFile(PSF_FILE) as m:
        for k in ("Nnum", "x1objspace", "x3objspace", "CAindex", "mask", "big_i16", "f64be", "fl32", "holes"):
            got = m.read(k)
            assert got.shape == EXPECTED[k].shape and np.array_equal(got, EXPECTED[k]), k
        assert np.array_equal(m.read("meta/scale"), EXPECTED["meta__scale"])
        assert m.read("empty").size == 0
    with MatFile(LATEST_FILE) as m:
        for k in ("Nnum", "one", "imp", "cont", "small"):
            got = m.read(k)
            assert got.shape == EXPECTED[k].shape and np.array_equal(got, EXPECTED[k]), k


def test_holes_take_the_fill_value():
    with MatFile(PSF_FILE) as m:
        h = m.read("holes")
    assert np.all(h[:, 4:] == 7.0) and not np.any(h[:, :4] == 7.0)


def test_refusals():
    with MatFile(PSF_FILE) as m:
        with pytest.raises(Unsupported, match="compound"):
            m.read("cplx")
        with pytest.raises(Unsupported, match="struct"):
            m.read("meta")
        with pytest.raises(InvalidInput, match="not found"):
            m.read("nope")
        with pytest.raises(InvalidInput, match="elements"):
            lib = m._lib
            buf = np.zeros(3)
            from mesolf_amd.lib import _check, _fptr
            import ctypes as C
            _check(lib.mlf_mat_read_f64(m.handle, b"Nnum", _fptr(buf, C.c_double), 3))


def test_missing_and_malformed_files(tmp_path):
    with pytest.raises(MatFileError, match="cannot open"):
        MatFile(str(tmp_path / "absent.mat"))
    p = tmp_path / "v5.mat"
    p.write_bytes(b"MATLAB 5.0 MAT-file" + b" " * 200)
    with pytest.raises(MatFileError, match="not an HDF5"):
        MatFile(str(p))


def test_truncated_and_corrupted_files_fail_cleanly(tmp_path):
    """Every prefix and a sweep of single-byte corruptions of a real file give
    a clean error or a correct read, never a crash: the reader checks every
    offset it follows against the file."""
    raw = open(PSF_FILE, "rb").read()
    H, _ = make_mat73.psf_arrays()
    p = tmp_path / "t.mat"
    for cut in list(range(520, 4096, 97)) + [len(raw) // 2, len(raw) - 1]:
        p.write_bytes(raw[:cut])
        try:
            with MatFile(str(p)) as m:
                m.read("H", np.float32)
        except (MatFileError, Unsupported, InvalidInput):
            pass
    rng = np.random.default_rng(5)
    for _ in range(300):
        b = bytearray(raw)
        pos = int(rng.integers(512, 8192))
        b[pos] ^= int(rng.integers(1, 256))
        p.write_bytes(bytes(b))
        try:
            with MatFile(str(p)) as m:
                for v in m.variables():
                    try:
                        inf = m.info(v)
                        if inf["numeric"] and inf["numel"] < 10_000_000:
                            m.read(v, np.float32)
                    except (MatFileError, Unsupported, InvalidInput):
                        pass
        except (MatFileError, Unsupported, InvalidInput):
            pass
