"""mlf_psf_load_mat: load(SI.psffile) (mesolf.m:20-21) of a -v7.3 PSF file
straight into device-resident tables.

The fixtures were written by libhdf5 in MATLAB's layout
(tests/golden/make_mat73.py): H single, Ht double (converted to single, as
the recon gateway converts double input), chunked with deflate (and shuffle),
in the library's default format and in its newest format.  The loaded PSF
must equal PSF.from_dense on the same arrays bit for bit (views, weights,
radii, and the volumes two FP32 iterations produce), and the oracle's tables
built from the same values.
"""
import math
import os

import numpy as np
import pytest

import mesolf_amd
from oracle import mesolf_oracle as O
from tests.golden import make_mat73
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def arrays():
    H, Ht = make_mat73.psf_arrays()
    H32, Ht32 = H.astype(np.float32), Ht.astype(np.float32)
    cp = O.compact_from_dense(H32.astype(np.float64), Ht32.astype(np.float64))
    return H32, Ht32, cp


@pytest.mark.parametrize("fname", ["mat73_psf.mat", "mat73_latest.mat"])
def test_loaded_psf_equals_from_dense_and_oracle(ctx, arrays, fname):
    H32, Ht32, cp = arrays
    print(f"case: load {fname}", flush=True)
    a = mesolf_amd.PSF.load_mat(ctx, os.path.join(GOLDEN, fname))
    b = mesolf_amd.PSF.from_dense(ctx, H32, Ht32)
    fi = a.file_info
    assert (fi["S"], fi["N"], fi["Z"]) == (45, 15, 2)
    if fname == "mat73_psf.mat":
        # |x1objspace(2) - x1objspace(1)| * 1e6, |x3objspace(2) - x3objspace(1)| * 1e6
        assert fi["x_pixel_um"] == pytest.approx(2.0, abs=1e-9)
        assert fi["z_pixel_um"] == pytest.approx(4.0, abs=1e-9)
    else:
        assert math.isnan(fi["x_pixel_um"]) and math.isnan(fi["z_pixel_um"])
    va, wa = a.views()
    vb, wb = b.views()
    assert va == vb == cp.views
    assert np.array_equal(wa, wb) and np.allclose(wa, cp.weights, rtol=1e-12, atol=0)
    ra, rb_ = a.radii(), b.radii()
    assert all(np.array_equal(x, y) for x, y in zip(ra, rb_))
    assert list(ra[0]) == cp.rf and list(ra[1]) == cp.rb

    LF = np.random.default_rng(31).poisson(40.0, (45, 60)).astype(np.float32)  # 3 x 4 lenslets
    vols = []
    for psf in (a, b):
        p = mesolf_amd.Patch(psf, 3, 4, precision="fp32")
        p.set_image(LF)
        p.init_volume()
        p.run(2)
        vols.append(p.get_volume(f64=True))
        p.close()
    assert np.array_equal(vols[0], vols[1])
    ref = O.rl_compact(LF.astype(np.float64), cp, 2)
    assert rel_l2(vols[0], ref) <= 1e-4
    a.close()
    b.close()


def test_load_refuses_missing_variables(ctx, tmp_path):
    # scipy's MATLAB-written v7.3 file holds only 'testdouble'
    import scipy
    path = os.path.join(os.path.dirname(scipy.__file__), "io", "matlab", "tests", "data",
                        "testhdf5_7.4_GLNX86.mat")
    if not os.path.exists(path):
        pytest.skip("scipy test data not installed")
    with pytest.raises(mesolf_amd.InvalidInput, match="'H' not found"):
        mesolf_amd.PSF.load_mat(ctx, path)
    with pytest.raises(mesolf_amd.MatFileError):
        mesolf_amd.PSF.load_mat(ctx, str(tmp_path / "absent.mat"))
