"""MEX operator on the reference's own test input at the production kernel width.

``src/cuda/mex_lfm_convolution_vec_test.m:61-72`` pads each ``vols_slice``
(376 x 376 x 77 single, the reference's only fixture) by 90 pixels at the end
of both image axes (``padarray(vols_slice, [90, 90], 'post')``) and projects
it with a 15 x 15-phase kernel.  The kernel that test used
(``psf_slice.mat``) is not in the reference, so a synthetic kernel of the
production width 241 stands in: the phase-space lobes of the synthetic PSF
family at z = 40 um (radius 8 lenslets = 241 px) with a 10 % random
non-separable perturbation.  Tolerance: relative L2 1e-5 against the FP64
oracle, per image.
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth
from oracle import mesolf_oracle as O
from tests.golden import load_vols_slice
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

KW = 241


def kernel_kw241(seed=4):
    rng = np.random.default_rng(seed)
    K = np.empty((KW, KW, 15, 15), np.float32, order="F")
    for a in range(15):
        for b in range(15):
            gx, gy = synth.view_profiles(a, b, 40)
            assert gx.size == KW
            K[:, :, a, b] = np.outer(gx, gy) * (1.0 + 0.1 * rng.random((KW, KW)))
    return K


@pytest.fixture(scope="module")
def fixture_input():
    vols = load_vols_slice()
    bm = np.zeros((376 + 90, 376 + 90, vols.shape[2]), np.float32, order="F")
    bm[:376, :376, :] = vols  # padarray(vols_slice, [90, 90], 'post')
    return bm


def test_mex_reference_fixture_kw241(fixture_input):
    bm = fixture_input
    K = kernel_kw241()
    got = mesolf_amd.mex_lfm_convolution_vec(bm, K)
    assert got.shape == bm.shape and got.dtype == np.float32
    nnz = np.count_nonzero(bm, axis=(0, 1))
    empty = np.flatnonzero(nnz == 0)
    assert np.all(got[:, :, empty] == 0)
    nonempty = np.flatnonzero(nnz)
    picks = sorted({int(np.argmax(nnz)), int(nonempty[len(nonempty) // 2]), int(nonempty[-1])})
    for n in picks:
        xs, ys = np.nonzero(bm[:, :, n])
        ref = O.mex_op_sparse(bm.shape[0], bm.shape[1], xs, ys, bm[xs, ys, n], K)
        err = rel_l2(got[:, :, n], ref)
        print(f"vols_slice image {n} ({nnz[n]} non-zeros), kw 241: rel L2 vs FP64 oracle {err:.2e}")
        assert err <= 1e-5, (n, err)


def test_mex_reference_fixture_accumulated_over_planes(fixture_input):
    """The reference test's loop (mex_lfm_convolution_vec_test.m:66-74): one
    MEX call per plane, summed; here as one segment call over all slices.
    Equal to the per-call sum added in the same order."""
    bm = fixture_input[:, :, :6]
    rng = np.random.default_rng(7)
    base = kernel_kw241()
    Kz = np.stack([base * np.float32(1.0 + 0.2 * z) for z in range(3)], axis=4)
    Kz = np.asfortranarray(Kz * (1.0 + 0.05 * rng.random(Kz.shape)).astype(np.float32))
    planes = [0, 1, 2, 0, 1, 2]
    got = mesolf_amd.lfm_convolution_segments(bm, planes, [0, 3, 6], Kz)
    for s in range(2):
        acc = np.zeros(bm.shape[:2], np.float32)
        for i in range(3 * s, 3 * s + 3):
            acc += mesolf_amd.mex_lfm_convolution_vec(bm[:, :, i], np.asfortranarray(Kz[..., planes[i]]))
        assert rel_l2(got[:, :, s], acc) <= 1e-6, s
