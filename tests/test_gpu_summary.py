"""GPU parity of the summary image and rank-1 background
(summary_image_generation_module.m:62-96, detrend mode 'pixelwise';
rank_1_factorization.m; compute_std_image.m) through the C ABI.

Tolerances: the background is FP64 on the device as in the oracle (rtol
1e-9); the detrended movie is held in FP32, so the summary image (normalised
to a max of 1) is compared to 1e-5 absolute.
"""
import numpy as np
import pytest

import mesolf_amd
from oracle import mesolf_oracle as O

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
BG_RTOL = 1e-9


def _movie(npix, T, seed):
    rng = np.random.default_rng(seed)
    base = rng.random(npix)[:, None] * 20 + 5
    trend = np.linspace(0, 1, T)[None, :] * rng.random(npix)[:, None] * 10
    spikes = (rng.random((npix, T)) < 0.02) * rng.random((npix, T)) * 50
    return (rng.poisson(base + trend + spikes)).astype(np.float32)


def _check(movie, h, w, *args):
    img, bs, bt = mesolf_amd.summary_image_pixelwise(movie, h, w, *args)
    ref_img, ref_s, ref_t = O.summary_pixelwise(movie.reshape(h * w, -1, order="F"), *args)
    assert img.shape == (h, w) and bs.shape == (h, w) and bt.shape == (1, movie.shape[-1])
    ri = ref_img.reshape(h, w, order="F")
    assert np.array_equal(np.isnan(img), np.isnan(ri))
    ok = ~np.isnan(ri)
    assert np.max(np.abs(img[ok] - ri[ok]), initial=0.0) <= IMG_TOL
    assert np.allclose(bs, ref_s.reshape(h, w, order="F"), rtol=BG_RTOL, atol=0)
    assert np.allclose(bt[0], ref_t, rtol=BG_RTOL, atol=1e-300)
    return img


# (9, 4, 60): windows of 100 and 20 frames clipped at both ends of a 60-frame movie
@pytest.mark.parametrize("h,w,T", [(20, 17, 150), (70, 65, 130), (9, 4, 60)])
def test_summary_defaults_vs_oracle(h, w, T):
    movie = _movie(h * w, T, seed=h + T)
    img = _check(movie, h, w, 100, 20, 3, 5)
    assert np.nanmax(img) == 1.0


@pytest.mark.parametrize("w1,w2,p,it", [(4, 2, 1, 1), (7, 5, 2, 3), (31, 10, 3, 2), (2, 1, 3, 1)])
def test_summary_windows_vs_oracle(w1, w2, p, it):
    h, w, T = 13, 11, 64
    _check(_movie(h * w, T, seed=w1), h, w, w1, w2, p, it)


def test_summary_module_dropin_and_errors():
    h, w, T = 16, 12, 90
    movie = _movie(h * w, T, seed=3).reshape(h, w, T, order="F")
    SI = {"detrend_mode": "pixelwise", "pixelwise_window_1": 30, "pixelwise_window_2": 6,
          "pixelwise_poly_index": 3, "bg_iter": 4}
    got = mesolf_amd.summary_image_generation_module(movie, {"size": [h, w]}, SI)
    want = mesolf_amd.summary_image_pixelwise(movie, h, w, 30, 6, 3, 4)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    with pytest.raises(mesolf_amd.Unsupported):
        mesolf_amd.summary_image_generation_module(movie, {"size": [h, w]}, dict(SI, detrend_mode="global"))
    with pytest.raises(mesolf_amd.InvalidInput):
        mesolf_amd.summary_image_generation_module(movie, {"size": [h, w]}, dict(SI, detrend_mode="x"))
    for bad in ((0, 20, 3, 5), (100, 0, 3, 5), (100, 20, -1, 5), (100, 20, 3, 0)):
        with pytest.raises(mesolf_amd.InvalidInput):
            mesolf_amd.summary_image_pixelwise(movie, h, w, *bad)
