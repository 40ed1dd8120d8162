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
from mesolf_amd import lib as L
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
    g = mesolf_amd.summary_image_generation_module(movie, {"size": [h, w]}, dict(SI, detrend_mode="global"))
    for a, b in zip(g, mesolf_amd.summary_image_global(movie, h, w, 4)):
        assert np.array_equal(a, b)
    with pytest.raises(mesolf_amd.InvalidInput):
        mesolf_amd.summary_image_generation_module(movie, {"size": [h, w]}, dict(SI, detrend_mode="x"))
    for bad in ((0, 20, 3, 5), (100, 0, 3, 5), (100, 20, -1, 5), (100, 20, 3, 0)):
        with pytest.raises(mesolf_amd.InvalidInput):
            mesolf_amd.summary_image_pixelwise(movie, h, w, *bad)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library loaded
    (test plumbing for the device-pointer entry points)."""

    def __init__(self):
        import ctypes as C
        L.load_library()
        path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line)
        self.C = C
        self.h = C.CDLL(path)
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.h.hipMalloc(self.C.byref(p), nbytes) == 0
        return p.value

    def put(self, a):
        d = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def get(self, d, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, d):
        self.h.hipFree(d)


def test_carpet_crop_then_summary_on_device():
    """Carpet patches of a sensor movie in HBM: each patch movie cut on the
    device equals the MATLAB slice bit for bit, and its summary image computed
    from device pointers equals the host entry point's."""
    import ctypes as C

    H, W, T = 200, 170, 40
    movie = np.asfortranarray(_movie(H * W, T, seed=9).reshape(H, W, T, order="F"))
    patches = mesolf_amd.carpet_patch_info(H, W, [2, 2])
    assert len(patches) == 4
    ctx = mesolf_amd.Context(0)
    hip = _Hip()
    d_movie = hip.put(movie.ravel(order="F").copy())
    bufs = [d_movie]
    try:
        for pi in patches:
            (t, l), (b, r) = pi["location"]
            h, w = pi["size"]
            d_out = hip.alloc(4 * h * w * T)
            d_img = hip.alloc(8 * h * w)
            bufs += [d_out, d_img]
            mesolf_amd.crop_movie_device(ctx, d_movie, H, W, T, pi, d_out)
            ctx.synchronize()
            want = movie[t - 1:b, l - 1:r, :]
            assert np.array_equal(hip.get(d_out, h * w * T, np.float32), want.ravel(order="F"))
            L._check(ctx._lib.mlf_summary_image_device(ctx.handle, C.c_void_p(d_out), h, w, T, 30, 6, 3, 2,
                                                       C.c_void_p(d_img), None, None))
            img = hip.get(d_img, h * w, np.float64).reshape(h, w, order="F")
            host = mesolf_amd.summary_image_pixelwise(want, h, w, 30, 6, 3, 2)[0]
            assert np.array_equal(img, host)
        with pytest.raises(mesolf_amd.InvalidInput):
            mesolf_amd.crop_movie_device(ctx, d_movie, H, W, T, {"location": [[150, 1], [215, 60]]}, bufs[1])
    finally:
        for d in bufs:
            hip.free(d)
        ctx.close()



@pytest.mark.parametrize("H,W,n,xc,yc,dx", [(61, 47, 3, 10.3, 7.6, 14.2), (40, 33, 1, 3.5, 10.25, 7.5),
                                             (90, 75, 2, 44.0, 45.0, 15.0)])
def test_rectify_vs_oracle(H, W, n, xc, yc, dx):
    """img_rectify_mod against scipy's linear interpolation with linear
    extrapolation (FP64 on both sides, single output: within 2e-7 relative)."""
    rng = np.random.default_rng(H + n)
    frames = (rng.random((H, W, n)) * 1000).astype(np.float32)
    got = mesolf_amd.img_rectify_mod(frames, xc, yc, 15, dx)
    want = O.rectify(frames.astype(np.float64), xc, yc, 15, dx)
    assert got.shape == want.shape and got.shape[0] % 15 == 0 and got.shape[1] % 15 == 0
    assert np.max(np.abs(got - want)) <= 2e-7 * np.max(np.abs(want))
    one = mesolf_amd.img_rectify_mod(frames[:, :, 0], xc, yc, 15, dx)
    assert np.array_equal(one, got[:, :, 0])
    with pytest.raises(mesolf_amd.InvalidInput):
        mesolf_amd.img_rectify_mod(frames, W + 5.0, yc, 15, dx)  # centre outside the image



@pytest.mark.parametrize("h,w,T,seed", [(20, 17, 150, 1), (70, 65, 130, 2), (13, 11, 64, 3)])
def test_summary_global_vs_oracle(h, w, T, seed):
    """'global' detrend: the std image with the rank-1 background (the square
    root's argument stays above 0.03 for these movies, so no NaN can flip)."""
    rng = np.random.default_rng(seed)
    movie = (rng.random((h * w, T)) * 50 + rng.random((h * w, 1)) * 20).astype(np.float32)
    img, bs, bt = mesolf_amd.summary_image_global(movie, h, w, 5)
    ref_img, ref_s, ref_t = O.summary_global(movie, 5)
    assert not np.isnan(img).any()
    assert np.max(np.abs(img - ref_img.reshape(h, w, order="F"))) <= IMG_TOL
    assert np.allclose(bs, ref_s.reshape(h, w, order="F"), rtol=BG_RTOL, atol=0)
    assert np.allclose(bt[0], ref_t, rtol=BG_RTOL, atol=0)
