"""GPU parity of the post-reconstruction step (LFM_reconstruction_module.m:45-53,
60-62 and the crop of mesolf.m:101-102) through the C ABI.

The device path computes in FP64 on the FP32 volume in MATLAB's evaluation
order, as the oracle does on the same values, so stacks and images are compared
bit for bit (NaN where both are NaN).
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


def _psf(ctx, wl):
    tabs = tables.synthetic_compact(wl)
    return mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def _volume(wl, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((wl.I1, wl.I2, wl.Z)).astype(np.float32)
    X *= (10.0 ** rng.uniform(-3, 3, wl.Z)).astype(np.float32)  # planes of unlike scale
    return X


@pytest.mark.parametrize("wl", [synth.TINY, synth.SMALL, synth.Workload("odd_37x21", 37, 21, (-20, 0, 12), 1, 1)],
                         ids=lambda w: w.name)
def test_postprocess_bitwise_vs_oracle(ctx, wl):
    psf = _psf(ctx, wl)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision="bf16")
    X = _volume(wl, 11)
    p.set_volume(X)
    ref_full, ref_tiff = O.post_recon(X.astype(np.float64))
    out, tiff = p.postprocess(0)
    assert _same(out, ref_full) and _same(tiff, ref_tiff)
    for v in range(1, wl.Z + 1, 2):
        lo, hi = O.valid_planes(wl.Z, v)
        out, none = p.postprocess(v, tiff=False)
        assert none is None and _same(out, ref_full[:, :, lo:hi])
    none, tiff = p.postprocess(1, stack=False)
    assert none is None and _same(tiff, ref_tiff)
    # the volume is not changed
    assert np.array_equal(p.get_volume(), X)
    p.close()
    psf.close()


def test_postprocess_zero_plane_nan_and_negative(ctx):
    wl = synth.SMALL
    psf = _psf(ctx, wl)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
    X = _volume(wl, 12)
    X[:, :, 2] = 0.0                 # 0/0 -> NaN plane, image 0
    X[5, 7, 1] = np.nan              # ignored by the maxima
    X[3, 3, 3] = -2.0                # negative value -> image 0
    X[:, :, 4] = -X[:, :, 4]         # all-negative plane: max < 0
    p.set_volume(X)
    ref = O.post_recon(X.astype(np.float64), 3)
    got = p.postprocess(3)
    assert _same(got[0], ref[0]) and _same(got[1], ref[1])
    assert np.all(np.isnan(got[0][:, :, 1]))  # planes 2..4 of 5 kept: the zero plane is the second
    p.close()
    psf.close()


def test_postprocess_after_reconstruction(ctx):
    wl = synth.SMALL
    psf = _psf(ctx, wl)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision="bf16")
    LF = np.random.default_rng(4).poisson(60.0, (wl.I1, wl.I2)).astype(np.float32) + 1
    p.set_image(LF)
    p.init_volume()
    p.run(1, views_limit=12)
    X = p.get_volume()
    ref = O.post_recon(X.astype(np.float64), 5)
    got = p.postprocess(5)
    assert _same(got[0], ref[0]) and _same(got[1], ref[1])
    assert np.nanmax(got[0]) == 1.0 and got[1].max() == 65535
    p.close()
    psf.close()


def test_postprocess_errors_and_timing(ctx):
    wl = synth.TINY
    psf = _psf(ctx, wl)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
    with pytest.raises(mesolf_amd.InvalidInput):
        p.postprocess(1)  # volume not set
    p.set_volume(_volume(wl, 3))
    for bad in (2, 5, -1):  # even, beyond the 3 planes, negative
        with pytest.raises(mesolf_amd.InvalidInput):
            p.postprocess(bad)
    p.set_timing(True)
    p.kernel_times()
    p.postprocess(3)
    kt = p.kernel_times()
    assert kt["postprocess"][1] == 1 and kt["postprocess"][0] > 0
    assert kt["forward"][1] == 0
    p.close()
    psf.close()


def test_module_dropin(ctx):
    wl = synth.Workload("dense", 4, 5, (-4, 0, 4), 2, 2)
    Hps = synth.dense_phase_space(wl)
    H = O.inverse_phasespace(Hps)
    Ht = synth.adjoint_ht(Hps)
    LF = np.random.default_rng(8).poisson(80.0, (wl.I1, wl.I2)).astype(np.float32)
    SI = {"recon_max_iter": 2, "gpu_ids": [1], "recon_mode": "phase_space",
          "vessel_mask_enable": False, "worker_ix": 3}
    stack, tiff = mesolf_amd.LFM_reconstruction_module(LF, H, Ht, SI)
    assert stack.shape == (wl.I1, wl.I2, wl.Z) and tiff.dtype == np.uint16
    cp = O.compact_from_dense(H, Ht)
    ref_stack, ref_tiff = O.post_recon(O.rl_compact(LF.astype(np.float64), cp, 2))
    assert rel_l2(stack, ref_stack) <= 1e-4
    assert np.abs(tiff.astype(np.int64) - ref_tiff.astype(np.int64)).max() <= 8
    mid, _ = mesolf_amd.LFM_reconstruction_module(LF, H, Ht, SI, valid_range=1)
    assert rel_l2(mid, ref_stack[:, :, 1:2]) <= 1e-4
    for bad, exc in ((dict(SI, vessel_mask_enable=True), mesolf_amd.Unsupported),
                     (dict(SI, recon_mode="phase_space_peeling"), mesolf_amd.Unsupported),
                     (dict(SI, recon_mode="other"), mesolf_amd.InvalidInput),
                     (dict(SI, gpu_ids=[]), mesolf_amd.Unsupported)):
        with pytest.raises(exc):
            mesolf_amd.LFM_reconstruction_module(LF, H, Ht, bad)
