"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerances (relative L2, FP32 device arithmetic vs FP64 oracle):
  forward projections  <= 1e-5
  volumes              <= 1e-4 (after the stated iterations)
  tables / view order  bit-exact
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O
from tests.helpers import oracle_compact_from_abi, rel_l2, synthetic_problem

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
VOL_TOL = 1e-4
# BF16 operands (FP32 accumulation, FP32 update): looser per-step bounds
TOL = {"fp32": (FWD_TOL, VOL_TOL), "bf16": (4e-3, 2e-2), "fp16": (1e-3, 5e-3)}


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def small_problem():
    return synthetic_problem(synth.SMALL, seed=3)


def _psf(ctx, tabs):
    return mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("wl", [synth.TINY, synth.SMALL, synth.Workload("odd", 37, 21, (-20, 0, 12), 1, 1)],
                         ids=lambda w: w.name)
def test_forward_projection_parity(ctx, wl, prec):
    tabs = tables.synthetic_compact(wl)
    cp = oracle_compact_from_abi(wl, tabs)
    psf = _psf(ctx, tabs)
    assert psf.nviews == 177
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
    rng = np.random.default_rng(11)
    X = rng.random((wl.I1, wl.I2, wl.Z)).astype(np.float32)
    p.set_volume(X)
    Xl = O.to_lenslet(X.astype(np.float64), 15)
    for k in (0, 1, 57, 176):
        y = p.forward_view(k)
        ref = O.forward_compact(Xl, cp.fwd[k], cp.rf)
        assert rel_l2(y, ref) <= TOL[prec][0], (k, rel_l2(y, ref))


WIDE = synth.Workload("wide_37x21", 37, 21, (-100, -84, -40, 0), 1, 1)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_forward_projection_parity_wide_radii(ctx, prec):
    """Planes of radius 17, 15, 8 and 2 (K = 35, 31, 17, 5 taps along an axis):
    every N-tile count of the 16x16 path and both ends of the 32x32 path, on a
    grid smaller than the widest kernel."""
    views = synth.aperture_views()[::30]
    tabs = tables.synthetic_compact(WIDE, views=views)
    assert sorted(int(r) for r in tabs["rf"]) == [2, 8, 15, 17]
    cp = oracle_compact_from_abi(WIDE, tabs)
    psf = _psf(ctx, tabs)
    p = mesolf_amd.Patch(psf, WIDE.n1, WIDE.n2, precision=prec)
    rng = np.random.default_rng(12)
    X = rng.random((WIDE.I1, WIDE.I2, WIDE.Z)).astype(np.float32)
    p.set_volume(X)
    Xl = O.to_lenslet(X.astype(np.float64), 15)
    for k in range(psf.nviews):
        y = p.forward_view(k)
        ref = O.forward_compact(Xl, cp.fwd[k], cp.rf)
        assert rel_l2(y, ref) <= TOL[prec][0], (k, rel_l2(y, ref))


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_view_steps_parity_wide_radii(ctx, prec):
    """Forward, ratio and backprojection + update of all six views of the
    wide-radius PSF (windows of up to 50 x 50 lenslets, up to 44 K-steps)."""
    tabs, cp, truth, LF = synthetic_problem(WIDE, seed=5, views=synth.aperture_views()[::30])
    p = mesolf_amd.Patch(_psf(ctx, tabs), WIDE.n1, WIDE.n2, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(1, filter=False)
    got = p.get_volume(f64=True)
    ref = O.rl_compact(LF.astype(np.float64), cp, 1, filter_planes=False)
    assert rel_l2(got, ref) <= TOL[prec][1] / 4


def test_volume_roundtrip_layout(ctx):
    wl = synth.TINY
    psf = _psf(ctx, tables.synthetic_compact(wl))
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
    X = np.random.default_rng(2).random((wl.I1, wl.I2, wl.Z)).astype(np.float32)
    p.set_volume(X)
    assert np.array_equal(p.get_volume(), X)
    assert np.array_equal(p.get_volume(f64=True), X.astype(np.float64))


def _edge_volume(wl, seed):
    """FP32 volume at the rounding edges of a 16-bit split of X + 2^15: exact
    ties of the upper half, mantissas that carry into the next binade, lower
    halves just below the tie, zeros, and normal values from 1e-30 to 1e30."""
    rng = np.random.default_rng(seed)
    n = wl.I1 * wl.I2 * wl.Z
    bits = rng.integers(0x0D800000, 0x71000000, n, dtype=np.uint64).astype(np.uint32)
    kind = rng.integers(0, 5, n)
    bits = np.where(kind == 0, (bits & 0xFFFF0000) | 0x8000, bits)
    bits = np.where(kind == 1, bits | 0x007FFFFF, bits)
    bits = np.where(kind == 2, (bits & 0xFFFF0000) | 0x7FFF, bits)
    bits = np.where(kind == 3, 0, bits).astype(np.uint32)
    return bits.view(np.float32).reshape((wl.I1, wl.I2, wl.Z), order="F")


def _decayed(X, w):
    """X after view steps with every ratio 0: X <- fl((1 - w_k) X) per view
    (recon:130-137 with the backprojected ratio 0)."""
    x = X.copy()
    for wk in w:
        x = (np.float32(1) - np.float32(wk)) * x
    return x


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_view_steps_keep_the_fp32_volume_exact(ctx, prec):
    """With a zero image every ratio is 0, so each view step multiplies X by
    (1 - w) in FP32, one rounding per step.  The BF16 path holds the volume
    between view steps as the two 16-bit halves of X + 2^15; the result must
    still be the FP32 arithmetic bit for bit, at the split's rounding edges,
    alone and in a batch (graph-captured view loop)."""
    wl = synth.TINY
    psf = _psf(ctx, tables.synthetic_compact(wl))
    _, w = psf.views()
    X = _edge_volume(wl, 11)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
    p.set_image(np.zeros((wl.I1, wl.I2), np.float32))
    p.set_volume(X)
    p.run(1, views_limit=3, filter=False)
    got = p.get_volume()
    want = _decayed(X, w[:3])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    p.run(1, views_limit=2, filter=False)  # continues from the merged volume
    want = _decayed(want, w[:2])
    assert np.array_equal(p.get_volume().view(np.uint32), want.view(np.uint32))

    X2 = _edge_volume(wl, 12)
    q = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
    q.set_image(np.zeros((wl.I1, wl.I2), np.float32))
    q.set_volume(X2)
    b = mesolf_amd.Batch(psf, 2, precision=prec)
    b.attach(q)
    b.run(2, filter=False)
    want = _decayed(_decayed(X2, w), w)
    assert np.array_equal(q.get_volume().view(np.uint32), want.view(np.uint32))
    b.detach(q)
    b.close()
    for x in (q, p):
        x.close()
    psf.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_view_steps_parity(ctx, small_problem, prec):
    tabs, cp, truth, LF = small_problem
    wl = synth.SMALL
    p = mesolf_amd.Patch(_psf(ctx, tabs), wl.n1, wl.n2, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(1, views_limit=5, filter=False)
    got = p.get_volume(f64=True)
    ref = O.rl_compact(LF.astype(np.float64), cp, 1, views_limit=5, filter_planes=False)
    assert rel_l2(got, ref) <= TOL[prec][1] / 4


def test_sample_filter_parity(ctx):
    wl = synth.SMALL
    p = mesolf_amd.Patch(_psf(ctx, tables.synthetic_compact(wl)), wl.n1, wl.n2)
    X = np.random.default_rng(4).random((wl.I1, wl.I2, wl.Z)).astype(np.float32)
    p.set_volume(X)
    p.sample_filter()
    got = p.get_volume(f64=True)
    bank = O.filter_bank(wl.I1, wl.I2, wl.Z, 15)
    ref = O.sample_filter_literal(X.astype(np.float64), bank)
    assert rel_l2(got, ref) <= 1e-6
    # energy of every plane preserved
    assert np.allclose(got.sum(axis=(0, 1)), X.astype(np.float64).sum(axis=(0, 1)), rtol=1e-6)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_rl_parity_small_two_iterations(ctx, small_problem, prec):
    tabs, cp, truth, LF = small_problem
    wl = synth.SMALL
    p = mesolf_amd.Patch(_psf(ctx, tabs), wl.n1, wl.n2, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(2)
    got = p.get_volume(f64=True)
    ref = O.rl_compact(LF.astype(np.float64), cp, 2)
    err = rel_l2(got, ref)
    print(f"{prec}: 2 iterations rel L2 vs FP64 = {err:.3e}")
    assert err <= TOL[prec][1]


def test_reference_kernels_match_mfma_fp32(ctx, small_problem, monkeypatch):
    """The FP32 reference kernels and the FP32 MFMA kernels agree."""
    tabs, cp, truth, LF = small_problem
    wl = synth.SMALL
    psf = _psf(ctx, tabs)
    outs = []
    for ref in ("1", "0"):
        monkeypatch.setenv("MLF_REFERENCE_KERNELS", ref)
        p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
        p.set_image(LF)
        p.init_volume()
        p.run(1, views_limit=20)
        outs.append(p.get_volume(f64=True))
    assert rel_l2(outs[1], outs[0]) <= 1e-5


def test_determinism_bitwise(ctx, small_problem):
    tabs, cp, truth, LF = small_problem
    wl = synth.SMALL
    psf = _psf(ctx, tabs)
    outs = []
    for _ in range(2):
        p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
        p.set_image(LF)
        p.init_volume()
        p.run(1, views_limit=40)
        outs.append(p.get_volume())
    assert np.array_equal(outs[0], outs[1])


def test_kernel_timing_stride(ctx, small_problem):
    """Timing every 4th view step brackets 4 of 16 view steps; timing changes
    no result."""
    tabs, cp, truth, LF = small_problem
    wl = synth.SMALL
    psf = _psf(ctx, tabs)
    outs = []
    for stride in (0, 1, 4):
        p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision="bf16")
        p.set_image(LF)
        p.init_volume()
        p.set_timing(stride > 0, stride=max(stride, 1))
        p.kernel_times()
        p.run(1, views_limit=16, filter=False)
        kt = p.kernel_times()
        p.set_timing(False)
        want = 0 if stride == 0 else 16 // stride
        for c in ("forward", "ratio", "backward_update"):
            assert kt[c][1] == want, (stride, c, kt[c])
            assert (kt[c][0] > 0) == (want > 0)
        outs.append(p.get_volume())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    with pytest.raises(ValueError):
        p.set_timing(True, stride=0)


def test_project_all_parity(ctx):
    wl = synth.TINY
    tabs = tables.synthetic_compact(wl)
    cp = oracle_compact_from_abi(wl, tabs)
    p = mesolf_amd.Patch(_psf(ctx, tabs), wl.n1, wl.n2)
    truth = synth.sphere_truth(wl, np.random.default_rng(5)).astype(np.float32)
    p.set_volume(truth)
    got = p.project_all()
    Xl = O.to_lenslet(truth.astype(np.float64), 15)
    ref = np.zeros((wl.I1, wl.I2))
    for k, (u, v) in enumerate(cp.views):
        ref[u::15, v::15] = O.forward_compact(Xl, cp.fwd[k], cp.rf)
    assert rel_l2(got, ref) <= FWD_TOL


def test_dense_psf_preparation(ctx):
    """PSF.from_dense reproduces the oracle's view order, weights and radii."""
    wl = synth.TINY
    Hps = synth.dense_phase_space(wl)
    H = O.inverse_phasespace(Hps)
    Ht = synth.adjoint_ht(Hps)
    psf = mesolf_amd.PSF.from_dense(ctx, H, Ht)
    cp = O.compact_from_dense(H, Ht)
    views, w = psf.views()
    assert views == cp.views                     # bit-exact order and mask
    assert np.allclose(w, cp.weights, rtol=1e-12)
    rf, rb = psf.radii()
    assert list(rf) == cp.rf and list(rb) == cp.rb


def test_recon_dropin_dense_matches_oracle(ctx):
    wl = synth.Workload("dense", 4, 5, (-4, 0, 4), 2, 2)
    Hps = synth.dense_phase_space(wl)
    H = O.inverse_phasespace(Hps)
    Ht = synth.adjoint_ht(Hps)
    rng = np.random.default_rng(8)
    LF = rng.poisson(80.0, (wl.I1, wl.I2)).astype(np.float32)
    st = {}
    got = mesolf_amd.recon_theory_phase_space_LFM_with_filter(LF, H, Ht, 2, True, 1, stats=st)
    assert got.dtype == np.float64 and got.shape == (wl.I1, wl.I2, wl.Z)
    cp = O.compact_from_dense(H, Ht)
    ref = O.rl_compact(LF.astype(np.float64), cp, 2)
    assert rel_l2(got, ref) <= VOL_TOL
    assert st["iterations"] == 2 and st["views_per_iteration"] == 177


@pytest.mark.parametrize("W,H,n,kw", [(40, 33, 1, 7), (64, 70, 3, 31), (50, 45, 2, 8),
                                      (30, 20, 1, 101)])
def test_mex_operator_parity(W, H, n, kw):
    rng = np.random.default_rng(W + kw)
    b = rng.random((W, H, n)).astype(np.float32)
    K = rng.random((kw, kw, 15, 15)).astype(np.float32)
    got = mesolf_amd.mex_lfm_convolution_vec(b if n > 1 else b[:, :, 0], K)
    ref = O.mex_op(b.astype(np.float64), K.astype(np.float64))
    if n == 1:
        ref = ref[:, :, 0]
    assert got.shape == (b.shape if n > 1 else b.shape[:2])
    assert rel_l2(got, ref) <= 1e-5


def test_segment_projection_matches_summed_mex_calls():
    """One call for all segments and planes == per-plane MEX calls summed per
    segment (gen_seed_library_local_with_ring_bg_same_size.m:263-318)."""
    rng = np.random.default_rng(12)
    W, H, Z, kw = 45, 38, 4, 9
    K = rng.random((kw, kw, 15, 15, Z)).astype(np.float32)
    planes = [0, 2, 3, 1, 2, 0, 1]
    seg_ptr = [0, 3, 3, 5, 7]               # segment 1 has no slices
    sl = rng.random((W, H, len(planes))).astype(np.float32)
    got = mesolf_amd.lfm_convolution_segments(sl, planes, seg_ptr, K)
    assert got.shape == (W, H, 4)
    for sgi in range(4):
        ref = np.zeros((W, H))
        for i in range(seg_ptr[sgi], seg_ptr[sgi + 1]):
            ref += O.mex_op(sl[:, :, i].astype(np.float64), K[..., planes[i]].astype(np.float64))
        if seg_ptr[sgi] == seg_ptr[sgi + 1]:
            assert np.all(got[:, :, sgi] == 0)
        else:
            assert rel_l2(got[:, :, sgi], ref) <= 1e-5
        # and equal to the drop-in MEX called per slice, added in order
        acc = np.zeros((W, H), np.float32)
        for i in range(seg_ptr[sgi], seg_ptr[sgi + 1]):
            acc += mesolf_amd.mex_lfm_convolution_vec(sl[:, :, i], K[..., planes[i]])
        assert np.array_equal(got[:, :, sgi], acc)


def test_mex_operator_wide_kernel_is_a_shift():
    """kw = 401 (> the reference's 341 limit): a single-tap kernel shifts."""
    kw, h = 401, 200
    rng = np.random.default_rng(12)
    b = rng.random((300, 260)).astype(np.float32)
    K = np.zeros((kw, kw, 15, 15), np.float32)
    K[h + 3, h - 2] = 1.0
    got = mesolf_amd.mex_lfm_convolution_vec(b, K)
    exp = np.zeros_like(b)
    exp[:-3, 2:] = b[3:, :-2]          # out[x, y] = in[x + 3, y - 2]
    assert np.array_equal(got, exp)


@pytest.fixture(scope="module")
def cfg1_reference():
    wl = synth.CFG1
    tabs, cp, truth, LF = synthetic_problem(wl, seed=1)
    inter = {}
    ref = O.rl_compact(LF.astype(np.float64), cp, 8,
                       callback=lambda it, X: inter.__setitem__(it + 1, X.copy()))
    return tabs, LF, ref, inter


@pytest.mark.slow
@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_rl_parity_cfg1_eight_iterations(ctx, cfg1_reference, prec):
    """cfg1 shapes (33x33 lenslets, 11 planes), 8 RL iterations."""
    wl = synth.CFG1
    tabs, LF, ref, _ = cfg1_reference
    p = mesolf_amd.Patch(_psf(ctx, tabs), wl.n1, wl.n2, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(8)
    got = p.get_volume(f64=True)
    err = rel_l2(got, ref)
    print(f"cfg1 8-iteration {prec} rel L2 vs FP64 oracle: {err:.3e}")
    assert err <= {"fp32": VOL_TOL, "bf16": 1e-3, "fp16": 1e-3}[prec]


@pytest.mark.slow
def test_precision_error_after_1_4_8_iterations(ctx, cfg1_reference):
    """Relative L2 error vs the FP64 oracle after 1, 4 and 8 iterations for
    every precision, side by side (33x33 lenslets, 11 planes)."""
    wl = synth.CFG1
    tabs, LF, _, inter = cfg1_reference
    psf = _psf(ctx, tabs)
    table = {}
    for prec in ("fp32", "bf16", "fp16"):
        p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
        p.set_image(LF)
        p.init_volume()
        done = 0
        for it in (1, 4, 8):
            p.run(it - done)
            done = it
            table[(prec, it)] = rel_l2(p.get_volume(f64=True), inter[it])
        p.close()
    print("rel L2 vs FP64 oracle   1 it      4 it      8 it")
    for prec in ("fp32", "bf16", "fp16"):
        print(f"  {prec:5s}              " + "  ".join(f"{table[(prec, it)]:.2e}" for it in (1, 4, 8)))
    for it in (1, 4, 8):
        assert table[("fp32", it)] <= VOL_TOL
        assert table[("bf16", it)] <= 1e-3
        assert table[("fp16", it)] <= table[("bf16", it)]  # 3 more mantissa bits


def test_handle_lifetimes_in_any_order():
    """Destroying a parent handle first only defers its release."""
    wl = synth.TINY
    tabs = tables.synthetic_compact(wl)
    c = mesolf_amd.Context(0)
    psf = _psf(c, tabs)
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2)
    c.close()
    psf.close()
    X = np.ones((wl.I1, wl.I2, wl.Z), np.float32)
    p.set_volume(X)
    assert np.isfinite(p.forward_view(0)).all()
    p.close()


@pytest.mark.parametrize("kw", [9, 8])
def test_segment_projection_order_and_rounds(kw):
    """Segments whose slices repeat a plane (several launches per plane on
    the MFMA path) or list planes out of order: each segment equals the sum of
    its per-slice MEX calls added in ascending plane order (slice order within
    a plane), bit for bit, and the FP64 oracle within 1e-5."""
    rng = np.random.default_rng(21 + kw)
    W, H, Z = 50, 41, 3
    K = rng.random((kw, kw, 15, 15, Z)).astype(np.float32)
    planes = [2, 0, 2, 1, 1, 1, 0]
    seg_ptr = [0, 4, 6, 6, 7]                # segment 2 is empty
    sl = rng.random((W, H, len(planes))).astype(np.float32)
    got = mesolf_amd.lfm_convolution_segments(sl, planes, seg_ptr, K)
    for sgi in range(4):
        idx = list(range(seg_ptr[sgi], seg_ptr[sgi + 1]))
        idx.sort(key=lambda i: planes[i])    # stable: slice order within a plane
        acc = np.zeros((W, H), np.float32)
        ref = np.zeros((W, H))
        for i in idx:
            acc += mesolf_amd.mex_lfm_convolution_vec(sl[:, :, i], np.asfortranarray(K[..., planes[i]]))
            ref += O.mex_op(sl[:, :, i].astype(np.float64), K[..., planes[i]].astype(np.float64))
        assert np.array_equal(got[:, :, sgi], acc), sgi
        if idx:
            assert rel_l2(got[:, :, sgi], ref) <= 1e-5
        else:
            assert np.all(got[:, :, sgi] == 0)
