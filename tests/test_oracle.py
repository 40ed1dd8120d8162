"""Pins for the FP64 oracle (CPU only).

Each oracle function is checked against something independent of it: a
brute-force loop, a library routine (scipy), a closed form derived separately,
or an algebraic identity (adjointness, energy conservation, round trips).
"""
import numpy as np
import pytest

from mesolf_amd import synth as S
from oracle import mesolf_oracle as O


# --------------------------------------------------------------------- spiral
def _ring_walk(N):
    """Independent spiral: walk each ring clockwise starting at its top-left
    corner, first down the first column, then along the last row, up the last
    column and back along the first row (recon:171-175 in words)."""
    out = []
    lo, hi = 0, N - 1
    while lo <= hi:
        if lo == hi:
            out.append((lo, lo))
            break
        for i in range(lo, hi + 1):
            out.append((i, lo))
        for j in range(lo + 1, hi + 1):
            out.append((hi, j))
        for i in range(hi - 1, lo - 1, -1):
            out.append((i, hi))
        for j in range(hi - 1, lo, -1):
            out.append((lo, j))
        lo, hi = lo + 1, hi - 1
    return out


@pytest.mark.parametrize("N", [1, 3, 4, 5, 7, 15])
def test_spiral_matches_ring_walk(N):
    i1, i2 = O.gen_spiral(N)
    got = [(int(a) - 1, int(b) - 1) for a, b in zip(i1, i2)]
    assert got == _ring_walk(N)
    assert sorted(got) == [(a, b) for a in range(N) for b in range(N)]


def test_spiral_known_entries_n15():
    i1, i2 = O.gen_spiral(15)
    pairs = list(zip(i1.tolist(), i2.tolist()))
    assert pairs[:3] == [(1, 1), (2, 1), (3, 1)]
    assert pairs[14:17] == [(15, 1), (15, 2), (15, 3)]
    assert pairs[-5:] == [(9, 9), (8, 9), (7, 9), (7, 8), (8, 8)]
    w = np.zeros((15, 15))
    for (u, v) in S.aperture_views():
        w[u, v] = 1
    act = [(a, b) for a, b in pairs if w[a - 1, b - 1] != 0]
    assert act[:6] == [(6, 1), (7, 1), (8, 1), (9, 1), (10, 1), (15, 6)]


# -------------------------------------------------------------- phase space
def _random_psf(rng, S_, N, Z):
    return rng.random((S_, S_, N, N, Z))


@pytest.mark.parametrize("S_,N", [(45, 3), (47, 3), (75, 5), (75, 15)])
def test_phasespace_literal_equals_closed_form(S_, N):
    rng = np.random.default_rng(S_ + N)
    H = _random_psf(rng, S_, N, 2)
    a, _ = O.gen_phasespace_literal(H, H)
    b = O.phase_space_closed_form(H)
    assert a.shape == b.shape
    assert np.array_equal(a, b)


def test_phasespace_crop_rule():
    assert O.phase_space_crop(75, 15) == (75, 0)
    # 80 - mod(80-15, 30) = 75, round(40) - fix(37.5) = 3 (1-based)
    assert O.phase_space_crop(80, 15) == (75, 2)
    assert O.phase_space_crop(525, 15) == (525, 0)


def test_inverse_phasespace_roundtrip():
    Hps = S.dense_phase_space(S.TINY)
    H = O.inverse_phasespace(Hps)
    back, _ = O.gen_phasespace_literal(H, H)
    assert np.array_equal(back, Hps)


# ----------------------------------------------------------------- weights
def test_disc_aperture_has_177_views():
    Hps = S.dense_phase_space(S.TINY)
    w = O.view_weights(O.normalize_depth(Hps))
    act = w != 0
    assert act.sum() == 177
    ref = np.zeros((15, 15), bool)
    for (u, v) in S.aperture_views():
        ref[u, v] = True
    assert np.array_equal(act, ref)
    # synthetic kernels all sum to 1 -> every active view has weight 0.3
    assert np.allclose(w[act], 0.3)


def test_weights_threshold_by_hand():
    s = np.zeros((3, 3))
    s[0, 0], s[1, 1], s[2, 2], s[0, 1] = 10.0, 2.0, 3.0, 1.0
    w = O.view_weights_from_sums(s)
    # /sum -> /16; -min(0); /max(10/16) -> 1, .2, .3, .1 ; <0.2 -> 0 ; *0.3
    exp = np.zeros((3, 3))
    exp[0, 0], exp[1, 1], exp[2, 2] = 0.3, 0.2 * 0.3, 0.3 * 0.3
    assert np.allclose(w, exp)
    assert w[0, 1] == 0


def test_weights_degenerate_all_equal_is_nan():
    w = O.view_weights_from_sums(np.ones((3, 3)))
    assert np.all(np.isnan(w))          # NaN views stay "active" (NaN == 0 false)


def test_normalize_depth_by_hand():
    """recon:13-16 divides each depth slab by its own total, not each view:
    two slabs whose views have unequal sums, worked out by hand."""
    H = np.zeros((1, 1, 2, 1, 2))
    H[0, 0, :, 0, 0] = [1.0, 3.0]     # slab 0: total 4
    H[0, 0, :, 0, 1] = [2.0, 2.0]     # slab 1: total 4 -> but the views differ from slab 0
    H[0, 0, :, 0, 1] *= 5.0           # slab 1: total 20
    got = O.normalize_depth(H)
    assert np.array_equal(got[0, 0, :, 0, 0], [0.25, 0.75])
    assert np.array_equal(got[0, 0, :, 0, 1], [0.5, 0.5])
    # per-view sums over depth after the normalisation: 0.75 and 1.25; the
    # weights of recon:30-38 from them: /2 -> .375, .625; -min -> 0, .25;
    # /max -> 0, 1; cut; *0.3
    w = O.view_weights_from_sums(got.sum(axis=(0, 1, 4)))
    assert np.allclose(w[:, 0], [0.0, 0.3])
    # a 3-D slab: every entry of slab z divided by the same total
    rng = np.random.default_rng(0)
    H3 = rng.random((4, 4, 3, 3, 2))
    n3 = O.normalize_depth(H3)
    for z in range(2):
        tot = sum(float(x) for x in H3[:, :, :, :, z].ravel())
        assert np.allclose(n3[:, :, :, :, z] * tot, H3[:, :, :, :, z], rtol=1e-14)
        # per-view sums stay unequal (a per-view normalisation would make them equal)
        assert np.ptp(n3[:, :, :, :, z].sum(axis=(0, 1))) > 1e-3


def test_weights_of_production_psf_cut_and_vary():
    """A raw PSF whose in-aperture views have unequal sums: one in-disc view
    falls under the 0.2 cut and the active weights differ (recon:30-38)."""
    from tests.helpers import random_production_psf
    H, Ht, info = random_production_psf(np.random.default_rng(3))
    cp = O.compact_from_dense(H, Ht)
    assert info["cut_view"] not in cp.views
    assert len(cp.views) == 176
    w = np.array(cp.weights)
    assert w.max() == pytest.approx(0.3) and w.min() >= 0.3 * 0.2 and np.ptp(w) > 0.1
    assert cp.rf == info["rf"] and cp.rb == info["rb"]
    assert any(b == f + 1 for f, b in zip(cp.rf, cp.rb))


# ------------------------------------------------------------------- conv2
def test_conv2_same_bruteforce():
    rng = np.random.default_rng(3)
    x = rng.random((7, 9))
    h = rng.random((5, 3))
    got = O.conv2_same(x, h)
    m1, m2 = 2, 1
    exp = np.zeros_like(x)
    for p in range(7):
        for q in range(9):
            acc = 0.0
            for a in range(7):
                for b in range(9):
                    i, j = p - a + m1, q - b + m2
                    if 0 <= i < 5 and 0 <= j < 3:
                        acc += x[a, b] * h[i, j]
            exp[p, q] = acc
    assert np.allclose(got, exp, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------- forward / backward
@pytest.fixture(scope="module")
def tiny():
    wl = S.TINY
    Hps = S.dense_phase_space(wl)
    Ht = S.adjoint_ht(Hps)
    cp = O.compact_from_phase_space(Hps, Ht)
    return wl, Hps, Ht, cp


def test_forward_compact_equals_literal(tiny):
    wl, Hps, Ht, cp = tiny
    rng = np.random.default_rng(4)
    X = rng.random((wl.I1, wl.I2, wl.Z))
    Xl = O.to_lenslet(X, 15)
    for k in (0, 17, 88, len(cp.views) - 1):
        u, v = cp.views[k]
        a = O.forward_literal(X, Hps, u, v, 15)
        b = O.forward_compact(Xl, cp.fwd[k], cp.rf)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)


def test_backward_compact_equals_literal_and_adjoint(tiny):
    wl, Hps, Ht, cp = tiny
    rng = np.random.default_rng(5)
    X = rng.random((wl.I1, wl.I2, wl.Z))
    for k in (3, 60, 150):
        u, v = cp.views[k]
        R = rng.random((wl.n1, wl.n2))
        a = O.backward_literal(Ht, R, u, v, 15)
        b = O.from_lenslet(O.backward_compact(R, cp.bwd[k], cp.rb), 15)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
        # <F x, r> == <x, F^T r> for the exact-adjoint Ht
        y = O.forward_literal(X, Hps, u, v, 15)
        assert abs((y * R).sum() - (X * a).sum()) <= 1e-12 * abs((y * R).sum())


def test_compact_table_matches_generator(tiny):
    wl, Hps, Ht, cp = tiny
    for k in (0, 40, 176):
        u, v = cp.views[k]
        for zi, z in enumerate(wl.z_um):
            t = S.compact_fwd_table(u, v, z)
            assert np.allclose(cp.fwd[k][zi], t, rtol=1e-15, atol=1e-18)
            # exact adjoint: backward table is the point reflection
            assert np.array_equal(cp.bwd[k][zi], cp.fwd[k][zi][::-1, ::-1])


# ---------------------------------------------------------- sample_filter
def test_sample_filter_literal_equals_separable():
    rng = np.random.default_rng(6)
    for (I1, I2, Z) in [(45, 60, 3), (30, 30, 11)]:
        Y = rng.random((I1, I2, Z))
        bank = O.filter_bank(I1, I2, Z, 15)
        a = O.sample_filter_literal(Y, bank)
        b = O.sample_filter_separable(Y, Z, 15)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-13)
        # renormalisation: every filtered plane keeps its sum
        assert np.allclose(a.sum(axis=(0, 1)), Y.sum(axis=(0, 1)), rtol=1e-13)


def test_filter_kernel_is_real_even_unit_sum():
    for I in (45, 60, 495, 1515):
        for om in (0.2, 6.5, 8.0):
            k = O.filter_kernel_1d(I, om)
            assert abs(k.sum() - 1) < 1e-12
            assert np.allclose(k[1:], k[1:][::-1], atol=1e-15)


def test_half_z_uses_matlab_round():
    assert O.matlab_round(2.5) == 3 and O.matlab_round(25.5) == 26 and O.matlab_round(-2.5) == -3
    # Z=51: half_z = 26, nf = 54 -> k = z + 1 (1-based), every plane filtered
    idx = O.filter_plane_index(51, 54)
    assert idx[0] == 1 and idx[-1] == 51
    idx = O.filter_plane_index(11, 12)
    assert idx == list(range(11))


def test_omega_array_cfg2():
    om = O.omega_array(51, 15)
    assert om.size == 54
    assert om.min() > 6.4 and om.max() <= 8.0


@pytest.mark.parametrize("Z,size,first_axis", [(11, 12, -5), (51, 54, -26), (101, 102, -50)])
def test_omega_array_closed_form(Z, size, first_axis):
    """recon:21-26,54-56: omega_k = (N-7) exp(-((k - 3R/2)/100 * 2 sqrt(ln 2))^2)
    = 8 * 2^(-a^2 / 2500) with a = k - 3R/2 (exp(-4 ln2 a^2/1e4) = 2^(-a^2/2500)),
    floored at 0.2.  Exact values at named indices; a dropped factor 2 in
    the exponent gives 8 * 2^(-a^2/10000) and fails."""
    om = O.omega_array(Z, 15)
    assert om.size == size
    a = np.arange(size) + first_axis
    ref = 8.0 * np.power(2.0, -(a.astype(np.float64) ** 2) / 2500.0)
    assert np.allclose(om, np.maximum(ref, 0.2), rtol=1e-14, atol=0)
    k0 = int(np.where(a == 0)[0][0])
    assert om[k0] == 8.0
    if Z == 11:   # 3R = 12: a runs -5 .. 6
        assert om[0] == pytest.approx(8.0 * 2.0 ** -0.01, rel=1e-15)
        assert om[-1] == pytest.approx(8.0 * 2.0 ** -(36 / 2500), rel=1e-15)
    if Z == 51:   # R = ceil(51/3) = 17 -> even 18, 3R = 54: a runs -26 .. 27
        assert om[0] == pytest.approx(8.0 * 2.0 ** -(676 / 2500), rel=1e-15)
        assert om[-1] == pytest.approx(6.53596, abs=1e-5)    # 8 * 2^-0.2916
    # a fifth of the Gaussian's half-maximum width: omega halves at |a| = 50
    if Z == 101:
        assert om[list(a).index(50)] == pytest.approx(4.0, rel=1e-15)
        assert om[list(a).index(-50)] == pytest.approx(4.0, rel=1e-15)


def test_filter_bank_is_analytic_gaussian_transform():
    """recon:57-60: |fftshift(fft2(exp(-(x/w)^2 - (y/w)^2)))| / max is the
    sampled transform of a Gaussian, exp(-(pi w f)^2) per axis (f in cycles
    per pixel), when the Gaussian is narrow against the image and the
    frequency grid covers its support."""
    I = 128
    bank = O.filter_bank(I, I, 11, 15)
    om = O.omega_array(11, 15)
    f = (np.arange(I) - I // 2) / I          # fftshift order
    for k in (0, 5, 11):
        g = np.exp(-(np.pi * om[k] * f) ** 2)
        ref = np.outer(g, g)
        assert np.allclose(bank[:, :, k], ref / ref.max(), atol=1e-9)


# ------------------------------------------------------------ update rules
def test_update_special_values():
    X = np.array([1.0, 2.0, 0.0, 3.0, 4.0, 5.0])
    B = np.array([2.0, 0.0, 1.0, 1.0, -1.0, 1.0])
    B1 = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 2.0])
    w = 0.3
    got = O.multiplicative_update(X, B, B1, w)
    C = np.array([2.0, 0.0, 0.0, 0.0, 0.0, 2.5])  # 0/0->0, 0*1/0=NaN->0, 3/0=inf->0, neg->0
    assert np.allclose(got, C * w + (1 - w) * X)
    r = O.ratio_update(np.array([1.0, 0.0, 2.0]), np.array([0.0, 0.0, 4.0]))
    assert np.array_equal(r, [0.0, 0.0, 0.5])


# ------------------------------------------------------------------ full RL
@pytest.mark.slow
def test_rl_compact_equals_literal_small():
    wl = S.Workload("t", 3, 4, (-4, 4), 1, 1)
    Hps = S.dense_phase_space(wl)
    H = O.inverse_phasespace(Hps)
    Ht = S.adjoint_ht(Hps)
    rng = np.random.default_rng(7)
    LF = rng.poisson(50.0, (wl.I1, wl.I2)).astype(np.float64)
    a = O.rl_literal(LF, H, Ht, 1, views_limit=6)
    cp = O.compact_from_dense(H, Ht)
    b = O.rl_compact(LF, cp, 1, views_limit=6)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


@pytest.mark.slow
def test_rl_compact_equals_literal_production_psf():
    """Same with a random raw H and an independent random Ht (not the
    adjoint; rb = rf + 1 in two planes) and unequal view weights: the compact
    backward tables read Ht at the view's own slot (recon:121,123) exactly
    as the full-resolution conv2 does."""
    from tests.helpers import random_production_psf
    rng = np.random.default_rng(9)
    H, Ht, info = random_production_psf(rng)
    LF = rng.poisson(50.0, (45, 60)).astype(np.float64)
    a = O.rl_literal(LF, H, Ht, 1, views_limit=8)
    cp = O.compact_from_dense(H, Ht)
    assert cp.rb == info["rb"]
    b = O.rl_compact(LF, cp, 1, views_limit=8)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)


# ------------------------------------------------------------------- MEX op
def _mex_loop(bitmap, K):
    """Per-output-pixel loop following mex:50-78 and mex:194-251 literally."""
    W, Hh = bitmap.shape
    kw = K.shape[0]
    out = np.zeros((W, Hh))
    for y in range(Hh):
        for x in range(W):
            bsx, bsy = x - kw // 2, y - kw // 2
            bex, bey = x + kw // 2 + 1, y + kw // 2 + 1
            ksx = -bsx if bsx < 0 else 0
            ksy = -bsy if bsy < 0 else 0
            kex = kw - (bex - W) if bex > W else kw
            key = kw - (bey - Hh) if bey > Hh else kw
            acc = 0.0
            for ky in range(ksy, key):
                for kx in range(ksx, kex):
                    acc += bitmap[bsx + kx, bsy + ky] * K[kx, ky, x % 15, y % 15]
            out[x, y] = acc
    return out


@pytest.mark.parametrize("W,Hh,kw", [(17, 13, 5), (20, 31, 7), (16, 16, 4), (9, 40, 31)])
def test_mex_op_equals_loop(W, Hh, kw):
    rng = np.random.default_rng(W * Hh + kw)
    b = rng.random((W, Hh))
    K = rng.random((kw, kw, 15, 15))
    assert np.allclose(O.mex_op(b, K), _mex_loop(b, K), rtol=1e-12, atol=1e-12)


def test_mex_op_is_adjoint_of_scatter_forward():
    rng = np.random.default_rng(9)
    X = rng.random((30, 45))
    Y = rng.random((30, 45))
    K = rng.random((11, 11, 15, 15))
    lhs = (O.forward_scatter(X, K) * Y).sum()
    rhs = (X * O.mex_op(Y, K)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_mex_op_batch_axis():
    rng = np.random.default_rng(10)
    b = rng.random((12, 14, 3))
    K = rng.random((5, 5, 15, 15))
    out = O.mex_op(b, K)
    for n in range(3):
        assert np.allclose(out[:, :, n], O.mex_op(b[:, :, n], K))


@pytest.mark.parametrize("W,Hh,kw", [(31, 40, 7), (20, 17, 31), (45, 33, 15)])
def test_mex_op_sparse_equals_dense(W, Hh, kw):
    """The scatter form over non-zero pixels equals the gather form
    (including kernels wider than the bitmap, where both clip at every edge)."""
    rng = np.random.default_rng(W * Hh + kw)
    b = np.zeros((W, Hh))
    xs = rng.integers(0, W, 25)
    ys = rng.integers(0, Hh, 25)
    b[xs, ys] = rng.random(25)
    xs, ys = np.nonzero(b)
    K = rng.random((kw, kw, 15, 15))
    got = O.mex_op_sparse(W, Hh, xs, ys, b[xs, ys], K)
    assert np.allclose(got, O.mex_op(b, K), rtol=1e-13, atol=1e-13)
    with pytest.raises(ValueError):
        O.mex_op_sparse(W, Hh, xs, ys, b[xs, ys], K[:-1, :-1])


def test_vols_slice_fixture_matches_the_reference_file_summary():
    """The decoded reference MEX input (src/cuda/vols_slice.mat): 376 x 376 x
    77 single, 72,050 non-zeros, max 1.0858954, sum 3570.611 -- the figures
    read from the file when the reference was surveyed."""
    from tests.golden import load_vols_slice
    a = load_vols_slice()
    assert a.shape == (376, 376, 77) and a.dtype == np.float32
    assert np.count_nonzero(a) == 72050
    assert float(a.max()) == pytest.approx(1.0858954, abs=1e-7)
    assert float(a.astype(np.float64).sum()) == pytest.approx(3570.611, abs=1e-3)
    assert a.min() >= 0


# --------------------------------------------------- post-reconstruction step


def test_post_recon_by_hand():
    # planes as [row i][column j]; stack[:, :, z]
    p1 = [[1.0, 2.0], [3.0, 4.0]]    # max 4
    p2 = [[0.0, 8.0], [2.0, 6.0]]    # max 8
    p3 = [[1.0, 1.0], [1.0, 0.5]]    # max 1
    X = np.stack([np.array(p1), np.array(p2), np.array(p3)], axis=2)
    out, tiff = O.post_recon(X)
    # plane 1 * 8 / 4 = [[2, 4], [6, 8]], then / 8; plane 2 / 8; plane 3 / 1; global max 1
    want = np.stack([np.array([[0.25, 0.5], [0.75, 1.0]]),
                     np.array([[0.0, 1.0], [0.25, 0.75]]),
                     np.array([[1.0, 1.0], [1.0, 0.5]])], axis=2)
    assert np.array_equal(out, want)
    # 65535 * [0.25, 0.5, 0.75] = 16383.75, 32767.5, 49151.25 -> 16384, 32768, 49151
    want_u16 = np.stack([np.array([[16384, 32768], [49151, 65535]]),
                         np.array([[0, 65535], [16384, 49151]]),
                         np.array([[65535, 65535], [65535, 32768]])], axis=2)
    assert tiff.dtype == np.uint16 and np.array_equal(tiff, want_u16)
    # central plane of 3 (mesolf.m:101: ceil(3/2) = 2)
    mid, _ = O.post_recon(X, 1)
    assert np.array_equal(mid, want[:, :, 1:2])


def test_post_recon_zero_plane_and_nan():
    X = np.zeros((2, 3, 3))
    X[:, :, 0] = [[1, 2, 3], [4, 5, 6]]        # max 6
    X[:, :, 1] = [[9, 3, 0], [np.nan, 3, 0]]   # max 9 (NaN ignored)
    out, tiff = O.post_recon(X)                # plane 3 all zero: 0/0 -> NaN
    assert np.all(np.isnan(out[:, :, 2])) and np.all(tiff[:, :, 2] == 0)
    # plane 1 * 9 / 6 -> max 9 -> / 9: the original plane / 6
    assert np.array_equal(out[:, :, 0], (X[:, :, 0] * 9 / 6) / 9)
    assert np.isnan(out[1, 0, 1]) and tiff[1, 0, 1] == 0
    assert out[0, 0, 1] == 1.0 and tiff[0, 0, 1] == 65535


def test_im2uint16_values():
    v = np.array([-1.0, 0.0, 1e-300, 0.25, 0.5, 1.0, 2.0, np.inf, np.nan, 1 / 65535, 0.5 / 65535 * 0.999])
    assert O.im2uint16(v).tolist() == [0, 0, 0, 16384, 32768, 65535, 65535, 65535, 0, 1, 0]


@pytest.mark.parametrize("Z,v,lo,hi", [(51, 51, 0, 51), (51, 31, 10, 41), (51, 1, 25, 26),
                                       (4, 3, 0, 3), (4, 1, 1, 2), (7, 0, 0, 7)])
def test_valid_planes(Z, v, lo, hi):
    # mesolf.m:101-102 with ceil(Z/2) +- floor(v/2), 1-based inclusive
    assert O.valid_planes(Z, v) == (lo, hi)


@pytest.mark.parametrize("Z,v", [(51, 53), (4, 5), (51, 2), (51, -1)])
def test_valid_planes_rejects(Z, v):
    with pytest.raises(ValueError):
        O.valid_planes(Z, v)


# ------------------------------------------------------------- summary image


def test_movmean_documented_example():
    # MATLAB's movmean documentation: movmean(A, 3) of this vector
    A = np.array([4, 8, 6, -1, -2, -3, -1, 3, 4, 5], float)
    want = [6, 6, 13 / 3, 1, -2, -2, -1 / 3, 2, 4, 4.5]
    assert np.allclose(O.movmean(A, 3), want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 7, 20, 23, 40])
def test_movmean_bruteforce(k):
    rng = np.random.default_rng(k)
    x = rng.standard_normal((3, 17))
    got = O.movmean(x, k, axis=1)
    T = x.shape[1]
    before, after = (k // 2, k // 2 - 1) if k % 2 == 0 else ((k - 1) // 2, (k - 1) // 2)
    for i in range(3):
        for j in range(T):
            win = [x[i, m] for m in range(T) if j - before <= m <= j + after]
            assert abs(got[i, j] - sum(win) / len(win)) < 1e-12


def test_movmean_even_window_by_hand():
    # k = 4 covers [j-2, j+1]; k = 2 covers [j-1, j]
    x = np.array([1.0, 2, 3, 4, 5])
    assert np.allclose(O.movmean(x, 4), [1.5, 2, 2.5, 3.5, 4])
    assert np.allclose(O.movmean(x, 2), [1, 1.5, 2.5, 3.5, 4.5])


def test_rank_1_factorization_closed_form_and_svd():
    rng = np.random.default_rng(5)
    u = rng.random(30) + 0.1
    v = rng.random(12) + 0.1
    s, t = O.rank_1_factorization(np.outer(u, v), 1, 2.0)
    # t = v / |v| exactly in direction; s = Y t / 2 = u |v| / 2
    assert np.allclose(t, v / np.linalg.norm(v), rtol=1e-12)
    assert np.allclose(s, u * np.linalg.norm(v) / 2.0, rtol=1e-12)
    Y = rng.random((40, 15))
    s, t = O.rank_1_factorization(Y, 200, 1.0)
    U, S, Vt = np.linalg.svd(Y)
    assert np.allclose(t, Vt[0] * np.sign(Vt[0, 0]), rtol=1e-9)
    assert np.allclose(s, U[:, 0] * S[0] * np.sign(U[0, 0]), rtol=1e-9)


def test_std_image_is_population_std():
    Y = np.random.default_rng(6).standard_normal((9, 31))
    assert np.allclose(O.std_image(Y), np.std(Y, axis=1), rtol=1e-12)


def test_summary_pixelwise_by_loops():
    rng = np.random.default_rng(7)
    npix, T, w1, w2, pw = 6, 25, 6, 3, 3
    movie = rng.random((npix, T)) * 50 + 10
    img, bs, bt = O.summary_pixelwise(movie, w1, w2, pw, 2)
    mx = movie.max()
    want = np.zeros(npix)
    for i in range(npix):
        x = movie[i] / mx
        d = [x[j] - np.mean(x[max(0, j - w1 // 2):min(T, j + w1 // 2)]) for j in range(T)]
        y = [np.mean(d[max(0, j - (w2 - 1) // 2):min(T, j + (w2 - 1) // 2 + 1)]) ** pw for j in range(T)]
        want[i] = np.std(y)
    assert np.allclose(img, want / want.max(), rtol=1e-10)
    s0, t0 = O.rank_1_factorization(movie, 2, mx)
    assert np.array_equal(bs, s0) and np.array_equal(bt, t0)


# --------------------------------------------------------------- patch tiling


def test_carpet_patches_by_hand():
    # 1000 / 3 = 333 -> 345 (multiple of 15); the last row/column ends at 1000,
    # 310 px, cut to 300
    got = O.carpet_patches(1000, 1000, [3, 3])
    rows = [(1, 345), (346, 690), (691, 990)]
    want = [((rows[i][0], rows[j][0]), (rows[i][1], rows[j][1])) for j in range(3) for i in range(3)]
    assert got == want
    # 100 / 2 = 50 -> 60; the second patch is 61..100 = 40 px -> 30 < 50: discarded
    assert O.carpet_patches(100, 100, 2) == [((1, 1), (60, 60))]
    # 700 x 451 in 2 x 3: heights 350 -> 360 (1..360, 361..700 = 340 -> 330);
    # widths 150 -> 150 (1..150, 151..300, 301..450; 451 is left out)
    got = O.carpet_patches(700, 451, [2, 3])
    assert got == [((1, 1), (360, 150)), ((361, 1), (690, 150)),
                   ((1, 151), (360, 300)), ((361, 151), (690, 300)),
                   ((1, 301), (360, 450)), ((361, 301), (690, 450))]


# -------------------------------------------------------------- rectification


def test_rectify_axis_by_hand():
    # center 8, step 15/15 = 1, 20 px: grid -1..20 (22 points), 8 is the 10th;
    # 10 - floor(15/2) = 3 -> start after 3 points (at 2), 19 left -> cut to 15
    assert np.array_equal(O.rectify_axis(8.0, 20, 15, 15.0), np.arange(2.0, 17.0))
    # center 3.5, step 0.5 (dx 7.5): first half 3.5 down to -0.5 (9 points),
    # 9 - 7 = 2 -> start at 0.5; up to 40.9 -> 0.5..40.5 (81 points) -> 75
    q = O.rectify_axis(3.5, 40, 15, 7.5)
    assert len(q) == 75 and q[0] == 0.5 and np.allclose(np.diff(q), 0.5)


def test_rectify_reproduces_planes_and_bilinear_by_hand():
    H, W = 23, 31
    yy, xx = np.meshgrid(np.arange(1, H + 1), np.arange(1, W + 1), indexing="ij")
    plane = 0.5 * xx - 0.25 * yy + 3.0
    out = O.rectify(plane, 10.3, 7.6, 15, 14.2)
    xq = O.rectify_axis(10.3, W, 15, 14.2)
    yq = O.rectify_axis(7.6, H, 15, 14.2)
    Y, X = np.meshgrid(yq, xq, indexing="ij")
    assert np.allclose(out, 0.5 * X - 0.25 * Y + 3.0, atol=1e-12)  # incl. extrapolated points
    # one bilinear cell by hand: corners 1, 2 (x), 3, 4 (y) at (y, x) = (1.25, 1.5)
    img = np.array([[1.0, 2.0], [3.0, 4.0]])
    from scipy.interpolate import RegularGridInterpolator
    gi = RegularGridInterpolator(([1, 2], [1, 2]), img, bounds_error=False, fill_value=None)
    assert abs(gi([[1.25, 1.5]])[0] - (0.75 * 1.5 + 0.25 * 3.5)) < 1e-15



def test_std_image_bg_by_hand():
    # Y = [[1, 2, 3], [2, 2, 2]], y_1 = [0.5, 1], y_2 = [1, 0, 2]: C = 1
    # pixel 1: A = 14, B = 2: (14 - 0.25)/3 - 4 + 2*2*0.5*1 - 0.25*1 = 0.33333...
    # pixel 2: A = 12, B = 2: (12 - 1)/3 - 4 + 2*2*1*1 - 1*1 = 2/3
    got = O.std_image_bg(np.array([[1.0, 2, 3], [2, 2, 2]]), np.array([0.5, 1.0]), np.array([1.0, 0, 2]))
    assert np.allclose(got, np.sqrt([13.75 / 3 - 4 + 2 - 0.25, 11 / 3 - 4 + 4 - 1]), rtol=1e-14)
    Y = np.random.default_rng(3).random((5, 9))
    assert np.allclose(O.std_image_bg(Y, np.zeros(5), np.random.default_rng(4).random(9)),
                       O.std_image(Y), rtol=1e-12)


def test_summary_global_background_is_divisor_form():
    movie = np.random.default_rng(8).random((30, 20)) * 100
    _, s, t = O.summary_global(movie, 3)
    s2, t2 = O.rank_1_factorization(movie, 3, movie.max())
    assert np.allclose(s, s2, rtol=1e-12) and np.allclose(t, t2, rtol=1e-12)


# ------------------------------------------------------ patch layout (raw cut)
HAND_SI = {"frames_x_offset": 100.0, "frames_y_offset": 90.0, "frames_dx": 15.0, "Nnum": 15,
           "frames_crop_border_microlenses": [1, 1, 1, 1]}


def test_determine_patch_info_carpet_by_hand():
    """A 200 x 230 frame, centre (x 100, y 90), dx = 15 (grid step 1 px, so
    the grid is the integers 1..230 / 1..200), crop 1 microlens per side:
    X grid: centre index 100, XqInit = 93 - 90 = 3 -> 227 columns -> 225 (15
    lenslets) -> crop -> 195; Y: index 90, YqInit = 83 - 75 = 8 -> 192 rows
    -> 180 -> 150.  Carpet [2, 1]: patch heights ceil(75/15)*15 = 75.
    Raw cut of patch 1 (margin ceil(15/2) - 2 = 6): rows 1 + 15 + 7 - 6 = 17 ..
    75 + 15 + 7 + 5 = 102, columns 1 + 15 + 2 - 6 = 12 .. 195 + 15 + 2 + 5 =
    217; centre row round((17 + 102)/2) = 60 (mod(90 - 60, 15) = 0), column
    round((12 + 217)/2) = 115 -> yCenter 60 - 17 = 43, xCenter 115 - 12 = 103."""
    got = O.determine_patch_info(200, 230, 100.0, 90.0, 15.0, 15, [1, 1, 1, 1], "carpet", n_patches=[2, 1])
    assert [p["location"] for p in got] == [[[1, 1], [75, 195]], [[76, 1], [150, 195]]]
    assert [p["raw_cut"] for p in got] == [[[17, 102], [12, 217]], [[92, 177], [12, 217]]]
    assert [(p["raw_cut_xCenter"], p["raw_cut_yCenter"]) for p in got] == [(103.0, 43.0), (103.0, 43.0)]
    # the raw cut rectifies about its centre to exactly the patch (image_load_module_aug.m:106)
    for p in got:
        (y0, y1), (x0, x1) = p["raw_cut"]
        assert len(O.rectify_axis(p["raw_cut_yCenter"], y1 - y0 + 1, 15, 15.0)) == p["size"][0]
        assert len(O.rectify_axis(p["raw_cut_xCenter"], x1 - x0 + 1, 15, 15.0)) == p["size"][1]


def test_roi_patches_by_hand():
    """ROI rows 16..120, columns 31..170 (105 x 140) with ROI_patch_size 64
    (-> 60): 2 x 3 patches in ind2sub order; second row 76..120 is 45 high
    (< 50, discarded); third column 151..170 is 20 wide -> 15 (discarded)."""
    got = O.roi_patches([[16, 31, 120, 170]], 64, 15)
    assert got == [((16, 31), (75, 90)), ((16, 91), (75, 150))]
    # an ROI lower than a patch gives one patch row of the ROI's own height
    # (80 -> 75); columns 1..120, 121..240 and 241..300 (60 wide)
    assert O.roi_patches([[1, 1, 80, 300]], 120, 15) == [((1, 1), (75, 120)), ((1, 121), (75, 240)),
                                                          ((1, 241), (75, 300))]


def test_determine_patch_info_stops_where_the_reference_does():
    with pytest.raises(IndexError):   # no right crop: the last patch's margin leaves the grid
        O.determine_patch_info(200, 230, 100.0, 90.0, 15.0, 15, [1, 0, 0, 1], "carpet", n_patches=[2, 1])
    with pytest.raises(ValueError):
        O.determine_patch_info(200, 230, 100.0, 90.0, 15.0, 15, [8, 8, 1, 1], "carpet", n_patches=[1, 1])
