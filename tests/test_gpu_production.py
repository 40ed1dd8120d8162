"""GPU parity on inputs shaped like production rather than like the synthetic
family, and at the benchmark's full shapes.

* A production-like PSF pair: random non-separable H, an independent random
  Ht (not the adjoint; rb = rf + 1 in two planes), unequal view weights and
  one in-aperture view under the 0.2 cut (recon:13-40, 121, 123), through
  ``PSF.from_dense`` and the recon drop-in.
* A plane of radius 25 (> 23) switches the forward to the Q-scatter MFMA
  kernel.
* ``relative_change`` of the drop-in's statistics against the oracle's
  volumes.
* 101 x 101 lenslets x 51 planes (the benchmark patch): forward projections
  and whole view steps against the FP64 oracle on a 4-view sub-table.
* Batches and serving with 51 planes and 61-121-lenslet patches.

Tolerances (relative L2 vs the FP64 oracle): FP32 forward 1e-5, FP32
volumes 1e-4; BF16 / FP16 as in test_gpu_parity.
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O
from tests.helpers import oracle_compact_from_abi, random_production_psf, rel_l2, synthetic_problem

pytestmark = pytest.mark.gpu

TOL = {"fp32": (1e-5, 1e-4), "bf16": (4e-3, 2e-2), "fp16": (1e-3, 5e-3)}


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


# ------------------------------------------------------- production PSF pair
@pytest.fixture(scope="module")
def prod():
    rng = np.random.default_rng(9)
    H, Ht, info = random_production_psf(rng)
    # the library takes single arrays: the oracle starts from the same values
    H, Ht = H.astype(np.float32), Ht.astype(np.float32)
    cp = O.compact_from_dense(H.astype(np.float64), Ht.astype(np.float64))
    LF = rng.poisson(60.0, (60, 75)).astype(np.float32)  # 4 x 5 lenslets
    return H, Ht, info, cp, LF


@pytest.fixture(scope="module")
def prod_psf(ctx, prod):
    H, Ht, info, cp, LF = prod
    p = mesolf_amd.PSF.from_dense(ctx, H, Ht)
    yield p
    p.close()


def test_production_psf_tables(prod, prod_psf):
    """View order and aperture bit-exact (the cut view is skipped), weights
    and per-plane radii equal to the oracle's."""
    H, Ht, info, cp, LF = prod
    views, w = prod_psf.views()
    assert views == cp.views
    assert info["cut_view"] not in views and len(views) == 176
    assert np.allclose(w, cp.weights, rtol=1e-12, atol=0)
    assert np.ptp(w) > 0.1
    rf, rb = prod_psf.radii()
    assert list(rf) == cp.rf and list(rb) == cp.rb
    assert list(rb) == [f + d for f, d in zip(cp.rf, (0, 1, 1))]


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_production_psf_forward(prod, prod_psf, prec):
    H, Ht, info, cp, LF = prod
    p = mesolf_amd.Patch(prod_psf, 4, 5, precision=prec)
    X = np.random.default_rng(21).random((60, 75, 3)).astype(np.float32)
    p.set_volume(X)
    Xl = O.to_lenslet(X.astype(np.float64), 15)
    for k in (0, 1, 80, 175):
        y = p.forward_view(k)
        ref = O.forward_compact(Xl, cp.fwd[k], cp.rf)
        assert rel_l2(y, ref) <= TOL[prec][0], (k, rel_l2(y, ref))
    p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_production_psf_rl_two_iterations(prod, prod_psf, prec):
    """The backward tables read from the real-shaped Ht (rb up to R + 1) and
    per-view weights in the update (recon:130-137)."""
    H, Ht, info, cp, LF = prod
    p = mesolf_amd.Patch(prod_psf, 4, 5, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(2)
    got = p.get_volume(f64=True)
    ref = O.rl_compact(LF.astype(np.float64), cp, 2)
    err = rel_l2(got, ref)
    print(f"production PSF, {prec}, 2 iterations: rel L2 vs FP64 oracle = {err:.3e}")
    assert err <= TOL[prec][1]
    p.close()


def test_production_psf_recon_dropin_and_relative_change(prod):
    """recon drop-in from the raw arrays; the per-iteration relative change
    ||X_k - X_{k-1}|| / ||X_{k-1}|| equals the oracle's."""
    H, Ht, info, cp, LF = prod
    st = {}
    got = mesolf_amd.recon_theory_phase_space_LFM_with_filter(LF, H, Ht, 3, True, 1, stats=st)
    inter = {0: O.initial_volume(LF.astype(np.float64), 3)}
    ref = O.rl_compact(LF.astype(np.float64), cp, 3,
                       callback=lambda it, X: inter.__setitem__(it + 1, X.copy()))
    assert rel_l2(got, ref) <= 1e-4
    assert st["iterations"] == 3 and st["views_per_iteration"] == 176
    for it in range(3):
        a, b = inter[it + 1], inter[it]
        rc = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert st["relative_change"][it] == pytest.approx(rc, rel=1e-4), (it, st["relative_change"], rc)
    assert st["relative_change"][0] > st["relative_change"][2] > 0


# ------------------------------------------------------ radius > 23 planes
WIDE25 = synth.Workload("wide25_12x10", 12, 10, (-148, -40, 0), 1, 1)


def test_radius_25_uses_qscatter_forward_and_matches(ctx):
    """A plane of radius 25 (z = -148 um) routes every plane's forward to the
    Q-scatter MFMA kernel; forward projections and one view step against the
    oracle, FP32 and BF16; FP16 (which needs radii <= 23) is refused."""
    assert WIDE25.radii == [25, 8, 2]
    views = synth.aperture_views()[::35]
    tabs, cp, truth, LF = synthetic_problem(WIDE25, seed=6, views=views)
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    rng = np.random.default_rng(13)
    X = rng.random((WIDE25.I1, WIDE25.I2, WIDE25.Z)).astype(np.float32)
    Xl = O.to_lenslet(X.astype(np.float64), 15)
    for prec in ("fp32", "bf16"):
        p = mesolf_amd.Patch(psf, WIDE25.n1, WIDE25.n2, precision=prec)
        p.set_volume(X)
        for k in range(psf.nviews):
            y = p.forward_view(k)
            ref = O.forward_compact(Xl, cp.fwd[k], cp.rf)
            assert rel_l2(y, ref) <= TOL[prec][0], (prec, k, rel_l2(y, ref))
        p.set_image(LF)
        p.init_volume()
        p.run(1, filter=False)
        got = p.get_volume(f64=True)
        ref = O.rl_compact(LF.astype(np.float64), cp, 1, filter_planes=False)
        assert rel_l2(got, ref) <= TOL[prec][1] / 4, (prec, rel_l2(got, ref))
        p.close()
    with pytest.raises(mesolf_amd.Unsupported):
        mesolf_amd.Patch(psf, WIDE25.n1, WIDE25.n2, precision="fp16")
    psf.close()


# ------------------------------------------ the benchmark patch, 101^2 x 51
CFG2_VIEWS = [synth.aperture_views()[i] for i in (0, 60, 88, 150)]


@pytest.fixture(scope="module")
def cfg2_problem():
    """cfg2 grid and planes with a 4-view sub-table (the oracle then builds
    only those views); measurement from the oracle forward of the 4 views."""
    wl = synth.CFG2
    return synthetic_problem(wl, seed=1, views=CFG2_VIEWS)


@pytest.fixture(scope="module")
def cfg2_forward_refs(cfg2_problem):
    wl = synth.CFG2
    tabs, cp, truth, LF = cfg2_problem
    X = truth.astype(np.float32)
    Xl = O.to_lenslet(X.astype(np.float64), 15)
    return X, [O.forward_compact(Xl, cp.fwd[k], cp.rf) for k in range(4)]


@pytest.fixture(scope="module")
def cfg2_steps_ref(cfg2_problem):
    tabs, cp, truth, LF = cfg2_problem
    return O.rl_compact(LF.astype(np.float64), cp, 1, views_limit=3, filter_planes=False)


@pytest.mark.slow
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_cfg2_forward_four_views(ctx, cfg2_problem, cfg2_forward_refs, prec):
    wl = synth.CFG2
    tabs, cp, truth, LF = cfg2_problem
    X, refs = cfg2_forward_refs
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    assert max(psf.radii()[0]) == 17
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
    p.set_volume(X)
    for k in range(4):
        err = rel_l2(p.forward_view(k), refs[k])
        print(f"cfg2 forward view {k} {prec}: rel L2 {err:.2e}")
        assert err <= TOL[prec][0], (k, err)
    p.close()
    psf.close()


@pytest.mark.slow
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_cfg2_three_view_steps(ctx, cfg2_problem, cfg2_steps_ref, prec):
    """Three whole view steps (forward, ratio, backprojection of the ratio and
    the normaliser, update) at 101 x 101 x 51 against the FP64 oracle."""
    wl = synth.CFG2
    tabs, cp, truth, LF = cfg2_problem
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    p = mesolf_amd.Patch(psf, wl.n1, wl.n2, precision=prec)
    p.set_image(LF)
    p.init_volume()
    p.run(1, views_limit=3, filter=False)
    err = rel_l2(p.get_volume(f64=True), cfg2_steps_ref)
    print(f"cfg2 3 view steps {prec}: rel L2 vs FP64 oracle {err:.2e}")
    assert err <= {"fp32": 1e-5, "bf16": 2e-3}[prec]
    p.close()
    psf.close()


# ---------------------------------- batches and serving at 51 planes, 61-121
BIG_GRIDS = [(61, 77), (121, 90), (97, 121)]


@pytest.fixture(scope="module")
def psf51(ctx):
    tabs = tables.synthetic_compact(synth.CFG2, views=synth.aperture_views()[::6])
    p = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    yield p
    p.close()


def _img(g, seed):
    return np.random.default_rng(seed).poisson(40.0, (15 * g[0], 15 * g[1])).astype(np.float32) + 1.0


def _alone(psf, g, img, iters, warm=None):
    p = mesolf_amd.Patch(psf, *g, precision="bf16")
    p.set_image(img)
    if warm is None:
        p.init_volume()
    else:
        p.set_volume(warm)
    p.run(iters)
    out = p.get_volume()
    p.close()
    return out


@pytest.mark.slow
def test_batch_51_planes_big_patches_equal_alone(psf51):
    """One BF16 iteration of three 61-121-lenslet patches, 51 planes (radii
    2..17), in one graph-captured batch equals each patch run alone, bit for bit."""
    assert psf51.Z == 51 and max(psf51.radii()[0]) == 17
    imgs = [_img(g, 300 + i) for i, g in enumerate(BIG_GRIDS)]
    ref = [_alone(psf51, g, im, 1) for g, im in zip(BIG_GRIDS, imgs)]
    b = mesolf_amd.Batch(psf51, capacity=3, precision="bf16", graph=True)
    ps = []
    for g, im in zip(BIG_GRIDS, imgs):
        p = mesolf_amd.Patch(psf51, *g, precision="bf16")
        p.set_image(im)
        p.init_volume()
        b.attach(p)
        ps.append(p)
    b.run(1)
    for p, r in zip(ps, ref):
        assert np.array_equal(p.get_volume(), r), (p.n1, p.n2)
    b.close()
    for p in ps:
        p.close()


@pytest.mark.slow
def test_server_51_planes_big_patches_warm_second_frame(psf51):
    """A serving round at 51 planes: two frames of two big patches, each
    result equal to the single-patch run (cold, then warm from it)."""
    grids = BIG_GRIDS[:2]
    frames = {i: [_img(g, 400 + 10 * i + f) for f in range(2)] for i, g in enumerate(grids)}
    ref = {}
    for i, g in enumerate(grids):
        v1 = _alone(psf51, g, frames[i][0], 1)
        ref[i] = (v1, _alone(psf51, g, frames[i][1], 1, warm=v1))
    srv = mesolf_amd.Server(psf51, precision="bf16", max_slots=2, iterations=1)
    ts = [(i, f, srv.submit(i, frames[i][f], want_volume=True)) for f in range(2) for i in range(2)]
    for i, f, t in ts:
        _, vol = srv.wait(t)
        assert np.array_equal(vol, ref[i][f]), (i, f)
    st = srv.stats()
    assert st["completed"] == 4 and st["cold_starts"] == 2
    srv.close()
