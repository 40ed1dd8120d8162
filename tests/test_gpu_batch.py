"""Batches and the serving loop on the GPU.

A patch's work entries are the same whether it runs alone or in a batch, so
the batched results must equal the single-patch results bit for bit; one
batched patch is also checked against the FP64 oracle.  Graph replays are
compared with eager launches iteration by iteration, with host work and
synchronisation between replays.
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O
from tests.helpers import oracle_compact_from_abi, rel_l2

pytestmark = pytest.mark.gpu

WL = synth.SMALL                       # 5 planes, radii 2..3
GRIDS = [(12, 12), (9, 14), (20, 7)]   # uneven patches sharing one PSF


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def tabs():
    return tables.synthetic_compact(WL)


@pytest.fixture(scope="module")
def psf(ctx, tabs):
    p = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    yield p
    p.close()


def _image(n1, n2, seed):
    rng = np.random.default_rng(seed)
    return rng.poisson(40.0, (15 * n1, 15 * n2)).astype(np.float32) + 1.0


def _alone(psf, grid, img, iters, prec, warm=None):
    p = mesolf_amd.Patch(psf, *grid, precision=prec)
    p.set_image(img)
    if warm is None:
        p.init_volume()
    else:
        p.set_volume(warm)
    p.run(iters, filter=True)
    out = p.get_volume()
    p.close()
    return out


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batch_equals_alone_bitwise(psf, prec, graph):
    imgs = [_image(*g, seed=10 + i) for i, g in enumerate(GRIDS)]
    ref = [_alone(psf, g, im, 2, prec) for g, im in zip(GRIDS, imgs)]
    b = mesolf_amd.Batch(psf, capacity=4, precision=prec, graph=graph)
    patches = []
    for g, im in zip(GRIDS, imgs):
        p = mesolf_amd.Patch(psf, *g, precision=prec)
        p.set_image(im)
        p.init_volume()
        b.attach(p)
        patches.append(p)
    assert len(b) == 3
    b.run(2, filter=True)
    for p, r in zip(patches, ref):
        got = p.get_volume()
        assert np.array_equal(got, r), (prec, graph, p.n1, p.n2, float(np.abs(got - r).max()))
    b.close()
    for p in patches:
        p.close()


def test_batched_patch_matches_oracle(psf, tabs):
    cp = oracle_compact_from_abi(WL, tabs)
    grid = GRIDS[2]
    img = _image(*grid, seed=5)
    b = mesolf_amd.Batch(psf, capacity=3, precision="fp32")
    ps = []
    for i, g in enumerate(GRIDS):
        p = mesolf_amd.Patch(psf, *g, precision="fp32")
        p.set_image(img if g == grid else _image(*g, seed=50 + i))
        p.init_volume()
        b.attach(p)
        ps.append(p)
    b.run(2, filter=True)
    got = ps[2].get_volume()
    ref = O.rl_compact(img.astype(np.float64), cp, 2)
    err = rel_l2(got, ref)
    print(f"batched 20x7 patch, 2 iterations: rel L2 vs FP64 oracle = {err:.3e}")
    assert err <= 1e-4
    b.close()
    for p in ps:
        p.close()


def test_graph_replays_match_eager_each_iteration(psf):
    imgs = [_image(*g, seed=30 + i) for i, g in enumerate(GRIDS)]
    sets = []
    for graph in (True, False):
        b = mesolf_amd.Batch(psf, capacity=3, precision="bf16", graph=graph)
        ps = []
        for g, im in zip(GRIDS, imgs):
            p = mesolf_amd.Patch(psf, *g, precision="bf16")
            p.set_image(im)
            p.init_volume()
            b.attach(p)
            ps.append(p)
        sets.append((b, ps))
    for it in range(4):
        vols = []
        for b, ps in sets:
            b.run(1, filter=True)
            psf.ctx.synchronize()
            vols.append([p.get_volume() for p in ps])  # eager host work between replays
        for a, e in zip(*vols):
            assert np.array_equal(a, e), f"replay {it} differs from eager launches"
    for b, ps in sets:
        b.close()
        for p in ps:
            p.close()


def test_attach_detach_midway(psf):
    """Patches join and leave between iterations; each result equals its own
    single-patch history."""
    prec = "bf16"
    imgs = [_image(*g, seed=70 + i) for i, g in enumerate(GRIDS)]
    # alone: patch 0 runs 3 iterations, patch 1 runs 1, patch 2 runs 2
    ref = [_alone(psf, GRIDS[0], imgs[0], 3, prec), _alone(psf, GRIDS[1], imgs[1], 1, prec),
           _alone(psf, GRIDS[2], imgs[2], 2, prec)]
    b = mesolf_amd.Batch(psf, capacity=2, precision=prec)
    ps = []
    for g, im in zip(GRIDS, imgs):
        p = mesolf_amd.Patch(psf, *g, precision=prec)
        p.set_image(im)
        p.init_volume()
        ps.append(p)
    b.attach(ps[0])
    b.attach(ps[1])
    with pytest.raises(mesolf_amd.BudgetExceeded):
        b.attach(ps[2])  # capacity 2
    b.run(1)
    b.detach(ps[1])
    b.attach(ps[2])
    b.run(2)
    for p, r in zip(ps, ref):
        assert np.array_equal(p.get_volume(), r)
    ps[0].close()  # destroying an attached patch detaches it
    assert len(b) == 1
    b.close()
    ps[1].close()
    ps[2].close()


def test_server_results_equal_alone_and_warm_start(psf):
    """Two frames for each of three patches, submitted together: each result
    equals the single-patch computation (cold start for the first frame,
    warm start from that result for the second)."""
    prec = "bf16"
    frames = {i: [_image(*g, seed=100 + 10 * i + f) for f in range(2)] for i, g in enumerate(GRIDS)}
    ref = {}
    for i, g in enumerate(GRIDS):
        v1 = _alone(psf, g, frames[i][0], 3, prec)
        v2 = _alone(psf, g, frames[i][1], 3, prec, warm=v1)
        ref[i] = (v1, v2)
    srv = mesolf_amd.Server(psf, precision=prec, max_slots=2, iterations=3)
    tickets = []
    for f in range(2):
        for i in range(len(GRIDS)):
            tickets.append((i, f, srv.submit(i, frames[i][f], want_volume=True)))
    for i, f, t in tickets:
        lat, vol = srv.wait(t)
        assert lat > 0
        assert np.array_equal(vol, ref[i][f]), (i, f)
    st = srv.stats()
    assert st["completed"] == 6 and st["cold_starts"] == 3 and st["evictions"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert 1.0 <= st["mean_slots"] <= 2.0
    srv.close()


def test_server_budget_evicts_idle_patches(psf):
    """A budget that holds one patch: every new patch evicts the idle one, so
    every request starts cold and still matches the cold single-patch run."""
    prec = "fp32"
    g = (40, 40)
    imgs = [_image(*g, seed=200 + k) for k in range(3)]
    probe = mesolf_amd.Server(psf, precision=prec, max_slots=1, iterations=1)
    probe.wait(probe.submit(0, imgs[0]))
    one = probe.stats()["resident_bytes"]  # device bytes of one 40x40 patch
    probe.close()
    budget = int(1.6 * one)
    srv = mesolf_amd.Server(psf, precision=prec, max_slots=2, iterations=2, hbm_budget_bytes=budget)
    ts = [srv.submit(k, imgs[k], want_volume=True) for k in range(3)]
    for k, t in enumerate(ts):
        _, vol = srv.wait(t)
        assert np.array_equal(vol, _alone(psf, g, imgs[k], 2, prec))
    st = srv.stats()
    assert st["cold_starts"] == 3 and st["evictions"] == 2 and st["resident_patches"] == 1
    assert st["peak_resident_bytes"] <= budget
    srv.close()


def test_server_refuses_request_over_budget_and_keeps_serving(psf):
    """A patch whose state alone exceeds the HBM budget is refused at
    submission (MLF_ERR_BUDGET); requests that fit still complete and equal
    their single-patch runs."""
    prec = "bf16"
    g = (20, 20)
    imgs = [_image(*g, seed=500 + k) for k in range(2)]
    probe = mesolf_amd.Server(psf, precision=prec, max_slots=1, iterations=1)
    probe.wait(probe.submit(0, imgs[0]))
    one = probe.stats()["resident_bytes"]
    probe.close()
    srv = mesolf_amd.Server(psf, precision=prec, max_slots=2, iterations=2, hbm_budget_bytes=int(2.5 * one))
    t0 = srv.submit(0, imgs[0], want_volume=True)
    with pytest.raises(mesolf_amd.BudgetExceeded):
        srv.submit(9, _image(60, 60, seed=9))     # 9x the pixels of the budgeted patches
    t1 = srv.submit(1, imgs[1], want_volume=True)
    for k, t in ((0, t0), (1, t1)):
        _, vol = srv.wait(t)
        assert np.array_equal(vol, _alone(psf, g, imgs[k], 2, prec))
    st = srv.stats()
    assert st["submitted"] == 2 and st["completed"] == 2
    srv.close()


def test_server_injected_faults_fail_one_request_and_keep_serving(psf):
    """Host-forced failures: a failed admission fails that request alone; a
    failed batched iteration fails the requests then in flight and drops their
    patch state; later requests run (cold) and match their single-patch runs."""
    prec = "bf16"
    frames = {i: [_image(*g, seed=600 + 10 * i + f) for f in range(3)] for i, g in enumerate(GRIDS)}
    srv = mesolf_amd.Server(psf, precision=prec, max_slots=3, iterations=2)
    # 1) admission of patch 1 fails; patches 0 and 2 complete
    srv.inject_fault("admit", 1)
    ts = [srv.submit(i, frames[i][0], want_volume=True) for i in range(3)]
    with pytest.raises(mesolf_amd.MesolfError) as ei:
        srv.wait(ts[1])
    assert ei.value.status == mesolf_amd.lib.MLF_ERR_DEVICE and "injected" in ei.value.message
    v0 = srv.wait(ts[0])[1]
    v2 = srv.wait(ts[2])[1]
    assert np.array_equal(v0, _alone(psf, GRIDS[0], frames[0][0], 2, prec))
    assert np.array_equal(v2, _alone(psf, GRIDS[2], frames[2][0], 2, prec))
    # 2) the next batched iteration fails: the request in flight fails ...
    srv.inject_fault("iteration")
    t = srv.submit(0, frames[0][1], want_volume=True)
    with pytest.raises(mesolf_amd.MesolfError) as ei:
        srv.wait(t)
    assert ei.value.status == mesolf_amd.lib.MLF_ERR_DEVICE
    # ... and patch 0's state was dropped: its next frame starts cold
    before = srv.stats()["cold_starts"]
    t = srv.submit(0, frames[0][2], want_volume=True)
    _, v = srv.wait(t)
    assert srv.stats()["cold_starts"] == before + 1
    assert np.array_equal(v, _alone(psf, GRIDS[0], frames[0][2], 2, prec))
    # patch 2 kept its warm state through patch 0's failure
    t = srv.submit(2, frames[2][1], want_volume=True)
    _, v = srv.wait(t)
    assert np.array_equal(v, _alone(psf, GRIDS[2], frames[2][1], 2, prec, warm=v2))
    assert srv._lib.mlf_server_inject_fault(srv.handle, 99, 0) == mesolf_amd.lib.MLF_ERR_INVALID_ARGUMENT
    srv.close()


def test_server_destroy_cancels_queued_requests(psf):
    """Requests still queued when the server is destroyed end with the
    cancellation status, not an argument error."""
    prec = "bf16"
    g = (40, 40)
    srv = mesolf_amd.Server(psf, precision=prec, max_slots=1, iterations=30)
    ts = [srv.submit(k, _image(*g, seed=700 + k)) for k in range(4)]
    lib = srv._lib
    import ctypes as C
    import threading
    results = {}

    def waiter(t):
        ms = C.c_double()
        results[t] = lib.mlf_server_wait(srv.handle, C.c_uint64(t), C.byref(ms))

    ths = [threading.Thread(target=waiter, args=(t,)) for t in ts]
    for th in ths:
        th.start()
    # every waiter must be inside mlf_server_wait before the handle is
    # destroyed (waiting on a destroyed server is a caller error); the request
    # in flight runs 30 iterations, far longer than this
    import time
    time.sleep(0.5)
    assert srv.stats()["pending"] >= 1
    h = srv.handle
    srv.handle = None
    assert lib.mlf_server_destroy(h) == 0
    for th in ths:
        th.join()
    # the request in flight completes; those behind it (one slot, 30
    # iterations each) were still queued and are cancelled
    assert results[ts[-1]] == mesolf_amd.lib.MLF_ERR_CANCELLED
    assert set(results.values()) <= {0, mesolf_amd.lib.MLF_ERR_CANCELLED}
