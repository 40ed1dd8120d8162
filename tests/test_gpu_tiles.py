"""Tile shapes and the autotuner on the GPU.

Every built tile shape must reproduce the FP64 oracle (the shapes only move
work between workgroups); the autotuner must pick one of them, record it in
the cache used by later patches, and a tuned batch must still match its
patches run alone bit for bit.
"""
import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O
from tests.helpers import oracle_compact_from_abi, rel_l2

pytestmark = pytest.mark.gpu

WL = synth.Workload("odd_37x21", 37, 21, (-20, 0, 12), 1, 1)
# (forward columns per wave, backprojection N-tiles, phase groups per entry,
# backprojection waves per SIMD), in the tuner's candidate order
SHAPES = {"fp32": [(mt, nt, g, 2) for g in (0, 1) for mt in (2, 1) for nt in (5, 3)],
          "bf16": [s for g in (0, 1) for mt in (4, 2)
                   for s in ((mt, 4, g, 2), (mt, 5, g, 2), (mt, 3, g, 2), (mt, 3, g, 3))]}
VOL_TOL = {"fp32": 1e-4, "bf16": 2e-3}


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def problem(ctx):
    tabs = tables.synthetic_compact(WL)
    cp = oracle_compact_from_abi(WL, tabs)
    psf = mesolf_amd.PSF.from_compact(ctx, tabs["uv"], tabs["weights"], tabs["rf"], tabs["fwd"])
    lf = np.random.default_rng(8).poisson(50.0, (WL.I1, WL.I2)).astype(np.float32) + 1.0
    ref = O.rl_compact(lf.astype(np.float64), cp, 2)
    yield psf, lf, ref
    psf.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_every_tile_shape_matches_oracle(problem, prec):
    psf, lf, ref = problem
    outs = []
    for shape in SHAPES[prec]:
        p = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision=prec)
        p.set_tile_shape(*shape)
        assert p.tile_shape() == shape
        p.set_image(lf)
        p.init_volume()
        p.run(2)
        x = p.get_volume()
        p.close()
        err = rel_l2(x, ref)
        print(f"{prec} tile shape {shape}: rel L2 vs FP64 oracle after 2 iterations = {err:.3e}")
        assert err <= VOL_TOL[prec], (shape, err)
        outs.append(x)
    if prec == "fp32":
        for x in outs[1:]:
            assert rel_l2(x, outs[0]) <= 1e-6
    # phase groups per entry and waves per SIMD only move work between
    # workgroups and registers: bit-identical
    half = len(SHAPES[prec]) // 2
    for i in range(half):
        assert np.array_equal(outs[i], outs[i + half]), SHAPES[prec][i]
    for i, s in enumerate(SHAPES[prec]):
        if s[3] == 3:
            j = SHAPES[prec].index((s[0], s[1], s[2], 2))
            assert np.array_equal(outs[i], outs[j]), (s, SHAPES[prec][j])


def test_unknown_tile_shape_rejected(problem):
    psf, _, _ = problem
    p = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision="bf16")
    with pytest.raises(mesolf_amd.InvalidInput):
        p.set_tile_shape(3, 5)
    with pytest.raises(mesolf_amd.InvalidInput):
        p.set_tile_shape(4, 5, 2)
    with pytest.raises(mesolf_amd.InvalidInput):
        p.set_tile_shape(4, 5, 0, 3)  # three waves per SIMD only with 3 N-tiles
    p.close()
    q = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision="fp32")
    with pytest.raises(mesolf_amd.InvalidInput):
        q.set_tile_shape(2, 3, 0, 3)  # FP32 has only the two-wave variant
    with pytest.raises(mesolf_amd.InvalidInput):
        q.set_tile_shape(2, 4)  # 4 N-tiles are built for BF16 only
    q.close()
    f = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision="fp16")
    with pytest.raises(mesolf_amd.InvalidInput):
        f.set_tile_shape(4, 4)
    f.close()


def test_autotune_picks_a_shape_and_caches_it(problem):
    psf, lf, ref = problem
    p = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision="bf16")
    p.set_image(lf)
    p.init_volume()
    before = p.get_volume()
    best = p.autotune(views=4)
    shape = (best["fwd_columns_per_wave"], best["bwd_phase_tiles"], best["bwd_groups_per_entry"],
             best["bwd_waves_per_simd"])
    assert shape in SHAPES["bf16"] and best["ms_per_view"] > 0
    assert p.tile_shape() == shape
    assert np.array_equal(p.get_volume(), before)  # tuning restores the volume
    q = mesolf_amd.Patch(psf, WL.n1, WL.n2, precision="bf16")
    assert q.tile_shape() == shape  # later patches of the same key start tuned
    p.run(2)
    assert rel_l2(p.get_volume(), ref) <= VOL_TOL["bf16"]
    p.close()
    q.close()


def test_tuned_batch_matches_alone(problem):
    psf, _, _ = problem
    grids = [(37, 21), (16, 30)]
    imgs = [np.random.default_rng(40 + i).poisson(30.0, (15 * g[0], 15 * g[1])).astype(np.float32) + 1.0
            for i, g in enumerate(grids)]
    b = mesolf_amd.Batch(psf, capacity=2, precision="bf16")
    ps = []
    for g, im in zip(grids, imgs):
        p = mesolf_amd.Patch(psf, *g, precision="bf16")
        p.set_image(im)
        p.init_volume()
        b.attach(p)
        ps.append(p)
    best = b.autotune(views=4)
    shape = (best["fwd_columns_per_wave"], best["bwd_phase_tiles"], best["bwd_groups_per_entry"],
             best["bwd_waves_per_simd"])
    assert shape in SHAPES["bf16"]
    b.run(2)
    for p, g, im in zip(ps, grids, imgs):
        assert p.tile_shape() == shape
        q = mesolf_amd.Patch(psf, *g, precision="bf16")
        q.set_tile_shape(*shape)
        q.set_image(im)
        q.init_volume()
        q.run(2)
        assert np.array_equal(p.get_volume(), q.get_volume())
        q.close()
    b.close()
    for p in ps:
        p.close()


def test_tune_cache_file_entries_are_validated(problem, tmp_path, monkeypatch):
    """Cached shapes are checked against the shapes this build has: a bogus
    entry (a shape never built) or an untagged line leaves the default; a
    valid tagged entry is applied; FP16 and BF16 have separate keys."""
    psf, _, _ = problem
    rf = ",".join(str(int(r)) for r in psf.radii()[0])
    path = tmp_path / "tune.txt"
    path.write_text(
        f"mlf-tune-v4 patch|bf16|{WL.n1}x{WL.n2}|rf,{rf} 3 7 0 2 0.1\n"   # not a built shape
        f"mlf-tune-v3 patch|bf16|16x30|rf,{rf} 2 3 0 0.1\n"               # older format tag
        f"mlf-tune-v4 patch|fp16|{WL.n1}x{WL.n2}|rf,{rf} 2 3 1 3 0.1\n"    # valid, fp16 key only
        f"mlf-tune-v4 patch|fp32|16x30|rf,{rf} 1 3 0 2 0.1\n")             # valid fp32 entry
    monkeypatch.setenv("MLF_TUNE_CACHE", str(path))
    shapes = {}
    for prec, grid in (("bf16", (WL.n1, WL.n2)), ("fp16", (WL.n1, WL.n2)), ("bf16", (16, 30)),
                       ("fp32", (16, 30))):
        p = mesolf_amd.Patch(psf, *grid, precision=prec)
        shapes[(prec, grid)] = p.tile_shape()
        p.close()
    assert shapes[("bf16", (WL.n1, WL.n2))] == (4, 4, 0, 2)   # bogus entry ignored: the default
    assert shapes[("fp16", (WL.n1, WL.n2))] == (2, 3, 1, 3)   # its own key
    assert shapes[("bf16", (16, 30))] == (4, 4, 0, 2)         # older-format line ignored
    assert shapes[("fp32", (16, 30))] == (1, 3, 0, 2)
