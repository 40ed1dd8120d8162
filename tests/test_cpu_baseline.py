"""Pins for the oracle's C/OpenMP restatement (the timed CPU baseline).

The FP32 C forms must reproduce the FP64 numpy oracle: the literal
full-resolution view step (recon:191-208 + 107-137), the strided view steps
on the compact tables, sample_filter, and conv2 'same' itself.
"""
import numpy as np
import pytest

from mesolf_amd import synth, tables
from oracle import cpu_baseline as CB
from oracle import mesolf_oracle as O
from tests.helpers import oracle_compact_from_abi, rel_l2, synthetic_problem


@pytest.fixture(scope="module", autouse=True)
def built():
    from oracle import build
    build.build()


@pytest.mark.parametrize("shape,k", [((31, 40), (7, 5)), ((45, 30), (61, 61)), ((20, 17), (3, 9))])
def test_conv2_same_matches_oracle(shape, k):
    rng = np.random.default_rng(shape[0] + k[0])
    x = rng.random(shape)
    h = rng.random(k)
    assert rel_l2(CB.conv2_same(x, h), O.conv2_same(x, h)) <= 2e-6


def test_literal_view_step_matches_oracle():
    wl = synth.TINY
    Hps = synth.dense_phase_space(wl)
    Ht = synth.adjoint_ht(Hps)
    rng = np.random.default_rng(3)
    X = rng.random((wl.I1, wl.I2, wl.Z))
    LF = rng.poisson(30.0, (wl.I1, wl.I2)).astype(np.float64)
    for (u, v) in ((7, 3), (2, 9)):
        y = O.forward_literal(X, Hps, u, v, 15)
        R = O.ratio_update(LF[u::15, v::15], y)
        ones = np.ones((wl.n1, wl.n2))
        ref = O.multiplicative_update(X, O.backward_literal(Ht, R, u, v, 15),
                                      O.backward_literal(Ht, ones, u, v, 15), 0.3)
        got = CB.view_step_literal(X, LF, Hps[:, :, u, v, :], Ht[:, :, u, v, :], u, v, 0.3)
        assert rel_l2(got, ref) <= 1e-5, (u, v)


def test_synth_view_slices_are_the_dense_family():
    wl = synth.TINY
    Hps = synth.dense_phase_space(wl)
    Ht = synth.adjoint_ht(Hps)
    Hv, Htv = CB.view_slices_from_synth(wl, 5, 11)
    assert np.array_equal(Hv, Hps[:, :, 5, 11, :]) and np.array_equal(Htv, Ht[:, :, 5, 11, :])


def test_strided_view_steps_match_oracle():
    wl = synth.SMALL
    tabs, cp, truth, LF = synthetic_problem(wl, seed=3)
    X0 = O.initial_volume(LF.astype(np.float64), wl.Z)
    st = CB.StridedTables(cp, views=range(6))
    got = CB.view_steps_strided(O.to_lenslet(X0, 15), LF.astype(np.float32), st)
    ref = O.rl_compact(LF.astype(np.float64), cp, 1, views_limit=6, filter_planes=False)
    assert rel_l2(O.from_lenslet(got.astype(np.float64), 15), ref) <= 1e-5


def test_filter_planes_match_oracle():
    wl = synth.SMALL
    X = np.random.default_rng(4).random((wl.I1, wl.I2, wl.Z))
    bank = O.filter_bank(wl.I1, wl.I2, wl.Z, 15)
    assert rel_l2(CB.filter_planes(X), O.sample_filter_literal(X, bank)) <= 1e-6
