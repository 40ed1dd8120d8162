"""Shared helpers for the parity tests."""
import numpy as np

from mesolf_amd import synth, tables
from oracle import mesolf_oracle as O


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_compact_from_abi(wl, tabs):
    """Oracle CompactPSF holding the same (normalised) tables as ``tabs``."""
    Z = wl.Z
    rf = [int(r) for r in tabs["rf"]]
    views = [tuple(int(x) for x in uv) for uv in tabs["uv"]]
    cp = O.CompactPSF(N=15, Z=Z, views=views, weights=[float(w) for w in tabs["weights"]],
                      rf=rf, rb=list(rf))
    fwd = tabs["fwd"].astype(np.float64)
    o = 0
    for k in range(len(views)):
        fz, bz = [], []
        for zi in range(Z):
            K = 2 * rf[zi] + 1
            blk = fwd[o:o + K * K * 225].reshape(K, K, 15, 15)  # [d2, d1, s2, s1]
            o += K * K * 225
            t = np.transpose(blk, (1, 0, 3, 2))                  # [d1, d2, s1, s2]
            fz.append(np.ascontiguousarray(t))
            bz.append(np.ascontiguousarray(t[::-1, ::-1]))
        cp.fwd.append(fz)
        cp.bwd.append(bz)
    return cp


def synthetic_problem(wl, seed=1, views=None):
    """(tabs, oracle CompactPSF, truth, LF) for a synthetic workload (all 177
    aperture views, or the subset ``views``); the measurement is made with the
    FP64 oracle forward."""
    rng = np.random.default_rng(seed)
    tabs = tables.synthetic_compact(wl, views=views)
    cp = oracle_compact_from_abi(wl, tabs)
    truth = synth.sphere_truth(wl, rng, dtype=np.float64)
    Xl = O.to_lenslet(truth, 15)
    clean = np.zeros((wl.I1, wl.I2))
    for k, (u, v) in enumerate(cp.views):
        clean[u::15, v::15] = O.forward_compact(Xl, cp.fwd[k], cp.rf)
    LF = synth.poisson_measurement(clean, rng)
    return tabs, cp, truth, LF
