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


def random_production_psf(rng, R=3, rf=(2, 1, 2), rb_support=(1, 1, 2), cut_view=(7, 3)):
    """Raw PSFs H, Ht [15(2R+1), 15(2R+1), 15, 15, Z] shaped like a real
    LFrecon PSF pair rather than the synthetic family:

    * H: every in-aperture view (disc r <= 7.5) holds independent uniform
      random (non-separable) kernels over a (2 rf + 1)-lenslet square, each
      view scaled by its own factor in [0.5, 1.5] so the view weights of
      recon:30-38 differ; view ``cut_view`` is scaled by 0.05 so it falls
      under the 0.2 weight cut.
    * Ht: independent uniform random values (not the adjoint of H) over a
      (2 s + 1)-lenslet square per plane; the backward radius of recon:121
      is then s + 1, i.e. rb = rf + 1 where s = rf.

    Returns (H, Ht, info) with info = {"rf", "rb", "cut_view"}.  H is given
    as the raw array whose gen_phasespace is the constructed phase-space PSF
    (supports stay 15 px inside the border, so the inverse is exact)."""
    Z = len(rf)
    IMG = 15 * (2 * R + 1)
    Hps = np.zeros((IMG, IMG, 15, 15, Z))
    for z, r in enumerate(rf):
        lo, hi = 15 * (R - r), 15 * (R + r + 1)
        Hps[lo:hi, lo:hi, :, :, z] = rng.random((hi - lo, hi - lo, 15, 15))
    for u in range(15):
        for v in range(15):
            if (u - 7) ** 2 + (v - 7) ** 2 > 7.5 ** 2:
                Hps[:, :, u, v, :] = 0.0
            else:
                Hps[:, :, u, v, :] *= 0.05 if (u, v) == tuple(cut_view) else rng.uniform(0.5, 1.5)
    H = O.inverse_phasespace(Hps)
    Ht = np.zeros_like(Hps)
    for z, s in enumerate(rb_support):
        lo, hi = 15 * (R - s), 15 * (R + s + 1)
        Ht[lo:hi, lo:hi, :, :, z] = rng.random((hi - lo, hi - lo, 15, 15))
    info = {"rf": list(rf), "rb": [s + 1 for s in rb_support], "cut_view": tuple(cut_view)}
    return H, Ht, info
