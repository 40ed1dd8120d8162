"""Host-side assembly of compact phase-space PSF tables (ABI order).

Used to feed ``PSF.from_compact`` from the synthetic PSF family without
materialising the dense [S,S,15,15,Z] arrays.  The preparation mirrors
recon:13-40: per-depth normalisation, view weights, spiral visiting order.
"""
from __future__ import annotations

import numpy as np

from . import synth


def spiral_order(N: int = 15):
    """recon:156-188 visiting order as 0-based (u, v) pairs."""
    out = []
    lo, hi = 0, N - 1
    while lo <= hi:
        if lo == hi:
            out.append((lo, lo))
            break
        out += [(i, lo) for i in range(lo, hi + 1)]
        out += [(hi, j) for j in range(lo + 1, hi + 1)]
        out += [(i, hi) for i in range(hi - 1, lo - 1, -1)]
        out += [(lo, j) for j in range(hi - 1, lo, -1)]
        lo, hi = lo + 1, hi - 1
    return out


def view_weights(sums: np.ndarray, cut: float = 0.2) -> np.ndarray:
    """recon:30-38 on a [15, 15] table of per-view PSF sums."""
    w = np.array(sums, dtype=np.float64, copy=True)
    w[np.isnan(w)] = 0
    w = w / w.sum()
    w = w - w.min()
    w = w / w.max()
    w[w < cut] = 0
    return w * 0.3


def synthetic_compact(wl: synth.Workload, views=None):
    """Compact ABI tables of the synthetic PSF for workload ``wl``.

    Returns dict(uv [V,2] int32, weights [V] f64, rf [Z] int32,
    fwd flat float32 (view-major, plane, ABI block order)).  The backward
    table is the exact adjoint (pass bwd=None to the library).
    """
    views = synth.aperture_views() if views is None else views
    Z = wl.Z
    radii = wl.radii
    # The synthetic kernels are separable: block[d1,d2,s1,s2] = tx[d1,s1] * ty[d2,s2].
    # Pass 1: per-depth normalisation (recon:13-16) and per-view sums (recon:30)
    # from the factor sums; pass 2 writes the normalised FP32 blocks.
    fac = {}
    depth_sum = np.zeros(Z)
    for (u, v) in views:
        for zi, z in enumerate(wl.z_um):
            r = radii[zi]
            gx, gy = synth.view_profiles(u, v, z)
            tx = synth._phase_table(gx, r)
            ty = synth._phase_table(gy, r)
            fac[(u, v, zi)] = (tx, ty)
            depth_sum[zi] += tx.sum() * ty.sum()
    sums = np.zeros((15, 15))
    for (u, v) in views:
        sums[u, v] = sum(fac[(u, v, zi)][0].sum() * fac[(u, v, zi)][1].sum() / depth_sum[zi]
                         for zi in range(Z))
    w = view_weights(sums)
    order = [(u, v) for (u, v) in spiral_order(15) if w[u, v] != 0]
    per_view = sum((2 * r + 1) ** 2 * 225 for r in radii)
    fwd = np.empty(per_view * len(order), np.float32)
    o = 0
    for (u, v) in order:
        for zi in range(Z):
            tx, ty = fac[(u, v, zi)]
            # ABI order [d2, d1, s2, s1]
            blk = np.einsum("bt,as->bats", ty / depth_sum[zi], tx)
            n = blk.size
            fwd[o:o + n] = blk.ravel()
            o += n
    return dict(uv=np.array(order, np.int32), weights=np.array([w[u, v] for u, v in order]),
                rf=np.array(radii, np.int32), fwd=fwd)
