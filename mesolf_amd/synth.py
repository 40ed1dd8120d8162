"""Seeded synthetic light-field workloads for the phase-space RL path.

Everything here is data generation (numpy), not the reconstruction: the
phase-space PSF family, the sphere ground truth and the Poisson measurement
used by the tests and the benchmark.

Conventions (0-based, MATLAB column-major axes: axis 0 is the first/"x" axis):

* N = 15 phases per lenslet axis, c0 = 7 the centre phase.
* Plane z (micrometres, integer) has kernel radius
  ``r_z = ceil(0.15 |z|) + 2`` lenslets, evaluated exactly in integers as
  ``(15|z| + 99) // 100 + 2``.
* View (u, v) kernel over pixel offsets (dx, dy), |dx|,|dy| <= 15 r_z:
  ``exp(-((dx-cx)^2 + (dy-cy)^2) / (2 sigma^2))`` with
  ``sigma = 15 (1 + 0.1|z|)`` px and parallax centre
  ``(cx, cy) = 15 * 0.08 * z * ((u-7)/7, (v-7)/7)`` px -- the (u, v) in the
  parallax term is the view coordinate normalised to [-1, 1], so every kernel
  keeps its centre inside its support.  Each kernel sums to 1.
* Views outside the disc (u-7)^2 + (v-7)^2 <= 7.5^2 have a zero kernel
  (177 views remain).
* The phase-space PSF array H_ps has IMG = 15 (2R + 1) px per side
  (R = max r_z) and centre c = 15 R + 7, so the compact forward table is
  ``fwd[d1, d2, s1, s2] = k(15 d1 + 7 - s1, 15 d2 + 7 - s2)``.
* Ht is the exact adjoint: ``Ht[q1,q2,u,v,z] = H_ps[IMG-1+7-u-q1, IMG-1+7-v-q2, u,v,z]``,
  whose compact backward table is the point reflection of the forward table.
* Truth: spheres of lateral radius 3 lenslets (45 px) and axial radius 3
  planes, amplitude U[0.5, 1], over a background of 0.01.
* Measurement: every view forward-projected (interleaved into the sensor
  image), scaled to a peak of 1000 counts, Poisson noise.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NNUM = 15
C0 = 7


def psf_radius(z_um: int) -> int:
    return (15 * abs(int(z_um)) + 99) // 100 + 2


@dataclass(frozen=True)
class Workload:
    name: str
    n1: int          # lenslets along axis 0
    n2: int          # lenslets along axis 1
    z_um: tuple      # plane positions (integer micrometres)
    n_spheres: int
    iters: int
    seed: int = 1

    @property
    def Z(self):
        return len(self.z_um)

    @property
    def I1(self):
        return NNUM * self.n1

    @property
    def I2(self):
        return NNUM * self.n2

    @property
    def radii(self):
        return [psf_radius(z) for z in self.z_um]

    @property
    def R(self):
        return max(self.radii)

    @property
    def IMG(self):
        return NNUM * (2 * self.R + 1)


def _zrange(a, b, step):
    return tuple(range(a, b + 1, step))


CFG1 = Workload("cfg1_33x33_11planes", 33, 33, _zrange(-20, 20, 4), 40, 8)
CFG2 = Workload("cfg2_101x101_51planes", 101, 101, _zrange(-100, 100, 4), 400, 8)
TINY = Workload("tiny_5x4_3planes", 5, 4, (-4, 0, 4), 2, 2)
SMALL = Workload("small_12x12_5planes", 12, 12, (-8, -4, 0, 4, 8), 6, 2)


def aperture_views():
    """(u, v) pairs inside the r = 7.5 disc, row-major."""
    return [(u, v) for u in range(NNUM) for v in range(NNUM)
            if (u - C0) ** 2 + (v - C0) ** 2 <= 7.5 ** 2]


def view_profiles(u: int, v: int, z_um: int):
    """1-D factors (gx, gy) over dx in [-15r, 15r] of the (u, v, z) kernel."""
    r = psf_radius(z_um)
    sigma = 15.0 * (1.0 + 0.1 * abs(z_um))
    cx = 15.0 * 0.08 * z_um * (u - C0) / 7.0
    cy = 15.0 * 0.08 * z_um * (v - C0) / 7.0
    d = np.arange(-15 * r, 15 * r + 1, dtype=np.float64)
    gx = np.exp(-((d - cx) ** 2) / (2 * sigma ** 2))
    gy = np.exp(-((d - cy) ** 2) / (2 * sigma ** 2))
    return gx / gx.sum(), gy / gy.sum()


def _phase_table(g: np.ndarray, r: int) -> np.ndarray:
    """t[d + r, s] = g(15 d + 7 - s) over d in [-r, r], s in [0, 15)."""
    d = np.arange(-r, r + 1)
    s = np.arange(NNUM)
    dx = 15 * d[:, None] + C0 - s[None, :]
    ok = np.abs(dx) <= 15 * r
    return np.where(ok, g[np.clip(dx + 15 * r, 0, g.size - 1)], 0.0)


def compact_fwd_table(u: int, v: int, z_um: int, dtype=np.float64) -> np.ndarray:
    """Forward table [d1+r, d2+r, s1, s2] for one view and plane."""
    r = psf_radius(z_um)
    gx, gy = view_profiles(u, v, z_um)
    tx = _phase_table(gx, r)
    ty = _phase_table(gy, r)
    return (tx[:, None, :, None] * ty[None, :, None, :]).astype(dtype)


def dense_phase_space(wl: Workload, views=None) -> np.ndarray:
    """H_ps [IMG, IMG, 15, 15, Z] (zero for views outside the aperture)."""
    views = aperture_views() if views is None else views
    IMG, R = wl.IMG, wl.R
    c = 15 * R + C0
    H = np.zeros((IMG, IMG, NNUM, NNUM, wl.Z))
    for (u, v) in views:
        for zi, z in enumerate(wl.z_um):
            r = psf_radius(z)
            gx, gy = view_profiles(u, v, z)
            H[c - 15 * r:c + 15 * r + 1, c - 15 * r:c + 15 * r + 1, u, v, zi] = gx[:, None] * gy[None, :]
    return H


def adjoint_ht(H_ps: np.ndarray) -> np.ndarray:
    """Exact-adjoint Ht (already cropped, i.e. the array gen_phasespace returns)."""
    IMG, N = H_ps.shape[0], H_ps.shape[2]
    c0 = (N - 1) // 2
    Ht = np.zeros_like(H_ps)
    q = np.arange(IMG)
    for u in range(N):
        src1 = IMG - 1 + c0 - u - q
        ok1 = (src1 >= 0) & (src1 < IMG)
        for v in range(N):
            src2 = IMG - 1 + c0 - v - q
            ok2 = (src2 >= 0) & (src2 < IMG)
            blk = H_ps[np.clip(src1, 0, IMG - 1)[:, None], np.clip(src2, 0, IMG - 1)[None, :], u, v, :]
            Ht[:, :, u, v, :] = blk * (ok1[:, None] & ok2[None, :])[:, :, None]
    return Ht


def sphere_truth(wl: Workload, rng: np.random.Generator, dtype=np.float32) -> np.ndarray:
    """Ground-truth volume [I1, I2, Z]."""
    I1, I2, Z = wl.I1, wl.I2, wl.Z
    vol = np.full((I1, I2, Z), 0.01, dtype=dtype)
    rl = 3 * NNUM
    rz = 3
    cx = rng.uniform(0, I1, wl.n_spheres)
    cy = rng.uniform(0, I2, wl.n_spheres)
    cz = rng.uniform(0, Z, wl.n_spheres)
    amp = rng.uniform(0.5, 1.0, wl.n_spheres)
    for k in range(wl.n_spheres):
        x0, x1 = max(0, int(cx[k] - rl)), min(I1, int(cx[k] + rl) + 1)
        y0, y1 = max(0, int(cy[k] - rl)), min(I2, int(cy[k] + rl) + 1)
        z0, z1 = max(0, int(cz[k] - rz)), min(Z, int(cz[k] + rz) + 1)
        xs = (np.arange(x0, x1) + 0.5 - cx[k]) / rl
        ys = (np.arange(y0, y1) + 0.5 - cy[k]) / rl
        zs = (np.arange(z0, z1) + 0.5 - cz[k]) / rz
        m = xs[:, None, None] ** 2 + ys[None, :, None] ** 2 + zs[None, None, :] ** 2 <= 1.0
        vol[x0:x1, y0:y1, z0:z1] += (amp[k] * m).astype(dtype)
    return vol


def poisson_measurement(clean_sensor: np.ndarray, rng: np.random.Generator, peak=1000.0):
    """Scale to ``peak`` counts and draw Poisson noise (float32 image)."""
    img = np.asarray(clean_sensor, dtype=np.float64)
    img = img * (peak / img.max())
    return rng.poisson(img).astype(np.float32)
