"""Synthetic movies and patch sets for the batched and serving workloads.

Time series (18 frames/s): the 400-sphere truth of the benchmark patch, at
the same sphere density per lenslet for other grids, with calcium-like
activity: every sphere group follows its own trace, a baseline of 1 plus
exponentially decaying transients (tau 0.5 s) at Poisson spike times
(0.5 Hz).  A frame is the Poisson-noised sum over groups of trace x forward
projection of the group's spheres, scaled so that the brightest frame of a
patch peaks at 1000 counts.  Spheres are split into `groups` groups (one
trace per group) so that each patch needs `groups` forward projections
rather than one per sphere.  Patches of other grids are lenslet-aligned
crops (at per-patch offsets) of one such movie on the largest grid, so a
set of uneven patches needs one set of projections.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import synth

FPS = 18.0


def calcium_traces(n: int, frames: int, rng: np.random.Generator, tau_s: float = 0.5,
                   rate_hz: float = 0.5, fps: float = FPS) -> np.ndarray:
    """[n, frames] traces: 1 + sum of exp(-(t - t_spike)/tau) over spikes."""
    dt = 1.0 / fps
    decay = np.exp(-dt / tau_s)
    out = np.empty((n, frames))
    for i in range(n):
        spikes = rng.poisson(rate_hz * dt, frames)
        c = 0.0
        for f in range(frames):
            c = c * decay + spikes[f]
            out[i, f] = 1.0 + c
    return out


@dataclasses.dataclass
class MoviePatch:
    n1: int
    n2: int
    projections: np.ndarray   # [groups, 15 n1, 15 n2] clean sensor images per group
    traces: np.ndarray        # [groups, frames]
    scale: float              # counts per unit of the clean image

    def frame(self, f: int, rng: np.random.Generator) -> np.ndarray:
        clean = np.tensordot(self.traces[:, f % self.traces.shape[1]], self.projections, axes=1)
        return rng.poisson(np.maximum(clean * self.scale, 0.0)).astype(np.float32)


def make_patch(psf, n1: int, n2: int, frames: int, rng: np.random.Generator, groups: int = 4,
               precision: str = "bf16") -> MoviePatch:
    """Sphere truth for an n1 x n2 patch split into `groups` activity groups,
    projected through `psf` (BF16 projections: the data only has to be
    plausible), with independent traces per group."""
    from .lib import Patch

    base = synth.CFG2
    n_sph = max(groups, int(round(base.n_spheres * n1 * n2 / (base.n1 * base.n2))))
    p = Patch(psf, n1, n2, precision=precision)
    proj = np.empty((groups, 15 * n1, 15 * n2), np.float32)
    for gi, cnt in enumerate(np.array_split(np.arange(n_sph), groups)):
        sub = synth.Workload(f"patch_{n1}x{n2}_g{gi}", n1, n2, base.z_um, len(cnt), 3)
        v = synth.sphere_truth(sub, np.random.default_rng(rng.integers(1 << 62)))
        if gi > 0:
            v -= np.float32(0.01)  # the background belongs to group 0
        p.set_volume(v)
        del v
        proj[gi] = p.project_all()
    p.close()
    traces = calcium_traces(groups, frames, rng)
    peak = max(float(np.tensordot(traces[:, f], proj, axes=1).max()) for f in range(frames))
    return MoviePatch(n1, n2, proj, traces, 1000.0 / peak)


def crop(img: np.ndarray, n1: int, n2: int, o1: int, o2: int) -> np.ndarray:
    """Lenslet-aligned crop [15 o1 : 15 (o1 + n1), 15 o2 : 15 (o2 + n2)]."""
    return np.asfortranarray(img[15 * o1:15 * (o1 + n1), 15 * o2:15 * (o2 + n2)])


def patch_grids(rng: np.random.Generator, count: int, lo: int, hi: int):
    """Lenslet grids drawn uniformly from [lo, hi] per side."""
    return [tuple(int(x) for x in rng.integers(lo, hi + 1, 2)) for _ in range(count)]
