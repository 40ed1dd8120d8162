"""FP64 CPU restatement of MesoLF's phase-space Richardson-Lucy deconvolution.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- as the checker or the timed CPU
baseline, never as the product.  The product path is the HIP library under
``mesolf_amd/``.

Two families of functions live here:

* ``*_literal`` functions follow the reference ``.m`` code line by line
  (dense 5-D PSF arrays, full-resolution ``conv2(..., 'same')`` through
  ``scipy.signal.convolve2d``, FFT-based ``sample_filter``).  They are only
  feasible on tiny shapes.
* ``*_compact`` functions compute the same mathematics on the compact
  phase-space tables (lenslet offsets x 15x15 phases), i.e. without the 225x
  redundant full-resolution outputs.  They are the oracle at mid sizes.

Pinning (each function is checked by something independent in ``tests/``):

==========================  ==================================================
function                    pinned by
==========================  ==================================================
gen_spiral                  brute-force clockwise ring walk; known entries
gen_phasespace_literal      closed form (A.2); round trip with the inverse
phase_space_closed_form     literal reshape/permute emulation
view_weights                hand evaluation on a disc aperture (177 views)
filter_bank / sample_        separable circular convolution with real kernels
  filter_literal            (independent derivation), energy conservation
conv2_same                  brute-force loops on tiny inputs
forward/backward_literal    compact forms; adjointness <Fx,r> = <x,F^T r>
forward/backward_compact    literal forms (tiny), brute-force loops
rl_literal / rl_compact     each other at tiny sizes
mex_op                      brute-force loop over the kernel's own clipping
                            rules; adjoint of forward_scatter (odd kw)
mex_op_sparse               mex_op on sparse bitmaps (odd kw, kernel widths
                            up to the bitmap's and beyond)
post_recon / im2uint16      a 2x2x3 stack and its outputs computed by hand;
                            NaN planes; crop indices worked out for Z = 51, 4
movmean                     MATLAB's documented movmean(A, 3) example; brute
                            force over odd and even windows
rank_1_factorization        rank-1 movies (closed form); the leading singular
                            vectors (numpy SVD) after many iterations
std_image                   numpy's population std of the same rows
std_image_bg                a 2x3 case worked out by hand; y_1 = 0 gives std_image
summary_global              rank-1 of movie/max equals the divisor form
summary_pixelwise           a per-pixel loop composition of brute-force pieces
carpet_patches              tilings of 1000x1000, 100x100 (a discarded patch)
                            and 700x451 worked out by hand
roi_patches,                a frame, its crop and two ROI layouts worked out by
determine_patch_info        hand; every patch's raw cut rectifies about its
                            centre to exactly the patch size (the reference's
                            own consistency check, image_load_module_aug.m:106)
rectify_axis / rectify      a grid worked out by hand; scipy's
                            RegularGridInterpolator (linear, extrapolating);
                            planar images reproduced at every query point
==========================  ==================================================

There is no reference test data to pin against: the reference ships no
golden outputs (its only fixture, ``src/cuda/vols_slice.mat``, is an input).
The MATLAB/cuFFT rounding of the reference is therefore not reproduced
bit-for-bit; parity is by tolerance (see ``tests/``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.signal import convolve2d

# ---------------------------------------------------------------------------
# small MATLAB semantics helpers
# ---------------------------------------------------------------------------


def matlab_round(x: float) -> int:
    """MATLAB ``round``: half away from zero (numpy rounds half to even)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


# ---------------------------------------------------------------------------
# spiral view order -- recon_theory_phase_space_LFM_with_filter.m:156-188
# ---------------------------------------------------------------------------


def gen_spiral(N: int):
    """Return the 1-based (i_index, j_index) view tables of recon:156-188.

    ``mode(Nnum) == 0`` at recon:158 is the statistical mode of a scalar,
    i.e. ``Nnum`` itself, so the test is false for every N > 0 and
    ``loop_n = (Nnum + 1) / 2`` always; ``for k = 1:loop_n`` runs to
    ``floor(loop_n)``.
    """
    loop_n = (N + 1) / 2.0
    i_index = np.zeros(N * N, dtype=np.int64)
    j_index = np.zeros(N * N, dtype=np.int64)
    start = 1
    for k in range(1, int(math.floor(loop_n)) + 1):
        cs = N - 2 * (k - 1)
        bias = k - 1
        if cs > 1:
            i_new = np.concatenate([np.arange(1, cs + 1), cs * np.ones(cs - 1, np.int64),
                                    np.arange(cs - 1, 0, -1), np.ones(cs - 2, np.int64)])
            j_new = np.concatenate([np.ones(cs, np.int64), np.arange(2, cs + 1),
                                    cs * np.ones(cs - 1, np.int64), np.arange(cs - 1, 1, -1)])
            n = 4 * (cs - 1)
            i_index[start - 1:start - 1 + n] = i_new + bias
            j_index[start - 1:start - 1 + n] = j_new + bias
            start += n
        else:
            i_index[start - 1] = 1 + bias
            j_index[start - 1] = 1 + bias
    return i_index, j_index


# ---------------------------------------------------------------------------
# PSF preparation -- recon:13-17 and gen_phasespace_PSF_theory.m
# ---------------------------------------------------------------------------


def normalize_depth(H: np.ndarray) -> np.ndarray:
    """recon:13-16: divide each depth slab H(:,:,:,:,z) by its total sum."""
    H = np.array(H, dtype=np.float64, copy=True)
    for z in range(H.shape[4]):
        H[:, :, :, :, z] /= H[:, :, :, :, z].sum()
    return H


def phase_space_crop(S: int, N: int):
    """gen_phasespace_PSF_theory.m:6-8 -> (IMG, first 0-based index)."""
    IMG = S - ((S - N) % (2 * N))
    index1 = matlab_round(S / 2.0) - int(math.trunc(IMG / 2.0))  # 1-based
    return IMG, index1 - 1


def im_shift3(img: np.ndarray, sx: int, sy: int) -> np.ndarray:
    """gen_phasespace_PSF_theory.m:34-57: integer shift with zero fill."""
    out = np.zeros_like(img)
    X, Y = img.shape[:2]
    xs_dst = slice(max(sx, 0), X + min(sx, 0))
    xs_src = slice(max(-sx, 0), X - max(sx, 0))
    ys_dst = slice(max(sy, 0), Y + min(sy, 0))
    ys_src = slice(max(-sy, 0), Y - max(sy, 0))
    out[xs_dst, ys_dst] = img[xs_src, ys_src]
    return out


def gen_phasespace_literal(H: np.ndarray, Ht: np.ndarray):
    """Line-by-line emulation of gen_phasespace_PSF_theory.m:1-31.

    MATLAB ``reshape`` is column-major (numpy ``order='F'``) and ``permute``
    takes 1-based axes.  Returns (H_ps [IMG,IMG,N,N,Z], Ht_crop).
    """
    S = H.shape[0]
    N = H.shape[2]
    Z = H.shape[4]
    IMG, i0 = phase_space_crop(S, N)
    c0 = (N + 1) / 2.0
    Hd = np.zeros((IMG, IMG, N, N, Z), dtype=np.float64)
    for z in range(Z):
        for ii in range(N):
            for jj in range(N):
                Hd[:, :, ii, jj, z] = im_shift3(H[i0:i0 + IMG, i0:i0 + IMG, ii, jj, z],
                                                int(round((ii + 1) - c0)), int(round((jj + 1) - c0)))
    m = IMG // N
    Hd = np.reshape(Hd, (N, m, N, m, N, N, Z), order="F")
    Hd = np.transpose(Hd, (0, 2, 1, 3, 4, 5, 6))
    Hd = Hd[:, :, :, :, ::-1, ::-1, :]
    Hd = np.transpose(Hd, (0, 1, 4, 2, 5, 3, 6))
    Hd = np.reshape(Hd, (N, N, IMG, IMG, Z), order="F")
    Hd = np.transpose(Hd, (2, 3, 0, 1, 4))
    Ht_c = np.array(Ht[i0:i0 + IMG, i0:i0 + IMG], dtype=np.float64)
    return np.ascontiguousarray(Hd), Ht_c


def phase_space_closed_form(H: np.ndarray):
    """Closed form of the permutation in gen_phasespace_PSF_theory.m:22-27.

    ``H_ps[X,Y,u,v,z] = Hs[u + N*(X//N), v + N*(Y//N), N-1-X%N, N-1-Y%N, z]``
    where ``Hs`` is the cropped, de-shifted PSF (:6-20), 0-based.
    """
    S, N, Z = H.shape[0], H.shape[2], H.shape[4]
    IMG, i0 = phase_space_crop(S, N)
    c0 = (N - 1) // 2
    Hs = np.zeros((IMG, IMG, N, N, Z))
    for a in range(N):
        for b in range(N):
            Hs[:, :, a, b, :] = im_shift3(H[i0:i0 + IMG, i0:i0 + IMG, a, b, :], a - c0, b - c0)
    X = np.arange(IMG)
    out = np.empty((IMG, IMG, N, N, Z))
    for u in range(N):
        for v in range(N):
            rows = u + N * (X // N)
            cols = v + N * (X // N)
            pa = N - 1 - X % N
            out[:, :, u, v, :] = Hs[rows[:, None], cols[None, :], pa[:, None], pa[None, :], :]
    return out


def inverse_phasespace(H_ps: np.ndarray) -> np.ndarray:
    """Raw PSF H [S,S,N,N,Z] (S = IMG) whose gen_phasespace is ``H_ps``.

    Exact when every de-shifted entry stays inside the array (true for kernels
    supported within 7 px of the IMG border).
    """
    IMG, N, Z = H_ps.shape[0], H_ps.shape[2], H_ps.shape[4]
    c0 = (N - 1) // 2
    X = np.arange(IMG)
    Hs = np.zeros_like(H_ps)
    for u in range(N):
        for v in range(N):
            rows = u + N * (X // N)
            cols = v + N * (X // N)
            pa = N - 1 - X % N
            Hs[rows[:, None], cols[None, :], pa[:, None], pa[None, :], :] = H_ps[:, :, u, v, :]
    H = np.zeros_like(H_ps)
    for a in range(N):
        for b in range(N):
            H[:, :, a, b, :] = im_shift3(Hs[:, :, a, b, :], -(a - c0), -(b - c0))
    return H


# ---------------------------------------------------------------------------
# view weights -- recon:19-40
# ---------------------------------------------------------------------------


def view_weights(H_ps: np.ndarray, weight_cut: float = 0.2) -> np.ndarray:
    """recon:30-38.  ``weight(u,v) == 0`` marks an inactive view (recon:104)."""
    w = H_ps.sum(axis=(0, 1, 4))
    w[np.isnan(w)] = 0
    w = w / w.sum()
    w = w - w.min()
    w = w / w.max()
    w[w < weight_cut] = 0
    return w * 0.3


def view_weights_from_sums(wsum: np.ndarray, weight_cut: float = 0.2) -> np.ndarray:
    """Same arithmetic as :func:`view_weights` starting from per-view sums."""
    w = np.array(wsum, dtype=np.float64, copy=True)
    w[np.isnan(w)] = 0
    w = w / w.sum()
    w = w - w.min()
    w = w / w.max()
    w[w < weight_cut] = 0
    return w * 0.3


def active_views(weight: np.ndarray):
    """Views visited by recon:98-105, in spiral order: list of (u, v) 0-based."""
    N = weight.shape[0]
    i1, i2 = gen_spiral(N)
    return [(int(i1[k] - 1), int(i2[k] - 1)) for k in range(N * N)
            if not (weight[i1[k] - 1, i2[k] - 1] == 0)]


# ---------------------------------------------------------------------------
# object-space low-pass filter bank -- recon:20-26, 44-61, 222-240, 251-254
# ---------------------------------------------------------------------------


def nip_range(Z: int) -> int:
    r = int(math.ceil(Z / 3.0))
    return r + 1 if r % 2 == 1 else r


def omega_array(Z: int, N: int):
    R = nip_range(Z)
    axis = np.arange(1, 3 * R + 1) - 3 * R / 2.0
    om = (N - 7) * np.exp(-(axis / 100.0 * 2 * np.sqrt(np.log(2))) ** 2)
    om[om < 0.2] = 0.2
    return om


def filter_bank(I1: int, I2: int, Z: int, N: int) -> np.ndarray:
    """recon:45-61: obj_space_ft_filter_bank, [I1, I2, 3R] FP64."""
    x = np.arange(1, I1 + 1) - I1 / 2.0
    y = np.arange(1, I2 + 1) - I2 / 2.0
    xx = x[:, None] * np.ones((1, I2))
    yy = np.ones((I1, 1)) * y[None, :]
    om = omega_array(Z, N)
    bank = np.zeros((I1, I2, om.size))
    for i, w in enumerate(om):
        g = np.exp(-((xx / w) ** 2 + (yy / w) ** 2))
        buf = np.abs(np.fft.fftshift(np.fft.fft2(g)))
        bank[:, :, i] = buf / buf.max()
    return bank


def filter_plane_index(Z: int, nf: int):
    """sample_filter (recon:224-231): 0-based filter index per plane or None."""
    half_z = matlab_round(Z / 2.0)
    out = []
    for z in range(1, Z + 1):
        if half_z - nf / 2.0 + 1 <= z <= half_z + nf / 2.0:
            out.append(int(z - half_z + nf / 2.0) - 1)
        else:
            out.append(None)
    return out


def sample_filter_literal(X: np.ndarray, bank: np.ndarray) -> np.ndarray:
    """recon:222-240 with freq_filter recon:251-254, FP64 FFTs."""
    out = np.zeros_like(X)
    idx = filter_plane_index(X.shape[2], bank.shape[2])
    for z in range(X.shape[2]):
        img = X[:, :, z]
        k = idx[z]
        if k is None:
            out[:, :, z] = img
            continue
        s_before = img.sum()
        f = np.fft.fftshift(np.fft.fft2(img))
        buf = np.real(np.fft.ifft2(np.fft.ifftshift(f * bank[:, :, k])))
        out[:, :, z] = s_before * buf / buf.sum()
    return out


def filter_kernel_1d(I: int, omega: float) -> np.ndarray:
    """Real-space circular kernel equivalent to one axis of the filter bank.

    The bank is separable: F = (|G_x|/max) (x) (|G_y|/max) with G = fft(g),
    g(x) = exp(-((x - I/2)/omega)^2), x = 1..I; freq_filter therefore equals a
    circular convolution with kappa_x (x) kappa_y, kappa = real(ifft(|G|/max)).
    Returned with kappa[0] at index 0 (circular indexing).
    """
    x = np.arange(1, I + 1) - I / 2.0
    G = np.abs(np.fft.fft(np.exp(-(x / omega) ** 2)))
    return np.real(np.fft.ifft(G / G.max()))


def sample_filter_separable(X: np.ndarray, Z: int, N: int) -> np.ndarray:
    """sample_filter computed as separable circular convolutions (FP64)."""
    I1, I2 = X.shape[0], X.shape[1]
    om = omega_array(Z, N)
    idx = filter_plane_index(Z, om.size)
    out = np.empty_like(X)
    for z in range(Z):
        k = idx[z]
        if k is None:
            out[:, :, z] = X[:, :, z]
            continue
        k1 = filter_kernel_1d(I1, om[k])
        k2 = filter_kernel_1d(I2, om[k])
        # circular convolution via explicit circulant products (exact algebra)
        C1 = np.stack([np.roll(k1, s) for s in range(I1)], axis=1)  # C1[p,q] = k1[p-q]
        C2 = np.stack([np.roll(k2, s) for s in range(I2)], axis=1)
        buf = C1 @ X[:, :, z] @ C2.T
        out[:, :, z] = X[:, :, z].sum() * buf / buf.sum()
    return out


# ---------------------------------------------------------------------------
# literal operators -- recon:191-208 (CPU branch)
# ---------------------------------------------------------------------------


def conv2_same(x: np.ndarray, h: np.ndarray) -> np.ndarray:
    """MATLAB ``conv2(x, h, 'same')`` for odd-sized h (zero padding)."""
    return convolve2d(x, h, mode="same")


def forward_literal(X: np.ndarray, H_ps: np.ndarray, u: int, v: int, N: int) -> np.ndarray:
    """forwardProjectACC_phasespace_theory, recon:191-198 (u, v 0-based)."""
    proj = np.zeros(X.shape[:2])
    for z in range(X.shape[2]):
        proj += conv2_same(X[:, :, z], H_ps[:, :, u, v, z])
    c0 = (N - 1) // 2
    return proj[c0::N, c0::N]


def backward_literal(Ht: np.ndarray, R: np.ndarray, u: int, v: int, N: int) -> np.ndarray:
    """backwardProject_phasespace_theory, recon:201-208 (u, v 0-based)."""
    big = np.zeros((N * R.shape[0], N * R.shape[1]))
    big[u::N, v::N] = R
    out = np.zeros(big.shape + (Ht.shape[4],))
    for z in range(Ht.shape[4]):
        out[:, :, z] = conv2_same(big, Ht[:, :, u, v, z])
    return out


def ratio_update(LFsub, y):
    """recon:116-117."""
    with np.errstate(divide="ignore", invalid="ignore"):
        r = LFsub / y
    r[~np.isfinite(r)] = 0
    return r


def multiplicative_update(X, B, B1, w):
    """recon:130-137 (left-to-right evaluation of X .* B ./ B1)."""
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        C = X * B / B1
    C[np.isnan(C)] = 0
    C[np.isinf(C)] = 0
    C[C < 0] = 0
    Xn = C * w + (1 - w) * X
    Xn[Xn < 0] = 0
    return Xn


def initial_volume(LF: np.ndarray, Z: int) -> np.ndarray:
    """recon:72-73."""
    X = np.ones((LF.shape[0], LF.shape[1], Z))
    return X / X.sum() * LF.sum()


def rl_literal(LF, H, Ht, max_iter, views_limit=None):
    """recon:10-151, CPU branch (``Xguessv`` typo at :112 read as ``Xguess``,
    host zeros in ``sample_filter``).  Dense arrays; tiny sizes only."""
    Hn = normalize_depth(H)
    Htn = normalize_depth(Ht)
    H_ps, Ht_c = gen_phasespace_literal(Hn, Htn)
    N = H_ps.shape[2]
    Z = H_ps.shape[4]
    weight = view_weights(H_ps)
    bank = filter_bank(LF.shape[0], LF.shape[1], Z, N)
    X = initial_volume(LF, Z)
    views = active_views(weight)
    if views_limit is not None:
        views = views[:views_limit]
    ones = np.ones((LF.shape[0] // N, LF.shape[1] // N))
    for _ in range(max_iter):
        for (u, v) in views:
            y = forward_literal(X, H_ps, u, v, N)
            R = ratio_update(LF[u::N, v::N], y)
            B = backward_literal(Ht_c, R, u, v, N)
            B1 = backward_literal(Ht_c, ones, u, v, N)
            X = multiplicative_update(X, B, B1, weight[u, v])
        X = sample_filter_literal(X, bank)
    return X


# ---------------------------------------------------------------------------
# compact phase-space tables and operators
# ---------------------------------------------------------------------------


@dataclass
class CompactPSF:
    """Per active view (spiral order) and plane: compact lenslet-offset kernels.

    ``fwd[k][z]`` has shape (2rf+1, 2rf+1, N, N) indexed [d1+rf, d2+rf, s1, s2]:
    ``y[i] = sum_z sum_{L,s} X[N*L+s, z] * fwd[k][z][i-L+rf, s]``.
    ``bwd[k][z]`` has shape (2rb+1, 2rb+1, N, N):
    ``B[N*a+s, z] = sum_i R[i] * bwd[k][z][a-i+rb, s]``.
    """

    N: int
    Z: int
    views: list
    weights: list
    rf: list
    rb: list
    fwd: list = field(default_factory=list)
    bwd: list = field(default_factory=list)


def _trim_radius(tab: np.ndarray, center: int) -> int:
    """Smallest symmetric lenslet radius containing all nonzero entries.

    ``tab`` has shape (V, K, K, N, N) (views stacked on axis 0)."""
    nz = np.nonzero(np.any(tab != 0, axis=(0, 3, 4)))
    if nz[0].size == 0:
        return 0
    return int(max(np.abs(nz[0] - center).max(), np.abs(nz[1] - center).max()))


def compact_from_dense(H: np.ndarray, Ht: np.ndarray) -> CompactPSF:
    """Build compact tables from the raw PSFs exactly as recon:13-40 prepares
    them (depth normalisation, gen_phasespace, weights, spiral order)."""
    Hn = normalize_depth(H)
    Htn = normalize_depth(Ht)
    H_ps, Ht_c = gen_phasespace_literal(Hn, Htn)
    return compact_from_phase_space(H_ps, Ht_c)


def compact_from_phase_space(H_ps: np.ndarray, Ht_c: np.ndarray) -> CompactPSF:
    IMG, N, Z = H_ps.shape[0], H_ps.shape[2], H_ps.shape[4]
    R = (IMG // N - 1) // 2
    c0 = (N - 1) // 2
    weight = view_weights(H_ps)
    views = active_views(weight)
    V = len(views)
    s = np.arange(N)
    # forward: fwd[d+R, s] = H_ps[N*(d+R) + N-1-s]
    kf = np.arange(2 * R + 1)
    xf = N * kf[:, None] + (N - 1) - s[None, :]          # (2R+1, N)
    # backward: bwd[d+R+1, s] = Ht[N*(d+R) + c0 + s - u]  (zero outside)
    kb = np.arange(2 * R + 3)
    cp = CompactPSF(N=N, Z=Z, views=views, weights=[float(weight[u, v]) for u, v in views],
                    rf=[], rb=[])
    F = np.zeros((V, Z, 2 * R + 1, 2 * R + 1, N, N))
    B = np.zeros((V, Z, 2 * R + 3, 2 * R + 3, N, N))
    for k, (u, v) in enumerate(views):
        hp = H_ps[:, :, u, v, :]
        F[k] = np.transpose(hp[xf[:, None, :, None], xf[None, :, None, :], :], (4, 0, 1, 2, 3))
        xb1 = N * (kb[:, None] - 1) + c0 + s[None, :] - u
        xb2 = N * (kb[:, None] - 1) + c0 + s[None, :] - v
        ok1 = (xb1 >= 0) & (xb1 < IMG)
        ok2 = (xb2 >= 0) & (xb2 < IMG)
        ht = Ht_c[:, :, u, v, :]
        vals = ht[np.clip(xb1, 0, IMG - 1)[:, None, :, None], np.clip(xb2, 0, IMG - 1)[None, :, None, :], :]
        mask = ok1[:, None, :, None] & ok2[None, :, None, :]
        B[k] = np.transpose(vals * mask[..., None], (4, 0, 1, 2, 3))
    for z in range(Z):
        rf = _trim_radius(F[:, z], R)
        rb = _trim_radius(B[:, z], R + 1)
        cp.rf.append(rf)
        cp.rb.append(rb)
    cp.fwd = [[np.ascontiguousarray(F[k, z, R - cp.rf[z]:R + cp.rf[z] + 1, R - cp.rf[z]:R + cp.rf[z] + 1])
               for z in range(Z)] for k in range(V)]
    cp.bwd = [[np.ascontiguousarray(B[k, z, R + 1 - cp.rb[z]:R + 2 + cp.rb[z], R + 1 - cp.rb[z]:R + 2 + cp.rb[z]])
               for z in range(Z)] for k in range(V)]
    return cp


def to_lenslet(X: np.ndarray, N: int) -> np.ndarray:
    """[I1, I2, Z] -> [n1, n2, Z, N*N] with phase index s1 + N*s2."""
    I1, I2, Z = X.shape
    n1, n2 = I1 // N, I2 // N
    Xr = X.reshape(n1, N, n2, N, Z)  # [L1, s1, L2, s2, z]
    return np.ascontiguousarray(np.transpose(Xr, (0, 2, 4, 3, 1)).reshape(n1, n2, Z, N * N))


def from_lenslet(Xl: np.ndarray, N: int) -> np.ndarray:
    n1, n2, Z, _ = Xl.shape
    Xr = Xl.reshape(n1, n2, Z, N, N)  # [L1, L2, z, s2, s1]
    return np.ascontiguousarray(np.transpose(Xr, (0, 4, 1, 3, 2)).reshape(n1 * N, n2 * N, Z))


def _flat(tab: np.ndarray) -> np.ndarray:
    """(K, K, N, N) [d1, d2, s1, s2] -> (K*K, N*N) with phase index s1 + N*s2."""
    K, N = tab.shape[0], tab.shape[2]
    return np.transpose(tab, (0, 1, 3, 2)).reshape(K * K, N * N)


def forward_compact(Xl: np.ndarray, fwd_v: list, rf: list) -> np.ndarray:
    """Strided forward for one view: y[i] = sum_{z,d,s} X[i-d, z, s] W[z][d, s]."""
    n1, n2, Z, P = Xl.shape
    y = np.zeros((n1, n2))
    for z in range(Z):
        r = rf[z]
        K = 2 * r + 1
        Q = Xl[:, :, z, :].reshape(n1 * n2, P) @ _flat(fwd_v[z]).T  # [(L1,L2), (d1,d2)]
        Q = Q.reshape(n1, n2, K, K)
        for d1 in range(-r, r + 1):
            i1a, i1b = max(0, d1), min(n1, n1 + d1)
            if i1a >= i1b:
                continue
            for d2 in range(-r, r + 1):
                i2a, i2b = max(0, d2), min(n2, n2 + d2)
                if i2a >= i2b:
                    continue
                y[i1a:i1b, i2a:i2b] += Q[i1a - d1:i1b - d1, i2a - d2:i2b - d2, d1 + r, d2 + r]
    return y


def _im2col(Rimg: np.ndarray, r: int) -> np.ndarray:
    """cols[(a1,a2), (d1,d2)] = R[a1-d1, a2-d2] (zero outside)."""
    n1, n2 = Rimg.shape
    K = 2 * r + 1
    pad = np.zeros((n1 + 2 * r, n2 + 2 * r))
    pad[r:r + n1, r:r + n2] = Rimg
    cols = np.empty((n1, n2, K, K))
    for d1 in range(-r, r + 1):
        for d2 in range(-r, r + 1):
            cols[:, :, d1 + r, d2 + r] = pad[r - d1:r - d1 + n1, r - d2:r - d2 + n2]
    return cols.reshape(n1 * n2, K * K)


def backward_compact(Rimg: np.ndarray, bwd_v: list, rb: list) -> np.ndarray:
    """B[a, z, s] = sum_i R[i] Wt[z][a-i, s]  -> [n1, n2, Z, N*N]."""
    n1, n2 = Rimg.shape
    Z = len(bwd_v)
    P = bwd_v[0].shape[2] ** 2
    out = np.empty((n1, n2, Z, P))
    cache = {}
    for z in range(Z):
        r = rb[z]
        if r not in cache:
            cache[r] = _im2col(Rimg, r)
        out[:, :, z, :] = (cache[r] @ _flat(bwd_v[z])).reshape(n1, n2, P)
    return out


def rl_compact(LF: np.ndarray, cp: CompactPSF, max_iter: int, X0=None, views_limit=None,
               filter_planes=True, callback=None):
    """Full phase-space RL (recon:72-146) on compact tables, FP64."""
    N, Z = cp.N, cp.Z
    I1, I2 = LF.shape
    n1, n2 = I1 // N, I2 // N
    X = initial_volume(LF, Z) if X0 is None else np.array(X0, dtype=np.float64)
    Xl = to_lenslet(X, N)
    ones = np.ones((n1, n2))
    nv = len(cp.views) if views_limit is None else views_limit
    om = omega_array(Z, N)
    fidx = filter_plane_index(Z, om.size)
    kern = {}
    for it in range(max_iter):
        for k in range(nv):
            u, v = cp.views[k]
            y = forward_compact(Xl, cp.fwd[k], cp.rf)
            Rimg = ratio_update(LF[u::N, v::N], y)
            B = backward_compact(Rimg, cp.bwd[k], cp.rb)
            B1 = backward_compact(ones, cp.bwd[k], cp.rb)
            Xl = multiplicative_update(Xl, B, B1, cp.weights[k])
        if filter_planes:
            X = from_lenslet(Xl, N)
            for z in range(Z):
                f = fidx[z]
                if f is None:
                    continue
                if f not in kern:
                    kern[f] = (np.fft.fft(filter_kernel_1d(I1, om[f])), np.fft.fft(filter_kernel_1d(I2, om[f])))
                k1, k2 = kern[f]
                img = X[:, :, z]
                buf = np.real(np.fft.ifft2(np.fft.fft2(img) * (k1[:, None] * k2[None, :])))
                X[:, :, z] = img.sum() * buf / buf.sum()
            Xl = to_lenslet(X, N)
        if callback is not None:
            callback(it, from_lenslet(Xl, N))
    return from_lenslet(Xl, N)


# ---------------------------------------------------------------------------
# MEX operator -- src/cuda/mex_lfm_convolution_vec.cu:43-266
# ---------------------------------------------------------------------------


def mex_op(bitmap: np.ndarray, K: np.ndarray, nphase: int = 15) -> np.ndarray:
    """Race-free definition of lfm_convolution_kernel_vec (mex:43-266).

    ``bitmap`` [W, H] or [W, H, n] (x = first, contiguous axis),
    ``K`` [kw, kw, 15, 15].  For output (x, y):
    ``out = sum_{ky,kx} in[x-h+kx, y-h+ky] * K[kx, ky, x%15, y%15]``, h = kw//2,
    taps clipped by mex:55-78 (for even kw the tap landing on the last
    column/row is clipped too, as the kernel does).
    """
    squeeze = bitmap.ndim == 2
    b = bitmap[:, :, None] if squeeze else bitmap
    W, Hh, n = b.shape
    kw = K.shape[0]
    h = kw // 2
    out = np.zeros((W, Hh, n))
    xs = np.arange(W)
    ys = np.arange(Hh)
    kxs = np.maximum(0, h - xs)
    kxe = kw - np.maximum(0, xs + h + 1 - W)
    kys = np.maximum(0, h - ys)
    kye = kw - np.maximum(0, ys + h + 1 - Hh)
    px = xs % nphase
    py = ys % nphase
    bd = b.astype(np.float64)
    for ky in range(kw):
        yok = (ky >= kys) & (ky < kye)
        yin = ys - h + ky
        for kx in range(kw):
            xok = (kx >= kxs) & (kx < kxe)
            xin = xs - h + kx
            m = xok[:, None] & yok[None, :]
            if not m.any():
                continue
            kv = K[kx, ky][px[:, None], py[None, :]].astype(np.float64)  # [W, H]
            src = np.zeros((W, Hh, n))
            xi = np.clip(xin, 0, W - 1)
            yi = np.clip(yin, 0, Hh - 1)
            src = bd[xi[:, None], yi[None, :], :]
            out += (m * kv)[:, :, None] * src
    return out[:, :, 0] if squeeze else out


def mex_op_sparse(W: int, Hh: int, xs, ys, vals, K: np.ndarray, nphase: int = 15) -> np.ndarray:
    """mex_op of one [W, Hh] bitmap given by its non-zero pixels (xs, ys, vals),
    odd kw only: mex:55-78 then clips exactly the taps whose input pixel
    falls outside the bitmap, so each non-zero pixel (p, q) adds
    ``v * K[p - x + h, q - y + h, x % 15, y % 15]`` to every output (x, y)
    within h of it.  Cost ~ nnz * kw^2 instead of W * Hh * kw^2."""
    kw = K.shape[0]
    if kw % 2 == 0:
        raise ValueError("mex_op_sparse covers odd kernel widths only")
    h = kw // 2
    out = np.zeros((W, Hh))
    Kd = K.astype(np.float64)
    for p, q, v in zip(np.asarray(xs).tolist(), np.asarray(ys).tolist(), np.asarray(vals, np.float64).tolist()):
        x = np.arange(max(0, p - h), min(W, p + h + 1))
        y = np.arange(max(0, q - h), min(Hh, q + h + 1))
        kx = p - x + h
        ky = q - y + h
        out[x[0]:x[-1] + 1, y[0]:y[-1] + 1] += v * Kd[kx[:, None], ky[None, :],
                                                      (x % nphase)[:, None], (y % nphase)[None, :]]
    return out


def forward_scatter(X: np.ndarray, K: np.ndarray, nphase: int = 15) -> np.ndarray:
    """forwardProjectGPU.m:7-18 for one plane: sum_{a,b} conv2(X masked to
    phase (a,b), K[:,:,a,b], 'same')."""
    out = np.zeros(X.shape)
    for a in range(nphase):
        for b in range(nphase):
            t = np.zeros(X.shape)
            t[a::nphase, b::nphase] = X[a::nphase, b::nphase]
            out += conv2_same(t, K[:, :, a, b])
    return out


# ---------------------------------------------------------------------------
# post-reconstruction step -- LFM_reconstruction_module.m:45-53,60-62 and
# mesolf.m:101-102
# ---------------------------------------------------------------------------


def matlab_max(a) -> float:
    """max(a, [], 'all'): NaN entries are ignored; NaN only if all are NaN."""
    a = np.asarray(a, np.float64)
    return float("nan") if np.all(np.isnan(a)) else float(np.nanmax(a))


def im2uint16(x) -> np.ndarray:
    """im2uint16 of a double array: 65535*x rounded half away from zero,
    saturated to [0, 65535]; NaN -> 0."""
    v = np.asarray(x, np.float64)
    with np.errstate(invalid="ignore"):
        s = np.where(v >= 1.0, 65535.0, np.where(v > 0.0, v * 65535.0, 0.0))
        fl = np.floor(s)
        r = fl + (s - fl >= 0.5)
    return r.astype(np.uint16)


def valid_planes(Z: int, valid_range: int):
    """0-based [lo, hi) of recon_stack(:,:, ceil(Z/2)-floor(v/2) : ceil(Z/2)+floor(v/2))
    (mesolf.m:101-102); valid_range = 0 keeps every plane."""
    if valid_range == 0:
        return 0, Z
    if valid_range < 0 or valid_range % 2 == 0:
        raise ValueError("valid_range must be positive and odd")
    c = -(-Z // 2)
    lo, hi = c - valid_range // 2, c + valid_range // 2  # 1-based, inclusive
    if lo < 1 or hi > Z:
        raise ValueError("valid_range selects planes outside the stack")
    return lo - 1, hi


def post_recon(X: np.ndarray, valid_range: int = 0):
    """Top-layer balance and max normalisations of LFM_reconstruction_module.m
    (:46 plane 1 scaled by max(plane 2)/max(plane 1), left to right; :49-53 each
    plane divided by its max; :61 the stack divided by its max), then the TIFF
    image of :62 (im2uint16(stack / max(stack))) and the crop of mesolf.m:101-102.

    Returns (cropped stack float64, uint16 image of the whole stack)."""
    R = np.array(X, np.float64, copy=True)
    Z = R.shape[2]
    with np.errstate(invalid="ignore", divide="ignore"):
        R[:, :, 0] = R[:, :, 0] * matlab_max(R[:, :, 1]) / matlab_max(R[:, :, 0])
        for i in range(Z):
            layer = R[:, :, i]
            R[:, :, i] = layer / matlab_max(layer)
        R = R / matlab_max(R)
        tiff = im2uint16(R / matlab_max(R))
    lo, hi = valid_planes(Z, valid_range)
    return R[:, :, lo:hi], tiff


# ---------------------------------------------------------------------------
# summary image -- summary_image_generation_module.m:62-96 ('pixelwise',
# the default detrend mode), rank_1_factorization.m, compute_std_image.m
# ---------------------------------------------------------------------------


def movmean(x: np.ndarray, k: int, axis: int = -1) -> np.ndarray:
    """smoothdata(x, dim, 'movmean', k) / movmean(x, k) with the default
    'shrink' endpoints: the mean over the window [j - a, j + b] clipped to the
    array, a = b = (k-1)/2 for odd k, a = k/2, b = k/2 - 1 for even k."""
    x = np.moveaxis(np.asarray(x, np.float64), axis, -1)
    T = x.shape[-1]
    a, b = ((k - 1) // 2, (k - 1) // 2) if k % 2 else (k // 2, k // 2 - 1)
    P = np.concatenate([np.zeros(x.shape[:-1] + (1,)), np.cumsum(x, axis=-1)], axis=-1)
    j = np.arange(T)
    lo = np.maximum(j - a, 0)
    hi = np.minimum(j + b, T - 1)
    out = (P[..., hi + 1] - P[..., lo]) / (hi - lo + 1)
    return np.moveaxis(out, -1, axis)


def rank_1_factorization(Y: np.ndarray, max_iter: int = 1, divisor: float = 1.0):
    """rank_1_factorization.m:31-41 on Y [pixels, frames]: returns
    (bg_spatial [pixels], bg_temporal [frames])."""
    Y = np.asarray(Y, np.float64)
    s = np.ones(Y.shape[0]) / math.sqrt(Y.shape[0])
    t = None
    for it in range(1, max_iter + 1):
        t = (s @ Y) / divisor
        t = t / np.linalg.norm(t)
        s = (Y @ t) / divisor
        if it < max_iter:
            s = s / np.linalg.norm(s)
    return s, t


def std_image(Y: np.ndarray) -> np.ndarray:
    """compute_std_image.m with y_1 = 0, y_2 = zeros(1, T):
    sqrt(sum(Y.^2, 2) / T - mean(Y, 2).^2)."""
    Y = np.asarray(Y, np.float64)
    A = np.sum(Y * Y, axis=1)
    B = np.mean(Y, axis=1)
    with np.errstate(invalid="ignore"):
        return np.sqrt(A / Y.shape[1] - B * B)


def summary_pixelwise(movie: np.ndarray, window_1: int = 100, window_2: int = 20,
                      poly_index: int = 3, bg_iter: int = 5):
    """summary_image_generation_module.m:62-96 and :130 for movie [pixels,
    frames]: returns (summary image normalised by its max, bg_spatial,
    bg_temporal)."""
    Y = np.asarray(movie, np.float64)
    mx = float(np.max(Y))
    bs, bt = rank_1_factorization(Y, bg_iter, mx)
    x = Y / mx
    d = x - movmean(x, window_1, axis=1)
    y = movmean(d, window_2, axis=1) ** poly_index
    img = std_image(y)
    return img / matlab_max(img), bs, bt


# ---------------------------------------------------------------------------
# patch tiling -- determine_patch_info_aug.m:26-30 and :103-143 ('carpet')
# ---------------------------------------------------------------------------


def carpet_patches(global_h: int, global_w: int, n_patches, M: int = 15, min_size: int = 50):
    """Patches of the 'carpet' mode in MATLAB order (ind2sub over
    [n_patches(1), n_patches(2)], first index fastest): a list of
    ((top, left), (bottom, right)) 1-based inclusive corners; patches whose
    height or width, cut down to a multiple of M, is below min_size are
    discarded."""
    n = list(np.atleast_1d(n_patches))
    if len(n) == 1:
        n = [n[0], n[0]]
    ph = -(-(global_h // n[0]) // M) * M
    pw = -(-(global_w // n[1]) // M) * M
    out = []
    for k in range(n[0] * n[1]):
        ci, cj = k % n[0] + 1, k // n[0] + 1
        top, left = (ci - 1) * ph + 1, (cj - 1) * pw + 1
        bottom, right = min(ci * ph, global_h), min(cj * pw, global_w)
        h = bottom - top + 1
        w = right - left + 1
        h -= h % M
        w -= w % M
        if h < min_size or w < min_size:
            continue
        out.append(((top, left), (top + h - 1, left + w - 1)))
    return out


def roi_patches(roi_list, roi_patch_size: int, M: int = 15, min_size: int = 50):
    """Patches of the 'roi' mode (determine_patch_info_aug.m:32-42, 144-218):
    each ROI [top, left, bottom, right] (1-based, in the cropped rectified
    image) is tiled with patches of roi_patch_size (cut down to a multiple of
    M); an ROI smaller than a patch along an axis gives one patch of the ROI's
    size there; patches below min_size after the cut to multiples of M are
    discarded.  Same corner list format as carpet_patches."""
    ps = roi_patch_size - roi_patch_size % M
    out = []
    for roi in roi_list:
        t0, l0, b0, r0 = [int(v) for v in roi]
        sh, sw = b0 - t0 + 1, r0 - l0 + 1
        nh, nw = -(-sh // ps), -(-sw // ps)
        ph, pw = ps, ps
        if sh < ps:
            ph, nh = sh, 1
        if sw < ps:
            pw, nw = sw, 1
        for j in range(nh * nw):
            ci, cj = j % nh + 1, j // nh + 1
            top, left = (ci - 1) * ph + t0, (cj - 1) * pw + l0
            bottom, right = min(ci * ph, sh) + t0 - 1, min(cj * pw, sw) + l0 - 1
            h = bottom - top + 1
            w = right - left + 1
            h -= h % M
            w -= w % M
            if h < min_size or w < min_size:
                continue
            out.append(((top, left), (top + h - 1, left + w - 1)))
    return out


def determine_patch_info(raw_h: int, raw_w: int, x_center: float, y_center: float, dx: float, M: int,
                         crop_border, patch_mode: str = "carpet", n_patches=None, roi_list=None,
                         roi_patch_size=None):
    """determine_patch_info_aug.m:12-281 on the frame size (the pixel values
    only feed a figure there).  The reference grid of the whole frame steps
    dx/M through (x_center, y_center) from 1 to the frame size (:51-52), is
    started at a whole lenslet (:60-69), cut to whole lenslets (:82-84) and
    cropped by crop_border = [left, right, up, down] lenslets (:87-103); the
    patches tile that image (carpet or roi); for each patch, raw_cut is the
    raw-frame region [y0, y1; x0, x1] (1-based, inclusive) to read, with
    (raw_cut_xCenter, raw_cut_yCenter) the centre of a microlens in it, to be
    rectified about (:221-275).  Returns a list of dicts."""
    s = dx / M
    Xr = np.concatenate([_colon(x_center, -s, 1.0)[::-1], _colon(x_center + s, s, float(raw_w))])
    Yr = np.concatenate([_colon(y_center, -s, 1.0)[::-1], _colon(y_center + s, s, float(raw_h))])
    Mdiff = M // 2
    pre_x = len(_colon(x_center, -s, 1.0))          # find(Xresample == xCenter), 1-based
    pre_y = len(_colon(y_center, -s, 1.0))
    xq_init = (pre_x - Mdiff) - M * math.floor((pre_x - Mdiff) / M)
    yq_init = (pre_y - Mdiff) - M * math.floor((pre_y - Mdiff) / M)
    h_res = M * ((len(Yr) - yq_init) // M)
    w_res = M * ((len(Xr) - xq_init) // M)
    cl, cr, cu, cd = [int(v) for v in crop_border]
    xs_ml, ys_ml = w_res // M, h_res // M
    if cl + cr >= xs_ml:
        raise ValueError("X-cut range is larger than the x-size of image")
    if cu + cd >= ys_ml:
        raise ValueError("Y-cut range is larger than the y-size of image")
    x_first, x_last = 1 + cl, xs_ml - cr
    y_first, y_last = 1 + cu, ys_ml - cd
    gh, gw = (y_last - y_first + 1) * M, (x_last - x_first + 1) * M
    if patch_mode == "carpet":
        corners = carpet_patches(gh, gw, n_patches, M)
    elif patch_mode == "roi":
        corners = roi_patches(roi_list, roi_patch_size, M)
    else:
        raise ValueError("patch_mode should be either 'carpet' or 'roi'")
    y_top, x_left = (y_first - 1) * M + 1, (x_first - 1) * M + 1
    margin = -(-M // 2) - 2
    out = []
    for (t, l), (b, r) in corners:
        pt = t + y_top - 1 + yq_init - 1 - margin
        pl = l + x_left - 1 + xq_init - 1 - margin
        pb = b + y_top - 1 + yq_init - 1 + margin - 1
        pr = r + x_left - 1 + xq_init - 1 + margin - 1
        if pt < 1 or pl < 1 or pb > len(Yr) or pr > len(Xr):
            # MATLAB stops at Yresample(patch_top_in_global_after) (:245-248)
            raise IndexError(f"patch {(t, l, b, r)}: raw cut outside the resample grid")
        y0 = max(math.floor(Yr[pt - 1]), 1)
        y1 = min(math.ceil(Yr[pb - 1]), raw_h)
        x0 = max(math.floor(Xr[pl - 1]), 1)
        x1 = min(math.ceil(Xr[pr - 1]), raw_w)
        cy = matlab_round((pt + pb) / 2)
        cx = matlab_round((pl + pr) / 2)
        my = (pre_y - cy) % M
        mx = (pre_x - cx) % M
        ycp = Yr[cy - (M - my) - 1] if my >= -(-M // 2) else Yr[cy + my - 1]
        xcp = Xr[cx - (M - mx) - 1] if mx >= -(-M // 2) else Xr[cx + mx - 1]
        out.append({"location": [[t, l], [b, r]], "size": [b - t + 1, r - l + 1],
                    "raw_cut": [[y0, y1], [x0, x1]],
                    "raw_cut_xCenter": float(xcp - Xr[pl - 1]),
                    "raw_cut_yCenter": float(ycp - Yr[pt - 1])})
    return out


# ---------------------------------------------------------------------------
# rectification -- image_load_module_aug.m:124-164 (img_rectify_mod)
# ---------------------------------------------------------------------------


def _colon(a: float, d: float, b: float) -> np.ndarray:
    """a:d:b as a + k*d (MATLAB computes the upper half from the end; the
    difference is rounding in the last bits of the grid)."""
    n = int(math.floor((b - a) / d + 1e-10)) + 1
    return a + d * np.arange(max(n, 0))


def rectify_axis(center: float, size: int, M: int, dx: float) -> np.ndarray:
    """Query coordinates (1-based pixels) along one axis: the resample grid
    of step dx/M through `center` from -dx/M to size + 0.9, started at a whole
    lenslet (img_rectify_mod :128-148) and cut to whole lenslets (:160-162)."""
    s = dx / M
    first = _colon(center, -s, -s)[::-1]
    second = _colon(center + s, s, size + 0.9)
    grid = np.concatenate([first, second])
    q_center = len(first) - M // 2  # find(grid == center) - floor(M/2), 1-based
    q_init = q_center - M * math.floor(q_center / M)
    q = grid[q_init:]
    return q[:M * (len(q) // M)]


def rectify(img: np.ndarray, x_center: float, y_center: float, M: int, dx: float) -> np.ndarray:
    """griddedInterpolant({1:H, 1:W}, img) (linear, linear extrapolation) at
    the rectification grid: [len(yq), len(xq)] (frames along a third axis)."""
    from scipy.interpolate import RegularGridInterpolator
    img = np.asarray(img, np.float64)
    H, W = img.shape[:2]
    xq = rectify_axis(x_center, W, M, dx)
    yq = rectify_axis(y_center, H, M, dx)
    Y, X = np.meshgrid(yq, xq, indexing="ij")
    pts = np.stack([Y.ravel(), X.ravel()], axis=1)
    frames = img.reshape(H, W, -1)
    out = np.empty((len(yq), len(xq), frames.shape[2]))
    for f in range(frames.shape[2]):
        gi = RegularGridInterpolator((np.arange(1, H + 1), np.arange(1, W + 1)), frames[:, :, f],
                                     method="linear", bounds_error=False, fill_value=None)
        out[:, :, f] = gi(pts).reshape(len(yq), len(xq))
    return out.reshape((len(yq), len(xq)) + img.shape[2:])



def std_image_bg(Y: np.ndarray, y_1: np.ndarray, y_2: np.ndarray) -> np.ndarray:
    """compute_std_image.m:23-28 as written:
    sqrt((A - y_1.^2) / length(y_2) - B.^2 + 2 * B .* y_1 * C - y_1.^2 * C^2),
    A = sum(Y.^2, 2), B = mean(Y, 2), C = mean(y_2); a negative argument gives
    NaN here (MATLAB would return a complex value)."""
    Y = np.asarray(Y, np.float64)
    y_1 = np.asarray(y_1, np.float64).ravel()
    y_2 = np.asarray(y_2, np.float64).ravel()
    A = np.sum(Y * Y, axis=1)
    B = np.mean(Y, axis=1)
    C = np.mean(y_2)
    with np.errstate(invalid="ignore"):
        return np.sqrt((A - y_1 ** 2) / len(y_2) - B ** 2 + 2 * B * y_1 * C - y_1 ** 2 * C ** 2)


def summary_global(movie: np.ndarray, bg_iter: int = 5):
    """summary_image_generation_module.m:43-60 and :130 ('global'): the movie
    divided by its max, its rank-1 background (divisor 1) and
    compute_std_image with that background; the Savitzky-Golay baseline of
    :45-48 only feeds a plot.  Returns (summary / max, bg_spatial, bg_temporal)."""
    Y = np.asarray(movie, np.float64)
    Y = Y / float(np.max(Y))
    bs, bt = rank_1_factorization(Y, bg_iter)
    img = std_image_bg(Y, bs, bt)
    return img / matlab_max(img), bs, bt
