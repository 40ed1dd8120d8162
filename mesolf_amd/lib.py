"""ctypes binding of libmesolf (the C ABI in ``include/mesolf.h``).

Arrays follow MATLAB semantics: a numpy array ``A`` with ``A[x, y, ...]``
is passed in column-major (Fortran) order, so ``A[:, :, z]`` is the MATLAB
slice ``A(:, :, z)``.

The drop-in functions keep the reference's argument order and meaning:

* :func:`recon_theory_phase_space_LFM_with_filter` --
  ``src/recon/recon_theory_phase_space_LFM_with_filter.m:10``
* :func:`mex_lfm_convolution_vec` -- ``src/cuda/mex_lfm_convolution_vec.cu:276``
* :func:`LFM_reconstruction_module` -- ``src/module/LFM_reconstruction_module.m:1``
* :func:`summary_image_generation_module` --
  ``src/module/summary_image_generation_module.m:1`` (detrend mode 'pixelwise')

There is no CPU fallback: if the shared library is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MLF_LIBRARY", os.path.join(HERE, "libmesolf.so"))

MLF_OK = 0
MLF_ERR_INVALID_ARGUMENT = 1
MLF_ERR_UNSUPPORTED = 2
MLF_ERR_DEVICE = 3
MLF_ERR_OUT_OF_MEMORY = 4
MLF_ERR_BUDGET = 5
MLF_ERR_INTERNAL = 6
MLF_ERR_CANCELLED = 7
MLF_ERR_IO = 8

# host-forced server failures (mlf_server_inject_fault)
MLF_FAULT_ADMIT = 1
MLF_FAULT_ITERATION = 2

PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}

# error identifier of the reference gateway (mex:296)
MEX_ERROR_ID = "parallel:gpu:mexGPUExample:InvalidInput"
MEX_ERROR_MSG = "Invalid input to mex_lfm_convolution_vec()"


class MesolfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


class InvalidInput(MesolfError, ValueError):
    """Raised for MLF_ERR_INVALID_ARGUMENT; carries the reference error id."""

    identifier = MEX_ERROR_ID


class Unsupported(MesolfError):
    pass


class BudgetExceeded(MesolfError):
    pass


class Cancelled(MesolfError):
    """A queued server request dropped before it ran (server destroyed)."""


class MatFileError(MesolfError, OSError):
    """MLF_ERR_IO: a MAT-file that is missing, unreadable or malformed."""


_lock = threading.Lock()
_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_F32P = C.POINTER(C.c_float)
_F64P = C.POINTER(C.c_double)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U64P = C.POINTER(C.c_uint64)

_SIGNATURES = {
    "mlf_last_error": (C.c_char_p, []),
    "mlf_status_string": (C.c_char_p, [C.c_int]),
    "mlf_abi_version": (_I32, []),
    "mlf_context_create": (C.c_int, [_I32, C.POINTER(_P)]),
    "mlf_context_destroy": (C.c_int, [_P]),
    "mlf_context_synchronize": (C.c_int, [_P]),
    "mlf_context_memory": (C.c_int, [_P, _U64P, _U64P]),
    "mlf_psf_create_dense": (C.c_int, [_P, _F32P, _F32P, _I64, _I64, _I64, C.POINTER(_P)]),
    "mlf_psf_create_compact": (C.c_int, [_P, _I64, _I64, _I64, _I32P, _F64P, _I32P, _F32P,
                                         _I32P, _F32P, C.POINTER(_P)]),
    "mlf_psf_info": (C.c_int, [_P, _I64P, _I64P, _I32P, _I32P]),
    "mlf_psf_views": (C.c_int, [_P, _I32P, _F64P]),
    "mlf_psf_destroy": (C.c_int, [_P]),
    "mlf_patch_create": (C.c_int, [_P, _I64, _I64, C.c_int, C.POINTER(_P)]),
    "mlf_patch_destroy": (C.c_int, [_P]),
    "mlf_patch_set_image": (C.c_int, [_P, _F32P]),
    "mlf_patch_init_volume": (C.c_int, [_P]),
    "mlf_patch_set_volume": (C.c_int, [_P, _F32P]),
    "mlf_patch_get_volume": (C.c_int, [_P, _F32P]),
    "mlf_patch_get_volume_f64": (C.c_int, [_P, _F64P]),
    "mlf_patch_run": (C.c_int, [_P, _I32, _I32, _I32]),
    "mlf_patch_forward_view": (C.c_int, [_P, _I32, _F32P]),
    "mlf_patch_project_all": (C.c_int, [_P, _F32P]),
    "mlf_patch_sample_filter": (C.c_int, [_P]),
    "mlf_patch_postprocess": (C.c_int, [_P, _I32, _F64P, C.POINTER(C.c_uint16)]),
    "mlf_patch_last_run_ms": (C.c_int, [_P, _F64P]),
    "mlf_patch_set_timing": (C.c_int, [_P, _I32]),
    "mlf_patch_kernel_times": (C.c_int, [_P, _F64P, _I64P]),
    "mlf_batch_create": (C.c_int, [_P, C.c_int, _I32, C.POINTER(_P)]),
    "mlf_batch_destroy": (C.c_int, [_P]),
    "mlf_batch_attach": (C.c_int, [_P, _P]),
    "mlf_batch_detach": (C.c_int, [_P, _P]),
    "mlf_batch_size": (C.c_int, [_P, _I32P]),
    "mlf_batch_set_graph": (C.c_int, [_P, _I32]),
    "mlf_batch_run": (C.c_int, [_P, _I32, _I32]),
    "mlf_batch_last_run_ms": (C.c_int, [_P, _F64P]),
    "mlf_patch_autotune": (C.c_int, [_P, _I32, _P]),
    "mlf_patch_tile_shape": (C.c_int, [_P, _P]),
    "mlf_patch_set_tile_shape": (C.c_int, [_P, _P]),
    "mlf_batch_autotune": (C.c_int, [_P, _I32, _P]),
    "mlf_server_create": (C.c_int, [_P, C.c_int, _P, C.POINTER(_P)]),
    "mlf_server_submit": (C.c_int, [_P, _I32, _I64, _I64, _F32P, _F32P, _U64P]),
    "mlf_server_wait": (C.c_int, [_P, C.c_uint64, _F64P]),
    "mlf_server_drain": (C.c_int, [_P]),
    "mlf_server_get_stats": (C.c_int, [_P, _P]),
    "mlf_server_destroy": (C.c_int, [_P]),
    "mlf_server_inject_fault": (C.c_int, [_P, _I32, _I32]),
    "mlf_recon_phase_space": (C.c_int, [_F32P, _I64, _I64, _F32P, _F32P, _I64, _I64, _I64, _I32,
                                        _I32, _I32, C.c_int, _F64P, _P]),
    "mlf_lfm_convolution_vec": (C.c_int, [_F32P, _I64, _I64, _I64, _F32P, _I64, _I64, _F32P, _I32]),
    "mlf_lfm_convolution_vec_device": (C.c_int, [_P, _P, _I64, _I64, _I64, _P, _I64, _I64, _P]),
    "mlf_lfm_convolution_vec_noalloc": (C.c_int, [_P, _P, _I64, _I64, _I64, _P, _I64, _I64, _P]),
    "mlf_rectify_size": (C.c_int, [_I64, _I64, C.c_double, C.c_double, _I32, C.c_double, _I64P, _I64P]),
    "mlf_rectify": (C.c_int, [_F32P, _I64, _I64, _I64, C.c_double, C.c_double, _I32, C.c_double, _F32P,
                              _I32]),
    "mlf_rectify_device": (C.c_int, [_P, _P, _I64, _I64, _I64, C.c_double, C.c_double, _I32, C.c_double,
                                     _P]),
    "mlf_carpet_patches": (C.c_int, [_I64, _I64, _I32, _I32, _I32, _I32, _I64P, _I32P]),
    "mlf_crop_movie_device": (C.c_int, [_P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P]),
    "mlf_determine_patch_info": (C.c_int, [_I64, _I64, C.c_double, C.c_double, C.c_double, _I32, _I32P, _I32,
                                           _I32P, _I64P, _I32, _I32, _P, _I32, _I32P]),
    "mlf_summary_image": (C.c_int, [_F32P, _I64, _I64, _I64, _I32, _I32, _I32, _I32, _F64P, _F64P,
                                    _F64P, _I32]),
    "mlf_summary_image_global": (C.c_int, [_F32P, _I64, _I64, _I64, _I32, _F64P, _F64P, _F64P, _I32]),
    "mlf_summary_image_global_device": (C.c_int, [_P, _P, _I64, _I64, _I64, _I32, _P, _P, _P]),
    "mlf_summary_image_device": (C.c_int, [_P, _P, _I64, _I64, _I64, _I32, _I32, _I32, _I32, _P, _P,
                                           _P]),
    "mlf_lfm_convolution_segments": (C.c_int, [_F32P, _I64, _I64, _I64, _I32P, _I64P, _I64, _F32P,
                                               _I64, _I64, _I64, _F32P, _I32]),
    "mlf_lfm_convolution_segments_device": (C.c_int, [_P, _P, _I64, _I64, _I64, _I32P, _I64P, _I64, _P,
                                                      _I64, _I64, _I64, _P]),
    "mlf_mat_open": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mlf_mat_close": (C.c_int, [_P]),
    "mlf_mat_count": (C.c_int, [_P, _I32P]),
    "mlf_mat_name": (C.c_int, [_P, _I32, C.c_char_p, C.c_size_t]),
    "mlf_mat_info": (C.c_int, [_P, C.c_char_p, _P]),
    "mlf_mat_read_f32": (C.c_int, [_P, C.c_char_p, _F32P, _I64]),
    "mlf_mat_read_f64": (C.c_int, [_P, C.c_char_p, _F64P, _I64]),
    "mlf_psf_load_mat": (C.c_int, [_P, C.c_char_p, C.POINTER(_P), _P]),
}


def exported_symbols():
    return sorted(_SIGNATURES)


def load_library(path: str | None = None):
    """Load libmesolf.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise MesolfError(MLF_ERR_INTERNAL,
                              f"{p} not found: build it with `python -m mesolf_amd.build`")
        lib = C.CDLL(p)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def _check(status: int):
    if status == MLF_OK:
        return
    lib = load_library()
    msg = (lib.mlf_last_error() or b"").decode(errors="replace")
    if status == MLF_ERR_INVALID_ARGUMENT:
        raise InvalidInput(status, msg)
    if status == MLF_ERR_UNSUPPORTED:
        raise Unsupported(status, msg)
    if status == MLF_ERR_BUDGET:
        raise BudgetExceeded(status, msg)
    if status == MLF_ERR_CANCELLED:
        raise Cancelled(status, msg)
    if status == MLF_ERR_IO:
        raise MatFileError(status, msg)
    raise MesolfError(status, msg)


def _fptr(a: np.ndarray, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype))


def _col(a, dtype):
    """Column-major contiguous copy (no copy if already so)."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# --------------------------------------------------------------------- handles
class Context:
    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = _P()
        _check(self._lib.mlf_context_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        _check(self._lib.mlf_context_synchronize(self.handle))

    def memory(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(self._lib.mlf_context_memory(self.handle, C.byref(f), C.byref(t)))
        return f.value, t.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _MatVar(C.Structure):
    _fields_ = [("matlab_class", C.c_char * 32), ("ndim", C.c_int32), ("dims", C.c_int64 * 32),
                ("numel", C.c_int64), ("numeric", C.c_int32), ("stored_bytes", C.c_int32)]


class _PsfFileInfo(C.Structure):
    _fields_ = [("S", C.c_int64), ("N", C.c_int64), ("Z", C.c_int64), ("x_pixel_um", C.c_double),
                ("z_pixel_um", C.c_double)]


class MatFile:
    """A MATLAB -v7.3 MAT-file opened by the native reader (no HDF5 library).

    ``read(name)`` returns the variable as a Fortran-ordered numpy array of
    MATLAB's shape (float64 by default, float32 with ``dtype=np.float32``);
    ``info(name)`` returns its class, shape and element count.  Names may be
    '/'-separated paths into structs."""

    def __init__(self, path: str):
        self._lib = load_library()
        h = _P()
        _check(self._lib.mlf_mat_open(os.fsencode(path), C.byref(h)))
        self.handle = h
        self.path = path

    def variables(self):
        n = C.c_int32()
        _check(self._lib.mlf_mat_count(self.handle, C.byref(n)))
        out = []
        buf = C.create_string_buffer(4096)
        for i in range(n.value):
            _check(self._lib.mlf_mat_name(self.handle, i, buf, len(buf)))
            out.append(buf.value.decode())
        return out

    def info(self, name: str):
        v = _MatVar()
        _check(self._lib.mlf_mat_info(self.handle, name.encode(), C.byref(v)))
        return {"class": v.matlab_class.decode(), "shape": tuple(v.dims[:v.ndim]), "numel": v.numel,
                "numeric": bool(v.numeric), "stored_bytes": v.stored_bytes}

    def read(self, name: str, dtype=np.float64):
        inf = self.info(name)
        dtype = np.dtype(dtype)
        out = np.empty(inf["shape"] if inf["numel"] else (0,), dtype=dtype, order="F")
        if dtype == np.float64:
            _check(self._lib.mlf_mat_read_f64(self.handle, name.encode(), _fptr(out, C.c_double), inf["numel"]))
        elif dtype == np.float32:
            _check(self._lib.mlf_mat_read_f32(self.handle, name.encode(), _fptr(out), inf["numel"]))
        else:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "dtype must be float32 or float64")
        return out

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_mat_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compact_block_layout(tab: np.ndarray) -> np.ndarray:
    """[d1+r, d2+r, s1, s2] table -> flat ABI order ((d2*K + d1)*225 + s2*15 + s1)."""
    return np.ascontiguousarray(np.transpose(tab, (1, 0, 3, 2)), dtype=np.float32).ravel()


class PSF:
    """Device-resident phase-space PSF tables (one handle per PSF)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._lib = ctx._lib
        self.handle = handle
        nv, z = C.c_int64(), C.c_int64()
        _check(self._lib.mlf_psf_info(handle, C.byref(nv), C.byref(z), None, None))
        self.nviews, self.Z = nv.value, z.value

    @classmethod
    def from_dense(cls, ctx: Context, H, Ht):
        """H, Ht [S, S, 15, 15, Z] raw PSFs (recon:13-40 preparation)."""
        H = _col(H, np.float32)
        Ht = _col(Ht, np.float32)
        if H.ndim == 4:
            H = H[..., None]
            Ht = Ht[..., None]
        if H.ndim != 5 or H.shape != Ht.shape or H.shape[0] != H.shape[1]:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "H and Ht must be [S,S,N,N,Z] and equal")
        h = _P()
        _check(ctx._lib.mlf_psf_create_dense(ctx.handle, _fptr(H), _fptr(Ht), H.shape[0],
                                             H.shape[2], H.shape[4], C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_mat(cls, ctx: Context, path: str):
        """load(SI.psffile) (mesolf.m:20-21) of a -v7.3 PSF file, then
        :meth:`from_dense` on its H and Ht, in native code.  Returns the PSF;
        ``psf.file_info`` holds S, N, Z and the pixel sizes in um
        (mesolf.m:231-232; NaN when x1objspace / x3objspace are absent)."""
        h = _P()
        info = _PsfFileInfo()
        _check(ctx._lib.mlf_psf_load_mat(ctx.handle, os.fsencode(path), C.byref(h), C.byref(info)))
        psf = cls(ctx, h)
        psf.file_info = {"S": info.S, "N": info.N, "Z": info.Z, "x_pixel_um": info.x_pixel_um,
                         "z_pixel_um": info.z_pixel_um}
        return psf

    @classmethod
    def from_compact(cls, ctx: Context, views_uv, weights, rf, fwd_flat, rb=None, bwd_flat=None):
        """Compact tables in ABI order (see ``compact_block_layout``)."""
        uv = np.ascontiguousarray(np.asarray(views_uv, dtype=np.int32).reshape(-1, 2))
        w = np.ascontiguousarray(weights, dtype=np.float64)
        rf = np.ascontiguousarray(rf, dtype=np.int32)
        fwd = np.ascontiguousarray(fwd_flat, dtype=np.float32)
        rbp = None if rb is None else np.ascontiguousarray(rb, dtype=np.int32)
        bwd = None if bwd_flat is None else np.ascontiguousarray(bwd_flat, dtype=np.float32)
        h = _P()
        _check(ctx._lib.mlf_psf_create_compact(
            ctx.handle, 15, rf.size, uv.shape[0], _fptr(uv, C.c_int32), _fptr(w, C.c_double),
            _fptr(rf, C.c_int32), _fptr(fwd), None if rbp is None else _fptr(rbp, C.c_int32),
            None if bwd is None else _fptr(bwd), C.byref(h)))
        return cls(ctx, h)

    def radii(self):
        rf = np.zeros(self.Z, np.int32)
        rb = np.zeros(self.Z, np.int32)
        _check(self._lib.mlf_psf_info(self.handle, None, None, _fptr(rf, C.c_int32), _fptr(rb, C.c_int32)))
        return rf, rb

    def views(self):
        uv = np.zeros((self.nviews, 2), np.int32)
        w = np.zeros(self.nviews, np.float64)
        _check(self._lib.mlf_psf_views(self.handle, _fptr(uv, C.c_int32), _fptr(w, C.c_double)))
        return [tuple(x) for x in uv.tolist()], w

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_psf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Patch:
    """One reconstruction problem (n1 x n2 lenslets, all PSF planes)."""

    def __init__(self, psf: PSF, n1: int, n2: int, precision: str = "fp32"):
        self.psf = psf
        self._lib = psf._lib
        self.n1, self.n2, self.Z = int(n1), int(n2), psf.Z
        h = _P()
        _check(self._lib.mlf_patch_create(psf.handle, self.n1, self.n2, PRECISIONS[precision], C.byref(h)))
        self.handle = h

    @property
    def shape(self):
        return (15 * self.n1, 15 * self.n2, self.Z)

    def set_image(self, lf):
        a = _col(lf, np.float32)
        if a.shape != self.shape[:2]:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"image shape {a.shape} != {self.shape[:2]}")
        _check(self._lib.mlf_patch_set_image(self.handle, _fptr(a)))

    def init_volume(self):
        _check(self._lib.mlf_patch_init_volume(self.handle))

    def set_volume(self, X):
        a = _col(X, np.float32)
        if a.shape != self.shape:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"volume shape {a.shape} != {self.shape}")
        _check(self._lib.mlf_patch_set_volume(self.handle, _fptr(a)))

    def get_volume(self, f64: bool = False):
        if f64:
            out = np.empty(self.shape, np.float64, order="F")
            _check(self._lib.mlf_patch_get_volume_f64(self.handle, _fptr(out, C.c_double)))
        else:
            out = np.empty(self.shape, np.float32, order="F")
            _check(self._lib.mlf_patch_get_volume(self.handle, _fptr(out)))
        return out

    def run(self, iterations: int, views_limit: int = 0, filter: bool = True):
        _check(self._lib.mlf_patch_run(self.handle, int(iterations), int(views_limit), int(bool(filter))))

    def synchronize(self):
        self.psf.ctx.synchronize()

    def last_run_ms(self):
        ms = C.c_double()
        _check(self._lib.mlf_patch_last_run_ms(self.handle, C.byref(ms)))
        return ms.value

    KERNEL_CLASSES = ("forward", "ratio", "backward_update", "sample_filter", "postprocess")

    def set_timing(self, enable: bool = True, stride: int = 1):
        """Bracket launches with HIP events: every `stride`-th view step (the
        sample filter always) while enabled."""
        if stride < 1:
            raise ValueError("stride must be >= 1")
        _check(self._lib.mlf_patch_set_timing(self.handle, int(stride) if enable else 0))

    def kernel_times(self):
        """{class: (total_ms, launches)} since the previous read."""
        ms = np.zeros(len(self.KERNEL_CLASSES), np.float64)
        n = np.zeros(len(self.KERNEL_CLASSES), np.int64)
        _check(self._lib.mlf_patch_kernel_times(self.handle, _fptr(ms, C.c_double), _fptr(n, C.c_int64)))
        return {c: (float(ms[i]), int(n[i])) for i, c in enumerate(self.KERNEL_CLASSES)}

    def autotune(self, views: int = 8):
        """Time every tile shape on `views` view steps; keeps the fastest."""
        t = _TileShape()
        _check(self._lib.mlf_patch_autotune(self.handle, int(views), C.byref(t)))
        return _shape_dict(t)

    def tile_shape(self):
        """(forward columns per wave, backprojection phase tiles, phase groups
        per backprojection entry (0: all), backprojection waves per SIMD)."""
        t = _TileShape()
        _check(self._lib.mlf_patch_tile_shape(self.handle, C.byref(t)))
        return (t.fwd_columns_per_wave, t.bwd_phase_tiles, t.bwd_groups_per_entry, t.bwd_waves_per_simd)

    def set_tile_shape(self, fwd_columns_per_wave: int, bwd_phase_tiles: int, bwd_groups_per_entry: int = 0,
                       bwd_waves_per_simd: int = 2):
        t = _TileShape(int(fwd_columns_per_wave), int(bwd_phase_tiles), int(bwd_groups_per_entry),
                       int(bwd_waves_per_simd), 0.0)
        _check(self._lib.mlf_patch_set_tile_shape(self.handle, C.byref(t)))

    def forward_view(self, k: int):
        out = np.empty((self.n1, self.n2), np.float32, order="F")
        _check(self._lib.mlf_patch_forward_view(self.handle, int(k), _fptr(out)))
        return out

    def project_all(self):
        out = np.empty(self.shape[:2], np.float32, order="F")
        _check(self._lib.mlf_patch_project_all(self.handle, _fptr(out)))
        return out

    def sample_filter(self):
        _check(self._lib.mlf_patch_sample_filter(self.handle))

    def postprocess(self, valid_range: int, stack: bool = True, tiff: bool = True):
        """Post-reconstruction step of LFM_reconstruction_module.m:45-53,60-62
        and the central-plane crop of mesolf.m:101-102, on the device volume.

        Returns (stack, tiff): the normalised stack's central `valid_range`
        planes (all Z planes for valid_range = 0), float64 [I1, I2, planes],
        and the uint16 image of the whole normalised stack [I1, I2, Z] (what
        the module writes with saveastiff); either is None when not asked for.
        """
        I1, I2, Z = self.shape
        nz = int(valid_range) if valid_range else Z
        if nz < 0:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"valid_range {valid_range} < 0")
        out = np.empty((I1, I2, nz), np.float64, order="F") if stack else None
        img = np.empty((I1, I2, Z), np.uint16, order="F") if tiff else None
        _check(self._lib.mlf_patch_postprocess(
            self.handle, int(valid_range),
            _fptr(out, C.c_double) if stack else None,
            img.ctypes.data_as(C.POINTER(C.c_uint16)) if tiff else None))
        return out, img

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_patch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------------- drop-ins
class _TileShape(C.Structure):
    _fields_ = [("fwd_columns_per_wave", C.c_int32), ("bwd_phase_tiles", C.c_int32),
                ("bwd_groups_per_entry", C.c_int32), ("bwd_waves_per_simd", C.c_int32),
                ("ms_per_view", C.c_double)]


def _shape_dict(t):
    return {"fwd_columns_per_wave": t.fwd_columns_per_wave, "bwd_phase_tiles": t.bwd_phase_tiles,
            "bwd_groups_per_entry": t.bwd_groups_per_entry, "bwd_waves_per_simd": t.bwd_waves_per_simd,
            "ms_per_view": t.ms_per_view}


class Batch:
    """Several patches of one PSF run in shared launches (``mlf_batch_*``).

    A patch's result does not depend on the batch it ran in."""

    def __init__(self, psf: PSF, capacity: int, precision: str = "bf16", graph: bool = True):
        self.psf = psf
        self._lib = psf._lib
        h = _P()
        _check(self._lib.mlf_batch_create(psf.handle, PRECISIONS[precision], int(capacity), C.byref(h)))
        self.handle = h
        self.set_graph(graph)

    def set_graph(self, enable: bool):
        _check(self._lib.mlf_batch_set_graph(self.handle, int(bool(enable))))

    def attach(self, patch: Patch):
        _check(self._lib.mlf_batch_attach(self.handle, patch.handle))

    def detach(self, patch: Patch):
        _check(self._lib.mlf_batch_detach(self.handle, patch.handle))

    def __len__(self):
        n = C.c_int32()
        _check(self._lib.mlf_batch_size(self.handle, C.byref(n)))
        return n.value

    def run(self, iterations: int, filter: bool = True):
        _check(self._lib.mlf_batch_run(self.handle, int(iterations), int(bool(filter))))

    def autotune(self, views: int = 8):
        """Time every tile shape on the attached patches; keeps the fastest."""
        t = _TileShape()
        _check(self._lib.mlf_batch_autotune(self.handle, int(views), C.byref(t)))
        return _shape_dict(t)

    def last_run_ms(self):
        ms = C.c_double()
        _check(self._lib.mlf_batch_last_run_ms(self.handle, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ServerConfig(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("iterations", C.c_int32), ("filter", C.c_int32),
                ("use_graph", C.c_int32), ("hbm_budget_bytes", C.c_uint64)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("pending", C.c_int64),
                ("in_flight", C.c_int64), ("resident_patches", C.c_int64),
                ("evictions", C.c_int64), ("cold_starts", C.c_int64),
                ("iterations_run", C.c_int64), ("resident_bytes", C.c_uint64),
                ("peak_resident_bytes", C.c_uint64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("volumes_per_s", C.c_double), ("mean_slots", C.c_double)]


class Server:
    """Continuous-batching request loop (``mlf_server_*``).

    ``submit`` queues one frame of a patch and returns a ticket; results are
    written into the caller's array (if given) when the request completes."""

    def __init__(self, psf: PSF, precision: str = "bf16", max_slots: int = 8, iterations: int = 3,
                 hbm_budget_bytes: int = 64 << 30, filter: bool = True, graph: bool = True):
        self.psf = psf
        self._lib = psf._lib
        cfg = _ServerConfig(int(max_slots), int(iterations), int(bool(filter)), int(bool(graph)),
                            int(hbm_budget_bytes))
        h = _P()
        _check(self._lib.mlf_server_create(psf.handle, PRECISIONS[precision], C.byref(cfg), C.byref(h)))
        self.handle = h
        self._outs = {}  # ticket -> output array kept alive until completion

    def submit(self, patch_id: int, lfimg, want_volume: bool = False):
        a = _col(lfimg, np.float32)
        if a.ndim != 2 or a.shape[0] % 15 or a.shape[1] % 15:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "image must be [15*n1, 15*n2]")
        n1, n2 = a.shape[0] // 15, a.shape[1] // 15
        out = np.empty((a.shape[0], a.shape[1], self.psf.Z), np.float32, order="F") if want_volume else None
        t = C.c_uint64()
        _check(self._lib.mlf_server_submit(self.handle, int(patch_id), n1, n2, _fptr(a),
                                           None if out is None else _fptr(out), C.byref(t)))
        self._outs[t.value] = out
        return t.value

    def wait(self, ticket: int):
        """(latency_ms, volume or None)."""
        ms = C.c_double()
        try:
            _check(self._lib.mlf_server_wait(self.handle, C.c_uint64(ticket), C.byref(ms)))
        finally:
            out = self._outs.pop(ticket, None)
        return ms.value, out

    def drain(self):
        _check(self._lib.mlf_server_drain(self.handle))

    def inject_fault(self, kind: str, patch_id: int = -1):
        """Force a failure from the host (failure-path tests): kind "admit"
        fails the admission of patch_id's next request, "iteration" fails the
        next batched iteration (every request then in flight)."""
        k = {"admit": MLF_FAULT_ADMIT, "iteration": MLF_FAULT_ITERATION}[kind]
        _check(self._lib.mlf_server_inject_fault(self.handle, k, int(patch_id)))

    def stats(self):
        st = _ServerStats()
        _check(self._lib.mlf_server_get_stats(self.handle, C.byref(st)))
        return {name: getattr(st, name) for name, _ in _ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mlf_server_destroy(self.handle)
            self.handle = None
            self._outs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def recon_theory_phase_space_LFM_with_filter(LFIMG, H, Ht, maxIter, gpuFlag=True, gpu_id=1,
                                             precision="fp32", stats=None):
    """Drop-in for recon_theory_phase_space_LFM_with_filter.m:10.

    ``gpu_id`` is 1-based like MATLAB's ``gpuDevice(gpu_id)`` (recon:76).
    Returns Xguess as a float64 [I1, I2, Z] array (recon:147 gathers a double).
    """
    lib = load_library()
    lf = _col(LFIMG, np.float32)
    Hc = _col(H, np.float32)
    Htc = _col(Ht, np.float32)
    if lf.ndim != 2 or Hc.ndim != 5 or Hc.shape != Htc.shape:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "LFIMG must be 2-D and H, Ht [S,S,N,N,Z]")
    Z = Hc.shape[4]
    out = np.empty((lf.shape[0], lf.shape[1], Z), np.float64, order="F")
    st = (C.c_byte * 2048)()
    _check(lib.mlf_recon_phase_space(_fptr(lf), lf.shape[0], lf.shape[1], _fptr(Hc), _fptr(Htc),
                                     Hc.shape[0], Hc.shape[2], Z, int(maxIter), int(bool(gpuFlag)),
                                     int(gpu_id) - 1, PRECISIONS[precision], _fptr(out, C.c_double),
                                     C.cast(st, C.c_void_p)))
    if stats is not None:
        raw = bytes(st)
        it, vpi = np.frombuffer(raw[:8], np.int32)
        stats["iterations"] = int(it)
        stats["views_per_iteration"] = int(vpi)
        stats["total_seconds"] = float(np.frombuffer(raw[8:16], np.float64)[0])
        stats["iteration_seconds"] = np.frombuffer(raw[16:16 + 8 * 64], np.float64)[:it].tolist()
        stats["relative_change"] = np.frombuffer(raw[16 + 8 * 64:16 + 16 * 64], np.float64)[:it].tolist()
    return out


def _si_get(SI, name, default=None):
    if isinstance(SI, dict):
        return SI.get(name, default)
    return getattr(SI, name, default)


def LFM_reconstruction_module(summary_image, H, Ht, SI, precision="fp32", valid_range=0):
    """Drop-in for LFM_reconstruction_module.m:1-62.

    Reads SI.recon_max_iter, SI.gpu_ids (1-based), SI.worker_ix,
    SI.recon_mode and SI.vessel_mask_enable like the module (:19-33).  Runs the
    phase-space reconstruction (:37-39), then the top-layer balance and the
    per-layer and global max normalisation (:45-53, :61) on the device.
    Returns (recon_stack, tiff): float64 [I1, I2, Z] and the uint16 image the
    module writes with saveastiff (:62), returned rather than written.  With
    valid_range > 0 the stack is cropped to the central planes as mesolf.m:101-102
    does after the module returns.
    """
    mode = _si_get(SI, "recon_mode", "phase_space")
    if mode == "phase_space_peeling":
        raise Unsupported(MLF_ERR_UNSUPPORTED, "recon_mode 'phase_space_peeling' is not built")
    if mode != "phase_space":
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "Unsupported value of SI.recon_mode")
    if _si_get(SI, "vessel_mask_enable", False):
        raise Unsupported(MLF_ERR_UNSUPPORTED, "the blood-vessel mask (BCOSFIRE) is not built")
    gpu_ids = list(np.atleast_1d(_si_get(SI, "gpu_ids", [])))
    if not gpu_ids:
        raise Unsupported(MLF_ERR_UNSUPPORTED, "the CPU branch (empty SI.gpu_ids) is not built")
    wix = _si_get(SI, "worker_ix", None)
    gpu_id = int(gpu_ids[int(wix) % len(gpu_ids)]) if wix is not None else int(gpu_ids[0])
    lf = _col(summary_image, np.float32)
    Hc = _col(H, np.float32)
    Htc = _col(Ht, np.float32)
    if lf.ndim != 2 or Hc.ndim != 5 or Hc.shape != Htc.shape:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "summary_image must be 2-D and H, Ht [S,S,N,N,Z]")
    N = Hc.shape[2]
    if lf.shape[0] % N or lf.shape[1] % N:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"image {lf.shape} is not whole lenslets of {N} px")
    with Context(gpu_id - 1) as ctx:
        psf = PSF.from_dense(ctx, Hc, Htc)
        try:
            p = Patch(psf, lf.shape[0] // N, lf.shape[1] // N, precision=precision)
            try:
                p.set_image(lf)
                p.init_volume()
                p.run(int(_si_get(SI, "recon_max_iter")))
                return p.postprocess(valid_range)
            finally:
                p.close()
        finally:
            psf.close()


def rectify_size(H: int, W: int, xCenter: float, yCenter: float, M: int, dx: float):
    """Rectified frame size (rows, columns) of img_rectify_mod (host only)."""
    lib = load_library()
    h, w = C.c_int64(), C.c_int64()
    _check(lib.mlf_rectify_size(int(H), int(W), float(xCenter), float(yCenter), int(M), float(dx),
                                C.byref(h), C.byref(w)))
    return h.value, w.value


def img_rectify_mod(IMG_BW, xCenter, yCenter, M, dx, device: int = 0):
    """Drop-in for img_rectify_mod (image_load_module_aug.m:124-164) on one
    frame [H, W] or a stack [H, W, n]: single output of the rectified size."""
    lib = load_library()
    a = _col(IMG_BW, np.float32)
    H, W = a.shape[:2]
    n = int(np.prod(a.shape[2:])) if a.ndim > 2 else 1
    h, w = rectify_size(H, W, xCenter, yCenter, M, dx)
    out = np.empty((h, w) + a.shape[2:], np.float32, order="F")
    _check(lib.mlf_rectify(_fptr(a), H, W, n, float(xCenter), float(yCenter), int(M), float(dx),
                           _fptr(out), int(device)))
    return out


def carpet_patch_info(global_h: int, global_w: int, n_patches, M: int = 15, min_size: int = 50):
    """The 'carpet' patches of determine_patch_info_aug.m:103-143 as patch_info
    dicts: {"location": [[top, left], [bottom, right]] (1-based, inclusive),
    "size": [h, w]}, in MATLAB's ind2sub order."""
    lib = load_library()
    n = list(np.atleast_1d(n_patches))
    if len(n) == 1:
        n = [n[0], n[0]]
    corners = np.zeros((int(n[0]) * int(n[1]), 4), np.int64)
    cnt = C.c_int32()
    _check(lib.mlf_carpet_patches(int(global_h), int(global_w), int(n[0]), int(n[1]), int(M),
                                  int(min_size), corners.ctypes.data_as(_I64P), C.byref(cnt)))
    out = []
    for t, l, b, r in corners[:cnt.value].tolist():
        out.append({"location": [[t, l], [b, r]], "size": [b - t + 1, r - l + 1]})
    return out


class _PatchInfo(C.Structure):
    _fields_ = [("location", C.c_int64 * 4), ("size", C.c_int64 * 2), ("raw_cut", C.c_int64 * 4),
                ("raw_cut_xCenter", C.c_double), ("raw_cut_yCenter", C.c_double)]


def determine_patch_info_aug(SI, raw_size):
    """Drop-in for determine_patch_info_aug.m:1-281 given the raw frame size
    (rows, columns) -- the reference reads the recording's first TIFF for it.

    Reads SI.frames_x_offset, frames_y_offset, frames_dx,
    frames_crop_border_microlenses (default [0 0 0 0]), Nnum, patch_mode
    ('carpet' with n_patches, or 'roi' with ROI_list and ROI_patch_size).
    Returns the patch_info list: dicts with "location" [[top, left],
    [bottom, right]] and "size" [h, w] in the cropped rectified image,
    "raw_cut" [[row0, row1], [col0, col1]] (1-based, inclusive) of the raw
    frame and "raw_cut_xCenter" / "raw_cut_yCenter", the microlens centre in
    the cut for img_rectify_mod."""
    lib = load_library()
    mode = _si_get(SI, "patch_mode", "carpet")
    M = int(_si_get(SI, "Nnum", 15))
    crop = np.ascontiguousarray(np.ravel(_si_get(SI, "frames_crop_border_microlenses", [0, 0, 0, 0])),
                                dtype=np.int32)
    if crop.size != 4:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "frames_crop_border_microlenses must have 4 entries")
    npat = np.zeros(2, np.int32)
    roi = np.zeros((1, 4), np.int64)
    n_roi, roi_ps = 0, 0
    if mode == "carpet":
        n = list(np.atleast_1d(_si_get(SI, "n_patches")))
        npat[:] = [n[0], n[0]] if len(n) == 1 else n[:2]
        m = 0
    elif mode == "roi":
        if _si_get(SI, "ROI_patch_size") is None:
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "In ROI mode, SI.ROI_patch_size is required")
        roi = np.ascontiguousarray(np.atleast_2d(_si_get(SI, "ROI_list")), dtype=np.int64)
        n_roi, roi_ps = roi.shape[0], int(_si_get(SI, "ROI_patch_size"))
        m = 1
    else:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "patch_mode should be either 'carpet' or 'roi'")
    args = (int(raw_size[0]), int(raw_size[1]), float(_si_get(SI, "frames_x_offset")),
            float(_si_get(SI, "frames_y_offset")), float(_si_get(SI, "frames_dx")), M,
            crop.ctypes.data_as(_I32P), m, npat.ctypes.data_as(_I32P), roi.ctypes.data_as(_I64P), n_roi, roi_ps)
    cnt = C.c_int32()
    _check(lib.mlf_determine_patch_info(*args, None, 0, C.byref(cnt)))
    arr = (_PatchInfo * max(1, cnt.value))()
    _check(lib.mlf_determine_patch_info(*args, C.cast(arr, C.c_void_p), cnt.value, C.byref(cnt)))
    out = []
    for p in arr[:cnt.value]:
        t, l, b, r = list(p.location)
        y0, y1, x0, x1 = list(p.raw_cut)
        out.append({"location": [[t, l], [b, r]], "size": list(p.size), "raw_cut": [[y0, y1], [x0, x1]],
                    "raw_cut_xCenter": p.raw_cut_xCenter, "raw_cut_yCenter": p.raw_cut_yCenter})
    return out


def crop_movie_device(ctx: "Context", d_movie: int, H: int, W: int, frames: int, patch_info,
                      d_out: int):
    """Patch movie sensor_movie(top:bottom, left:right, :) on the device: raw
    device pointers (ints) to a single [H, W, frames] movie and a
    [h, w, frames] output; asynchronous on ctx's stream."""
    (t, l), (b, r) = patch_info["location"]
    _check(ctx._lib.mlf_crop_movie_device(ctx.handle, C.c_void_p(d_movie), int(H), int(W), int(frames),
                                          int(t), int(l), int(b - t + 1), int(r - l + 1),
                                          C.c_void_p(d_out)))


def summary_image_pixelwise(movie, height: int, width: int, window_1: int = 100, window_2: int = 20,
                            poly_index: int = 3, bg_iter: int = 5, device: int = 0):
    """Summary image and rank-1 background of a movie (summary_image_generation_module.m:62-96,
    :130, detrend mode 'pixelwise').

    movie: single [height*width, frames] (or [height, width, frames]), column-major.
    Returns (summary [height, width], bg_spatial [height, width], bg_temporal [1, frames]),
    float64, the summary normalised by its max.
    """
    lib = load_library()
    m = _col(movie, np.float32)
    if m.ndim == 3:
        if m.shape[:2] != (height, width):
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"movie {m.shape} is not [{height}, {width}, T]")
        m = m.reshape(height * width, m.shape[2], order="F")
    if m.ndim != 2 or m.shape[0] != height * width:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"movie {m.shape} is not [{height}*{width}, T]")
    T = m.shape[1]
    img = np.empty((height, width), np.float64, order="F")
    bs = np.empty((height, width), np.float64, order="F")
    bt = np.empty((1, T), np.float64, order="F")
    _check(lib.mlf_summary_image(_fptr(m), height, width, T, int(window_1), int(window_2),
                                 int(poly_index), int(bg_iter), _fptr(img, C.c_double),
                                 _fptr(bs, C.c_double), _fptr(bt, C.c_double), int(device)))
    return img, bs, bt


def summary_image_global(movie, height: int, width: int, bg_iter: int = 5, device: int = 0):
    """Summary image and rank-1 background for detrend mode 'global'
    (summary_image_generation_module.m:43-60, :130); arguments and returns as
    summary_image_pixelwise."""
    lib = load_library()
    m = _col(movie, np.float32)
    if m.ndim == 3:
        if m.shape[:2] != (height, width):
            raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"movie {m.shape} is not [{height}, {width}, T]")
        m = m.reshape(height * width, m.shape[2], order="F")
    if m.ndim != 2 or m.shape[0] != height * width:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, f"movie {m.shape} is not [{height}*{width}, T]")
    T = m.shape[1]
    img = np.empty((height, width), np.float64, order="F")
    bs = np.empty((height, width), np.float64, order="F")
    bt = np.empty((1, T), np.float64, order="F")
    _check(lib.mlf_summary_image_global(_fptr(m), height, width, T, int(bg_iter), _fptr(img, C.c_double),
                                        _fptr(bs, C.c_double), _fptr(bt, C.c_double), int(device)))
    return img, bs, bt


def summary_image_generation_module(sensor_movie, patch_info, SI, device: int = 0):
    """Drop-in for summary_image_generation_module.m:1-130 (detrend modes
    'pixelwise', the default, and 'global'): returns (summary_image,
    bg_spatial_init, bg_temporal_init).  The images and plots the module
    saves are left out (in 'global' mode the smoothed baseline only feeds one)."""
    mode = _si_get(SI, "detrend_mode", "pixelwise")
    if mode not in ("pixelwise", "global"):
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, 'Invalid value for param "detrend_mode"')
    size = _si_get(patch_info, "size")
    h, w = int(size[0]), int(size[1])
    if mode == "global":
        return summary_image_global(sensor_movie, h, w, _si_get(SI, "bg_iter", 5), device)
    return summary_image_pixelwise(sensor_movie, h, w, _si_get(SI, "pixelwise_window_1", 100),
                                   _si_get(SI, "pixelwise_window_2", 20),
                                   _si_get(SI, "pixelwise_poly_index", 3), _si_get(SI, "bg_iter", 5),
                                   device)


def mex_lfm_convolution_vec(bitmap, kernels, device: int = 0):
    """Drop-in for the MEX gateway (mex:276-421), host arrays in and out.

    bitmap: single [W, H] or [W, H, n]; kernels: single [kw, kw, 15, 15].
    Invalid input raises :class:`InvalidInput` whose ``identifier`` is the
    reference's error id.
    """
    lib = load_library()
    b = np.asarray(bitmap)
    k = np.asarray(kernels)
    if b.ndim not in (2, 3) or k.ndim != 4:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, MEX_ERROR_MSG)
    if b.dtype != np.float32 or k.dtype != np.float32:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, MEX_ERROR_MSG)
    if k.shape[0] != k.shape[1] or k.shape[2] != 15 or k.shape[3] != 15:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, MEX_ERROR_MSG)
    b = np.asfortranarray(b)
    k = np.asfortranarray(k)
    W, Hh = b.shape[0], b.shape[1]
    n = b.shape[2] if b.ndim == 3 else 1
    out = np.empty(b.shape, np.float32, order="F")
    _check(lib.mlf_lfm_convolution_vec(_fptr(b), W, Hh, n, _fptr(k), k.shape[0], 15, _fptr(out),
                                       int(device)))
    return out


def lfm_convolution_segments(slices, planes, seg_ptr, kernels, device: int = 0):
    """All MEX projections of the seed-library builder in one call
    (``gen_seed_library_local_with_ring_bg_same_size.m:263-318``).

    slices: single [W, H, nslices]; planes: 0-based plane per slice;
    seg_ptr: [nseg+1] slice ranges per segment; kernels: single
    [kw, kw, 15, 15, Z].  Returns [W, H, nseg]: per segment, the sum over its
    slices of ``mex_lfm_convolution_vec(slice, kernels[..., plane])``, added
    in ascending plane order (the reference's loop order)."""
    lib = load_library()
    s = _col(slices, np.float32)
    if s.ndim == 2:
        s = s[:, :, None]
    K = _col(kernels, np.float32)
    if K.ndim == 4:
        K = K[..., None]
    pl = np.ascontiguousarray(planes, dtype=np.int32)
    sp = np.ascontiguousarray(seg_ptr, dtype=np.int64)
    if s.ndim != 3 or K.ndim != 5 or K.shape[0] != K.shape[1] or pl.size != s.shape[2] or sp.ndim != 1:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "bad argument shapes")
    nseg = sp.size - 1
    out = np.empty((s.shape[0], s.shape[1], max(nseg, 0)), np.float32, order="F")
    _check(lib.mlf_lfm_convolution_segments(_fptr(s), s.shape[0], s.shape[1], s.shape[2],
                                            _fptr(pl, C.c_int32), _fptr(sp, C.c_int64), nseg,
                                            _fptr(K), K.shape[0], K.shape[2], K.shape[4],
                                            _fptr(out), int(device)))
    return out


def lfm_convolution_segments_device(ctx: "Context", d_slices: int, W: int, H: int, nslices: int, planes,
                                    seg_ptr, d_kernels: int, kw: int, Z: int, d_out: int):
    """``lfm_convolution_segments`` on device buffers given as raw pointers
    (ints): slices [W, H, nslices], kernels [kw, kw, 15, 15, Z] and the output
    [W, H, nseg], all single; planes and seg_ptr are host sequences."""
    pl = np.ascontiguousarray(planes, dtype=np.int32)
    sp = np.ascontiguousarray(seg_ptr, dtype=np.int64)
    if pl.size != int(nslices) or sp.ndim != 1:
        raise InvalidInput(MLF_ERR_INVALID_ARGUMENT, "bad argument shapes")
    _check(ctx._lib.mlf_lfm_convolution_segments_device(
        ctx.handle, C.c_void_p(d_slices), int(W), int(H), int(nslices), _fptr(pl, C.c_int32),
        _fptr(sp, C.c_int64), sp.size - 1, C.c_void_p(d_kernels), int(kw), 15, int(Z),
        C.c_void_p(d_out)))
