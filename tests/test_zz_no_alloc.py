"""The no-allocation MEX entry point against the allocating one.

mlf_lfm_convolution_vec_noalloc keeps the MFMA form's GEMM operand in the
library's file-scope device array instead of a buffer it allocates; its
results must equal mlf_lfm_convolution_vec_device's bit for bit.  Both are
called through the C ABI on device buffers (torch tensors serve only as
device memory), twice in a row, the second call smaller than the first, so
the array's leftover contents from the first call must not leak into the
second.
"""
import ctypes as C

import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import lib as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _to_dev(a):
    return torch.from_numpy(np.asfortranarray(a, dtype=np.float32).ravel(order="F").copy()).to("cuda:0")


def _run(lib, ctx, fn, d_in, shape, d_k, kw):
    W, H, n = shape
    out = torch.full((W * H * n,), float("nan"), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    st = getattr(lib, fn)(ctx.handle, C.c_void_p(d_in.data_ptr()), W, H, n, C.c_void_p(d_k.data_ptr()),
                          kw, 15, C.c_void_p(out.data_ptr()))
    L._check(st)
    ctx.synchronize()
    return out.cpu().numpy().reshape((W, H, n), order="F")


@pytest.fixture(scope="module")
def ctx():
    c = mesolf_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("calls", [[((160, 131, 3), 241), ((70, 64, 2), 31)],
                                   [((90, 75, 2), 8), ((40, 33, 1), 255)]],
                         ids=["odd-then-smaller", "even-then-odd"])
def test_noalloc_mex_equals_allocating_entry(ctx, calls):
    lib = L.load_library()
    rng = np.random.default_rng(31)
    for shape, kw in calls:
        b = rng.random(shape).astype(np.float32)
        K = rng.random((kw, kw, 15, 15)).astype(np.float32)
        d_b, d_k = _to_dev(b), _to_dev(K)
        print(f"no-alloc MEX: bitmap {shape}, kw {kw}", flush=True)
        got = _run(lib, ctx, "mlf_lfm_convolution_vec_noalloc", d_b, shape, d_k, kw)
        ref = _run(lib, ctx, "mlf_lfm_convolution_vec_device", d_b, shape, d_k, kw)
        assert np.isfinite(got).all()
        assert np.array_equal(got, ref), (shape, kw, float(np.abs(got - ref).max()))


def test_noalloc_mex_rejects_kernels_beyond_its_scratch(ctx):
    lib = L.load_library()
    kw = 259
    d_b = _to_dev(np.ones((20, 20, 1), np.float32))
    d_k = _to_dev(np.ones((kw, kw, 15, 15), np.float32))
    out = torch.empty(400, dtype=torch.float32, device="cuda:0")
    st = lib.mlf_lfm_convolution_vec_noalloc(ctx.handle, C.c_void_p(d_b.data_ptr()), 20, 20, 1,
                                             C.c_void_p(d_k.data_ptr()), kw, 15, C.c_void_p(out.data_ptr()))
    assert st == L.MLF_ERR_INVALID_ARGUMENT
    assert b"257" in lib.mlf_last_error()
