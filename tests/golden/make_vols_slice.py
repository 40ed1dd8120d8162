"""Decode the reference's MEX test input into a sparse fixture.

``src/cuda/vols_slice.mat`` (read by ``src/cuda/mex_lfm_convolution_vec_test.m:61-72``)
is a MATLAB 5.0 MAT-file holding one gpuArray variable ``vols_slice``
(single, 376 x 376 x 77).  A saved gpuArray is an MCOS object: the MAT-file's
``__function_workspace__`` element is a nested MAT stream whose object
property cell holds the array.  scipy's MAT-5 reader parses both levels (no
code in the file is executed).  The array is mostly zeros, so the fixture
stores it as (linear column-major index, value) pairs.

Run from the repository root:  python tests/golden/make_vols_slice.py
Writes tests/golden/vols_slice_sparse.npz.
"""
import io
import os

import numpy as np
import scipy.io as sio
from scipy.io.matlab._mio5 import MatFile5Reader

SRC = "/root/reference/src/cuda/vols_slice.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vols_slice_sparse.npz")


def _find_single_3d(x):
    if isinstance(x, np.ndarray) and x.dtype == np.float32 and x.ndim == 3:
        return x
    if isinstance(x, np.ndarray) and x.dtype == object:
        for y in x.ravel():
            r = _find_single_3d(y)
            if r is not None:
                return r
    if isinstance(x, np.ndarray) and x.dtype.names:
        for f in x.dtype.names:
            for y in np.atleast_1d(x[f]).ravel():
                r = _find_single_3d(y)
                if r is not None:
                    return r
    return None


def load_vols_slice(path=SRC):
    top = sio.loadmat(path)
    assert top["__header__"].startswith(b"MATLAB 5.0 MAT-file")
    obj = top["None"][0]
    assert obj["s0"] == b"vols_slice" and obj["s2"] == b"gpuArray"
    ws = top["__function_workspace__"].tobytes()
    # the workspace is a MAT stream without the 116-byte text header
    hdr = b"MATLAB 5.0 MAT-file".ljust(116, b" ") + b"\x00" * 8 + ws[:4]
    rd = MatFile5Reader(io.BytesIO(hdr + ws[8:]), squeeze_me=False, struct_as_record=True)
    rd.initialize_read()
    rd.mat_stream.seek(128)
    h, _ = rd.read_var_header()
    a = _find_single_3d(rd.read_var_array(h, process=True))
    assert a is not None and a.shape == (376, 376, 77), None if a is None else a.shape
    return np.asfortranarray(a)


def main():
    a = load_vols_slice()
    flat = a.ravel(order="F")
    idx = np.flatnonzero(flat).astype(np.uint32)
    val = flat[idx].astype(np.float32)
    np.savez_compressed(OUT, shape=np.array(a.shape, np.int64), index=idx, value=val)
    print(f"{OUT}: shape {a.shape}, nnz {idx.size}, max {val.max():.7f}, "
          f"sum {val.astype(np.float64).sum():.3f}, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
