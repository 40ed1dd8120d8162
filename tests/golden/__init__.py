"""Fixtures decoded from the reference's own test data (see make_*.py)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_vols_slice() -> np.ndarray:
    """The input of src/cuda/mex_lfm_convolution_vec_test.m (vols_slice.mat):
    single [376, 376, 77], column-major, rebuilt from the sparse fixture."""
    d = np.load(os.path.join(HERE, "vols_slice_sparse.npz"))
    shape = tuple(int(x) for x in d["shape"])
    flat = np.zeros(int(np.prod(shape)), np.float32)
    flat[d["index"].astype(np.int64)] = d["value"]
    return flat.reshape(shape, order="F")
