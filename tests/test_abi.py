"""C-ABI library checks that need no GPU: load, exported symbols, argument
validation paths that return before touching a device."""
import os
import re

import numpy as np
import pytest

import mesolf_amd
from mesolf_amd import lib as L


def _header_functions(root):
    txt = open(os.path.join(root, "include", "mesolf.h")).read()
    return sorted(set(re.findall(r"^\s*(?:mlf_status|const char\*|int32_t)\s+(mlf_\w+)\(", txt, re.M)))


def test_library_is_built(root_dir):
    assert os.path.exists(os.path.join(root_dir, "mesolf_amd", "libmesolf.so")), \
        "build with python -m mesolf_amd.build"


def test_library_exports_every_header_symbol(root_dir):
    lib = L.load_library()
    declared = _header_functions(root_dir)
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes binding covers exactly the header
    assert declared == L.exported_symbols()


def test_abi_version_and_status_strings():
    lib = L.load_library()
    assert lib.mlf_abi_version() == 2
    assert lib.mlf_status_string(1) == b"invalid argument"


def test_mex_rejects_bad_input_with_reference_error_id():
    b = np.zeros((20, 20), np.float32)
    with pytest.raises(L.InvalidInput) as ei:
        mesolf_amd.mex_lfm_convolution_vec(b, np.zeros((5, 5, 15), np.float32))
    assert ei.value.identifier == "parallel:gpu:mexGPUExample:InvalidInput"
    with pytest.raises(L.InvalidInput):   # class must be single (mex:353)
        mesolf_amd.mex_lfm_convolution_vec(b.astype(np.float64), np.zeros((5, 5, 15, 15), np.float32))
    with pytest.raises(L.InvalidInput):   # 1-D bitmap (mex:313)
        mesolf_amd.mex_lfm_convolution_vec(np.zeros(20, np.float32), np.zeros((5, 5, 15, 15), np.float32))


def test_c_abi_validates_before_device():
    lib = L.load_library()
    b = np.zeros((8, 8), np.float32)
    k = np.zeros((3, 3, 15, 15), np.float32)
    out = np.zeros_like(b)
    # nphase != 15 -> invalid argument, independent of any device
    st = lib.mlf_lfm_convolution_vec(L._fptr(b), 8, 8, 1, L._fptr(k), 3, 14, L._fptr(out), 0)
    assert st == L.MLF_ERR_INVALID_ARGUMENT
    assert b"15" in lib.mlf_last_error()
    # gpuFlag = false -> unsupported (reference CPU branch cannot run)
    X = np.zeros((15, 15, 1))
    H = np.zeros((15, 15, 15, 15, 1), np.float32)
    st = lib.mlf_recon_phase_space(L._fptr(b), 15, 15, L._fptr(H), L._fptr(H), 15, 15, 1, 1, 0, 0, 0,
                                   L._fptr(X, L.C.c_double), None)
    assert st == L.MLF_ERR_UNSUPPORTED
    assert b"Xguessv" in lib.mlf_last_error()


def test_segment_projection_validates_before_device():
    s = np.zeros((8, 8, 3), np.float32)
    K = np.zeros((3, 3, 15, 15, 2), np.float32)
    with pytest.raises(mesolf_amd.InvalidInput, match="seg_ptr"):
        mesolf_amd.lfm_convolution_segments(s, [0, 1, 1], [0, 2, 1, 3], K)   # decreasing
    with pytest.raises(mesolf_amd.InvalidInput, match="seg_ptr"):
        mesolf_amd.lfm_convolution_segments(s, [0, 1, 1], [0, 2], K)         # does not end at 3
    with pytest.raises(mesolf_amd.InvalidInput, match="plane"):
        mesolf_amd.lfm_convolution_segments(s, [0, 2, 1], [0, 3], K)         # plane 2 of Z = 2


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU every compute entry point fails loudly."""
    try:
        ctx = mesolf_amd.Context(0)
    except L.MesolfError as e:
        assert e.status == L.MLF_ERR_DEVICE
    else:
        ctx.close()
        pytest.skip("a GPU is present")
    b = np.zeros((20, 20), np.float32)
    with pytest.raises(L.MesolfError):
        mesolf_amd.mex_lfm_convolution_vec(b, np.zeros((5, 5, 15, 15), np.float32))



@pytest.mark.parametrize("gh,gw,n", [(1000, 1000, [3, 3]), (100, 100, 2), (700, 451, [2, 3]),
                                     (1515, 1515, [4, 5]), (60, 2000, [1, 7]), (333, 77, 1)])
def test_carpet_tiling_matches_oracle(gh, gw, n):
    # host-only entry point: runs without a GPU
    from oracle import mesolf_oracle as O
    got = mesolf_amd.carpet_patch_info(gh, gw, n)
    want = O.carpet_patches(gh, gw, n)
    assert [tuple(map(tuple, g["location"])) for g in got] == want
    for g in got:
        assert g["size"][0] % 15 == 0 and g["size"][1] % 15 == 0 and min(g["size"]) >= 50



@pytest.mark.parametrize("H,W,xc,yc,M,dx", [(20, 20, 8.0, 8.0, 15, 15.0), (40, 33, 3.5, 10.25, 15, 7.5),
                                             (1515, 1400, 700.3, 651.9, 15, 14.87), (61, 47, 10.3, 7.6, 15, 14.2)])
def test_rectify_size_matches_oracle(H, W, xc, yc, M, dx):
    # host-only entry point: runs without a GPU
    from oracle import mesolf_oracle as O
    assert mesolf_amd.rectify_size(H, W, xc, yc, M, dx) == (len(O.rectify_axis(yc, H, M, dx)),
                                                            len(O.rectify_axis(xc, W, M, dx)))
