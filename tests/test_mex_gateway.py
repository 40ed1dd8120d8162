"""The MATLAB MEX gateways, compiled against a mock MEX API (no MATLAB here).

CPU tests exercise argument validation and the error identifiers callers
match on; GPU tests run the gateways end to end and compare with the oracle.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from mesolf_amd import synth
from oracle import mesolf_oracle as O
from tests.helpers import rel_l2

REF_ID = "parallel:gpu:mexGPUExample:InvalidInput"


@pytest.fixture(scope="session")
def drivers(root_dir, tmp_path_factory):
    out = tmp_path_factory.mktemp("mockmx")
    subprocess.run(["sh", os.path.join(root_dir, "tests", "mockmx", "build.sh"), str(out)], check=True)
    return {"conv": str(out / "mex_lfm_convolution_vec"), "recon": str(out / "recon_phase_space_mex")}


def _arg(tmp, name, a):
    a = np.asfortranarray(a)
    path = str(tmp / f"{name}.bin")
    a.ravel(order="F").tofile(path)
    code = "s" if a.dtype == np.float32 else "d"
    return f"{code}:{'x'.join(str(d) for d in a.shape)}:{path}"


def _call(driver, out, *args):
    r = subprocess.run([driver, out, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("case", ["nrhs", "double_bitmap", "kernels_3d", "kernels_14_phases",
                                  "kernels_not_square", "bitmap_4d"])
def test_conv_gateway_rejects_bad_input_with_reference_id(drivers, tmp_path, case):
    b = np.zeros((20, 18), np.float32)
    k = np.zeros((5, 5, 15, 15), np.float32)
    if case == "nrhs":
        args = [_arg(tmp_path, "b", b)]
    elif case == "double_bitmap":
        args = [_arg(tmp_path, "b", b.astype(np.float64)), _arg(tmp_path, "k", k)]
    elif case == "kernels_3d":
        args = [_arg(tmp_path, "b", b), _arg(tmp_path, "k", np.zeros((5, 5, 15), np.float32))]
    elif case == "kernels_14_phases":
        args = [_arg(tmp_path, "b", b), _arg(tmp_path, "k", np.zeros((5, 5, 14, 15), np.float32))]
    elif case == "kernels_not_square":
        args = [_arg(tmp_path, "b", b), _arg(tmp_path, "k", np.zeros((5, 7, 15, 15), np.float32))]
    else:
        args = [_arg(tmp_path, "b", np.zeros((4, 4, 2, 2), np.float32)), _arg(tmp_path, "k", k)]
    res = _call(drivers["conv"], "-", *args)
    assert res["status"] == "error"
    assert res["id"] == REF_ID
    assert res["msg"] == "Invalid input to mex_lfm_convolution_vec()"


def test_recon_gateway_validates_and_rejects_cpu_flag(drivers, tmp_path):
    res = _call(drivers["recon"], "-", "S:1", "S:2")
    assert res["status"] == "error" and res["id"] == "mesolf:recon:InvalidInput"
    lf = np.zeros((30, 30), np.float32)
    H = np.zeros((15, 15, 15, 15, 2), np.float32)
    args = [_arg(tmp_path, "lf", lf), _arg(tmp_path, "H", H), _arg(tmp_path, "Ht", H), "S:1", "S:0", "S:1"]
    res = _call(drivers["recon"], "-", *args)
    assert res["status"] == "error" and "Xguessv" in res["msg"]


@pytest.mark.gpu
def test_conv_gateway_end_to_end(drivers, tmp_path):
    rng = np.random.default_rng(21)
    b = rng.random((40, 36, 2)).astype(np.float32)
    k = rng.random((9, 9, 15, 15)).astype(np.float32)
    out = str(tmp_path / "out.bin")
    res = _call(drivers["conv"], out, _arg(tmp_path, "b", b), _arg(tmp_path, "k", k))
    assert res["status"] == "ok" and res["dims"] == [40, 36, 2] and res["class"] == 7
    got = np.fromfile(out, np.float32).reshape((40, 36, 2), order="F")
    ref = O.mex_op(b.astype(np.float64), k.astype(np.float64))
    assert rel_l2(got, ref) <= 1e-5


@pytest.mark.gpu
def test_recon_gateway_end_to_end_double_inputs(drivers, tmp_path):
    wl = synth.Workload("gw", 3, 4, (-4, 0, 4), 1, 1)
    Hps = synth.dense_phase_space(wl)
    H = O.inverse_phasespace(Hps)
    Ht = synth.adjoint_ht(Hps)
    lf = np.random.default_rng(3).poisson(60.0, (wl.I1, wl.I2)).astype(np.float64)
    out = str(tmp_path / "x.bin")
    res = _call(drivers["recon"], out, _arg(tmp_path, "lf", lf), _arg(tmp_path, "H", H),
                _arg(tmp_path, "Ht", Ht), "S:2", "S:1", "S:1")
    assert res["status"] == "ok" and res["class"] == 6 and res["dims"] == [wl.I1, wl.I2, wl.Z]
    got = np.fromfile(out, np.float64).reshape((wl.I1, wl.I2, wl.Z), order="F")
    # the gateway converts double inputs to single, as the library takes single
    cp = O.compact_from_dense(H.astype(np.float32).astype(np.float64),
                              Ht.astype(np.float32).astype(np.float64))
    ref = O.rl_compact(lf.astype(np.float32).astype(np.float64), cp, 2)
    assert rel_l2(got, ref) <= 1e-4
