"""determine_patch_info_aug (host only): the library against the oracle, the
hand-worked layouts, and the reference's own consistency rule -- a patch's
raw cut, rectified about its microlens centre, has exactly the patch's size
(image_load_module_aug.m:106-108)."""
import numpy as np
import pytest

import mesolf_amd
from oracle import mesolf_oracle as O


def _si(x, y, dx, crop, **kw):
    d = {"frames_x_offset": x, "frames_y_offset": y, "frames_dx": dx, "Nnum": 15,
         "frames_crop_border_microlenses": crop}
    d.update(kw)
    return d


def test_hand_worked_carpet_and_roi():
    got = mesolf_amd.determine_patch_info_aug(_si(100.0, 90.0, 15.0, [1, 1, 1, 1], patch_mode="carpet",
                                                  n_patches=[2, 1]), (200, 230))
    assert [p["raw_cut"] for p in got] == [[[17, 102], [12, 217]], [[92, 177], [12, 217]]]
    assert [(p["raw_cut_xCenter"], p["raw_cut_yCenter"]) for p in got] == [(103.0, 43.0), (103.0, 43.0)]
    roi = mesolf_amd.determine_patch_info_aug(_si(100.0, 90.0, 15.0, [1, 1, 1, 1], patch_mode="roi",
                                                  ROI_list=[[16, 31, 120, 170]], ROI_patch_size=64), (200, 230))
    assert [p["location"] for p in roi] == [[[16, 31], [75, 90]], [[16, 91], [75, 150]]]


def test_library_equals_oracle_on_random_layouts():
    rng = np.random.default_rng(0)
    checked = 0
    for trial in range(300):
        H, W = int(rng.integers(300, 2500)), int(rng.integers(300, 2500))
        dx = float(rng.uniform(13.5, 16.5))
        xc, yc = float(rng.uniform(0.3 * W, 0.7 * W)), float(rng.uniform(0.3 * H, 0.7 * H))
        crop = [int(v) for v in rng.integers(0, 3, 4)]
        if trial % 2:
            kw = {"patch_mode": "roi", "ROI_list": [[31, 46, 400, 500], [100, 20, 300, 220]],
                  "ROI_patch_size": int(rng.integers(100, 300))}
            okw = {"roi_list": kw["ROI_list"], "roi_patch_size": kw["ROI_patch_size"]}
        else:
            kw = {"patch_mode": "carpet", "n_patches": [int(rng.integers(1, 5)), int(rng.integers(1, 5))]}
            okw = {"n_patches": kw["n_patches"]}
        try:
            ref = O.determine_patch_info(H, W, xc, yc, dx, 15, crop, kw["patch_mode"], **okw)
        except (ValueError, IndexError):
            with pytest.raises(mesolf_amd.InvalidInput):
                mesolf_amd.determine_patch_info_aug(_si(xc, yc, dx, crop, **kw), (H, W))
            continue
        got = mesolf_amd.determine_patch_info_aug(_si(xc, yc, dx, crop, **kw), (H, W))
        assert len(got) == len(ref)
        for a, b in zip(got, ref):
            assert a["location"] == [[int(v) for v in r] for r in b["location"]]
            assert a["raw_cut"] == b["raw_cut"]
            assert a["raw_cut_xCenter"] == b["raw_cut_xCenter"] and a["raw_cut_yCenter"] == b["raw_cut_yCenter"]
            checked += 1
    assert checked > 500


def test_raw_cut_rectifies_to_the_patch_size():
    """For carpet layouts of typical frames every patch passes the reference's
    own check: the rectified raw cut has the patch's size."""
    rng = np.random.default_rng(1)
    n = 0
    for _ in range(100):
        H, W = int(rng.integers(600, 2500)), int(rng.integers(600, 2500))
        dx = float(rng.uniform(13.5, 16.5))
        si = _si(float(rng.uniform(0.3 * W, 0.7 * W)), float(rng.uniform(0.3 * H, 0.7 * H)), dx,
                 [int(v) for v in rng.integers(1, 3, 4)], patch_mode="carpet",
                 n_patches=[int(rng.integers(1, 5)), int(rng.integers(1, 5))])
        for p in mesolf_amd.determine_patch_info_aug(si, (H, W)):
            (y0, y1), (x0, x1) = p["raw_cut"]
            assert mesolf_amd.rectify_size(y1 - y0 + 1, x1 - x0 + 1, p["raw_cut_xCenter"], p["raw_cut_yCenter"],
                                           15, dx) == tuple(p["size"])
            n += 1
    assert n > 300


def test_argument_errors():
    si = _si(100.0, 90.0, 15.0, [8, 8, 1, 1], patch_mode="carpet", n_patches=[1, 1])
    with pytest.raises(mesolf_amd.InvalidInput, match="X-cut"):
        mesolf_amd.determine_patch_info_aug(si, (200, 230))
    with pytest.raises(mesolf_amd.InvalidInput, match="carpet"):
        mesolf_amd.determine_patch_info_aug(dict(si, patch_mode="grid"), (200, 230))
    with pytest.raises(mesolf_amd.InvalidInput, match="ROI_patch_size"):
        mesolf_amd.determine_patch_info_aug(dict(si, patch_mode="roi", ROI_list=[[1, 1, 60, 60]]), (200, 230))
