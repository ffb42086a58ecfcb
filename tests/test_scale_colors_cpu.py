"""scale_colors' multiplier logic (core/libraw_wrapper.cpp:612-862) through the
C ABI (host code, no device): one known-answer case per branch, with the
values worked out by hand from the reference's statements."""
from __future__ import annotations

import numpy as np
import pytest

from libraw_enhanced_amd.accelerator import (LibRawInvalidArgument, scale_colors_multipliers, wb_meta, wb_needs_auto)

RGGB = 0x94949494  # colours 0 1 / 1 2: the second green folds to 1
RGBG = 0xB4B4B4B4  # colours 0 1 / 3 2: the second green keeps index 3


def _meta(**kw):
    kw.setdefault("filters", RGGB)
    kw.setdefault("maximum", 16383)
    return wb_meta(**kw)


def _run(m, auto=None):
    try:
        return 0, scale_colors_multipliers(m, auto)
    except LibRawInvalidArgument:
        return -1, None


def f32(x):
    return np.float32(x)


def test_user_multipliers():
    m = _meta(user_mul=(2.0, 1.0, 1.5, 1.0), black=512)
    assert not wb_needs_auto(m)
    rc, s = _run(m)
    assert rc == 0 and list(m.pre_mul) == [2.0, 1.0, 1.5, 1.0]
    assert np.array_equal(s, np.float32([2.0, 1.0, 1.5, 1.0]))  # pre_mul / min
    assert m.maximum == 16383 - 512


def test_camera_wb_from_white_samples():
    # 16 samples of each of R, G, B, G2 in the 8x8 block; cblack 100 everywhere
    cols = np.array([[0, 1] * 4, [3, 2] * 4] * 4)
    white = np.choose(cols, [1100, 1000, 1300, 1000])
    m = _meta(use_camera_wb=1, filters=RGBG, cam_mul=(2.2, 1.0, 1.4, 1.0), white=white,
              cblack=(100, 100, 100, 100))
    rc, s = _run(m)
    want_pre = [f32(16) / f32(16 * 1000), f32(16) / f32(16 * 900), f32(16) / f32(16 * 1200), f32(16) / f32(16 * 900)]
    assert rc == 0 and np.array_equal(np.float32(m.pre_mul[:]), np.float32(want_pre))
    dmin = min(float(v) for v in want_pre)
    assert np.array_equal(s, np.float32([float(v) / dmin for v in want_pre]))
    # a three-colour pattern never sees colour 3, so its sum stays 0 and cam_mul is taken
    m3 = _meta(use_camera_wb=1, cam_mul=(2.2, 1.0, 1.4, 1.0), white=white, cblack=(100, 100, 100, 100))
    _run(m3)
    assert np.array_equal(np.float32(m3.pre_mul[:]), np.float32([2.2, 1.0, 1.4, 1.0]))
    # samples at or below the black level add nothing (the difference wraps to a negative int)
    low = np.where(cols == 2, 50, white)
    mb = _meta(use_camera_wb=1, filters=RGBG, cam_mul=(2.2, 1.0, 1.4, 1.0), white=low,
               cblack=(100, 100, 100, 100))
    _run(mb)
    assert np.array_equal(np.float32(mb.pre_mul[:]), np.float32([2.2, 1.0, 1.4, 1.0]))


def test_camera_wb_without_samples_takes_cam_mul():
    m = _meta(use_camera_wb=1, cam_mul=(2.2, 1.0, 1.4, 1.0))  # white all 0: every val <= 0
    rc, s = _run(m)
    assert rc == 0 and np.array_equal(np.float32(m.pre_mul[:]), np.float32([2.2, 1.0, 1.4, 1.0]))
    assert np.array_equal(s, np.float32([2.2, 1.0, 1.4, 1.0]))


def test_bad_camera_wb_warns():
    m = _meta(use_camera_wb=1, cam_mul=(2.2, 1.0, 0.0, 1.0))
    rc, _ = _run(m)
    assert rc == 0 and m.warnings & 0x4 and list(m.pre_mul) == [1.0, 1.0, 1.0, 1.0]


@pytest.mark.parametrize("cam0,fallback,auto", [(-1.0, 0, True), (0.0, 0, True), (0.0, 1, False)])
def test_automatic_path(cam0, fallback, auto):
    m = _meta(use_camera_wb=1, camerawb_daylight_fallback=fallback, cam_mul=(cam0, 1.0, 1.0, 1.0))
    assert wb_needs_auto(m) == auto
    if auto:
        assert _run(_meta(use_camera_wb=1, camerawb_daylight_fallback=fallback,
                          cam_mul=(cam0, 1.0, 1.0, 1.0)))[0] != 0  # the sums are required
        rc, s = _run(m, (1.8, 1.0, 1.3, 1.0))
        assert rc == 0 and np.array_equal(s, np.float32([1.8, 1.0, 1.3, 1.0]))  # not divided by the minimum


def test_nikon_sraw_daylight_and_normalisation():
    m = _meta(as_shot_wb_applied=1, cam_mul=(2.0, 1.0, 1.5, 1.0), pre_mul=(4.0, 0.0, 3.0, 0.0))
    rc, s = _run(m)
    # pre_mul /= cam_mul for R, G, B, then G = 0 -> 1 and G2 = 0 -> G (three colours)
    assert rc == 0 and np.array_equal(np.float32(m.pre_mul[:]), np.float32([2.0, 1.0, 2.0, 1.0]))
    m4 = _meta(pre_mul=(2.0, 0.5, 1.0, 0.0), colors=4)
    _run(m4)
    assert m4.pre_mul[3] == 1.0


def test_cblack_pattern_folded():
    m = _meta(cblack=(1, 2, 3, 4, 2, 2, 10, 20, 30, 40))
    rc, _ = _run(m)
    # 0 1 / 1 2: (0,0) R += 10, (0,1) and (1,0) G += 20 + 30, (1,1) B += 40
    assert rc == 0 and list(m.cblack[:6]) == [11, 52, 43, 4, 0, 0]
    m4 = _meta(filters=RGBG, cblack=(1, 2, 3, 4, 2, 2, 10, 20, 30, 40))
    _run(m4)
    assert list(m4.cblack[:6]) == [11, 22, 43, 34, 0, 0]  # 0 1 / 3 2: G2 keeps its own entry
    m1 = _meta(cblack=(1, 2, 3, 4, 1, 1, 7, 0, 0, 0))  # a 1x1 pattern adds its value once per site
    _run(m1)
    assert list(m1.cblack[:6]) == [8, 16, 10, 4, 0, 0]  # both greens of 0 1 / 1 2 land on G
    m3 = _meta(cblack=(1, 2, 3, 4, 3, 3, 7, 0, 0, 0))  # 3x3 patterns are left alone
    _run(m3)
    assert list(m3.cblack[:6]) == [1, 2, 3, 4, 3, 3]


def test_black_above_maximum_wraps_and_keeps_the_multipliers():
    m = _meta(user_mul=(2.0, 1.0, 1.5, 1.0), black=20000)
    rc, s = _run(m)
    assert rc == 0 and m.maximum == (16383 - 20000) % 2 ** 32
    assert np.array_equal(s, np.float32([2.0, 1.0, 1.5, 1.0]))
