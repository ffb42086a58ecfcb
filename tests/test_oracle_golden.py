"""The CPU restatement (oracle) against the golden fixtures written from the
reference's own CPU code (tests/golden/make_golden.py), plus invariants that the
mathematics fixes.  Runs without a GPU and without the reference tree."""
import os

import numpy as np
import pytest

from _common import (TILES_CLEAN, TILES_SEQUENTIAL, oracle_amaze, oracle_detail_tonemap,
                     oracle_tone_mapping, oracle_white_balance)
from libraw_enhanced_amd import scene

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "develop_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", ["amaze_rggb_64x96", "amaze_bggr_gain_80x112", "amaze_grbg_144x176"])
def test_amaze_golden(gold, name):
    cfa = gold[name + "/cfa"]
    filters, r, g, b, g2, mx = gold[name + "/params"]
    raw3 = scene.cfa_to_rgb3(cfa, int(filters))
    got = oracle_amaze(raw3, int(filters), (r, g, b, g2), mx, TILES_SEQUENTIAL)
    assert np.array_equal(got, gold[name + "/out"])
    # single-tile images: the clean-tile contract gives the same result
    if cfa.shape[0] <= 128 and cfa.shape[1] <= 128:
        assert np.array_equal(oracle_amaze(raw3, int(filters), (r, g, b, g2), mx, TILES_CLEAN),
                              gold[name + "/out"])


def test_detail_tm_golden(gold):
    got = oracle_detail_tonemap(gold["detail_tm_40x56/in"])
    assert np.array_equal(got, gold["detail_tm_40x56/out"])


def test_tone_mapping_golden(gold):
    got = oracle_tone_mapping(gold["tone_16x24/in"], 1.3)
    assert np.array_equal(got, gold["tone_16x24/out"])


@pytest.mark.parametrize("name,filters", [("wb_bayer_24x36", 0x94949494), ("wb_xtrans_24x36", 9)])
def test_white_balance_golden(gold, name, filters):
    got = oracle_white_balance(gold[name + "/img4"], (2.1, 1.0, 1.5, 1.0), filters,
                               scene.XTRANS if filters == 9 else None)
    assert np.array_equal(got, gold[name + "/out"])


@pytest.mark.parametrize("name", ["defringe_oop_64x96", "defringe_inplace_green_64x96"])
def test_defringe_golden(gold, name):
    import ctypes
    from _common import oracle, p
    r_, e_, c_, s_, g_, gs_, inplace = gold[name + "/params"]
    a = gold[name + "/in"].copy()
    res = a if inplace else np.zeros_like(a)
    assert oracle().oracle_defringe(p(a), p(res), 96, 64, ctypes.c_float(r_), ctypes.c_float(e_),
                                    ctypes.c_float(c_), ctypes.c_float(s_), int(g_), ctypes.c_float(gs_)) == 0
    assert np.array_equal(res, gold[name + "/out"])


def test_ca_golden(gold):
    import ctypes
    from _common import oracle, p
    src = gold["axial_96x128/in"]
    res = np.zeros_like(src)
    assert oracle().oracle_axial_ca(p(src), p(res), 128, 96, 6, ctypes.c_float(1e-4), ctypes.c_float(0.3)) == 0
    assert np.array_equal(res, gold["axial_96x128/out"])
    src = gold["lateral_96x128/in"]
    res = np.zeros_like(src)
    assert oracle().oracle_lateral_ca(p(src), p(res), 128, 96, 32, 3, ctypes.c_float(6.0),
                                      ctypes.c_float(0.02), 2) == 0
    assert np.array_equal(res, gold["lateral_96x128/out"])


def test_xtrans_golden(gold):
    import ctypes
    from _common import I8P, oracle, p
    cfa = gold["xtrans_96x120/cfa"]
    raw3 = scene.cfa_to_rgb3(cfa, 0, scene.XTRANS)
    res = np.zeros_like(raw3)
    xt = scene.XTRANS.reshape(-1).astype(np.int8).copy()
    assert oracle().oracle_xtrans3(p(raw3), p(res), 120, 96, p(xt, I8P), ctypes.c_float(1.0), TILES_SEQUENTIAL) == 0
    assert np.array_equal(res, gold["xtrans_96x120/out"])


def test_amaze_flat_field_stays_flat():
    """A uniform grey CFA is reproduced on every channel (to the eps terms of the
    colour-ratio interpolation, ~1e-5 relative)."""
    v = 0.37
    raw3 = scene.cfa_to_rgb3(np.full((150, 170), v, np.float32))
    out = oracle_amaze(raw3)
    assert np.abs(out - v).max() < 1e-4


def test_amaze_clip_and_nonnegative():
    cfa = scene.bayer_scene(96, 128, scene.BASE_SEED, noise=True)
    out = oracle_amaze(scene.cfa_to_rgb3(cfa))
    assert out.min() >= 0.0 and np.isfinite(out).all()


def test_aces_closed_form():
    x = np.linspace(0, 4, 97, dtype=np.float32).reshape(1, -1, 1).repeat(3, 2).copy()
    got = oracle_tone_mapping(x, 2.0)
    xd = x.astype(np.float64)
    want = (xd * (2.51 * xd + 0.03)) / (xd * (2.43 * xd + 0.59) + 0.14) * 2.0
    assert np.abs(got - want).max() < 1e-6


def test_detail_tm_flat_below_knee_is_identity():
    """Constant image: variance 0 -> a = 0, base = mean; below 0.55 the blend
    weight is 0 and the image is returned unchanged."""
    rgb = np.full((30, 41, 3), 0.3, np.float32)
    assert np.array_equal(oracle_detail_tonemap(rgb), rgb)


def test_detail_tm_flat_above_knee_closed_form():
    v = 0.8
    rgb = np.full((17, 23, 3), v, np.float32)
    got = oracle_detail_tonemap(rgb)
    aces = (v * (2.51 * v + 0.03)) / (v * (2.43 * v + 0.59) + 0.14)
    t = min(max((v - 0.55) / 0.35, 0), 1)
    t = t * t * (3 - 2 * t)
    gain = 1 + (min(max(aces / v, 0), 1) - 1) * t
    assert np.abs(got - v * gain).max() < 1e-6
