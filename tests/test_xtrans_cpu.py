"""X-Trans 3-pass: the CPU restatement against the reference build."""
import ctypes

import numpy as np
import pytest

from _common import I8P, TILES_CLEAN, TILES_SEQUENTIAL, oracle, p, ref
from libraw_enhanced_amd import scene

XT = scene.XTRANS.reshape(-1).astype(np.int8).copy()


def oracle_xtrans(raw3, mode=TILES_CLEAN, maximum=1.0):
    h, w, _ = raw3.shape
    out = np.zeros_like(raw3)
    assert oracle().oracle_xtrans3(p(raw3), p(out), w, h, p(XT, I8P), ctypes.c_float(maximum), mode) == 0
    return out


def ref_xtrans(raw3, threads=1, maximum=1.0):
    h, w, _ = raw3.shape
    out = np.zeros_like(raw3)
    r = ref()
    r.ref_set_threads(threads)
    assert r.ref_xtrans3(p(raw3), p(out), w, h, p(XT, I8P), ctypes.c_float(maximum)) == 0
    return out


def xtrans_raw3(h, w, seed, scale=1.0):
    cfa = scene.xtrans_scene(h, w, seed) * scale
    return scene.cfa_to_rgb3(cfa, 0, scene.XTRANS)


@pytest.mark.parametrize("h,w", [(120, 150), (300, 500), (640, 960), (1040, 1560)])
def test_xtrans_sequential_bit_exact(h, w):
    raw3 = xtrans_raw3(h, w, scene.BASE_SEED + w)
    assert np.array_equal(oracle_xtrans(raw3, TILES_SEQUENTIAL), ref_xtrans(raw3, 1))


def test_xtrans_maximum_scaling():
    raw3 = xtrans_raw3(200, 260, 7, 4000.0)
    assert np.array_equal(oracle_xtrans(raw3, TILES_SEQUENTIAL, 4000.0), ref_xtrans(raw3, 1, 4000.0))


def test_xtrans_clean_tiles_differ_only_in_bottom_right_band():
    """Stale homogeneity rows/columns of the last tile row/column (read before
    written in the reference) reach only the 12-px bottom/right band."""
    h, w = 640, 960
    raw3 = xtrans_raw3(h, w, scene.BASE_SEED + 2)
    d = np.any(oracle_xtrans(raw3, TILES_CLEAN) != ref_xtrans(raw3, 1), axis=2)
    ys, xs = np.nonzero(d)
    assert np.all((h - ys <= 12) | (w - xs <= 12))


def test_xtrans_flat_grey_stays_flat():
    raw3 = scene.cfa_to_rgb3(np.full((150, 170), 0.3, np.float32), 0, scene.XTRANS)
    out = oracle_xtrans(raw3)
    assert np.abs(out - 0.3).max() < 1e-5
