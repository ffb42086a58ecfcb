"""Integer preprocessing: known answers for black subtraction/clip and the
adjust_maximum rule, and the fused ingest checked against the reference's own
white balance (oracle/_ref) applied after black subtraction."""
import ctypes

import numpy as np
import pytest

from _common import I8P, RGGB, U16P, oracle, p, place_native_u16, ref
from libraw_enhanced_amd import scene

INT4 = ctypes.c_int * 4


def black_level(img4, cblack=(0, 0, 0, 0), black=0, maximum=16383, pos=None):
    img = img4.copy()
    h, w, _ = img.shape
    if pos is not None:
        pt = np.ascontiguousarray(pos, np.int32)
        ptr, pr, pc = pt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), pt.shape[0], pt.shape[1]
    else:
        ptr, pr, pc = None, 0, 0
    assert oracle().oracle_black_level(p(img, U16P), w, h, INT4(*cblack), ptr, pr, pc, black, maximum) == 0
    return img


def ingest(raw, filters=RGGB, xtrans=None, cblack=(0, 0, 0, 0), black=512, maximum=16383,
           wb=(2.1, 1.0, 1.5, 1.0)):
    h, w = raw.shape
    out = np.zeros((h, w), np.float32)
    dm = ctypes.c_float()
    xt = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(-1), np.int8)
    wbv = np.asarray(wb, np.float32)
    o = oracle()
    o.oracle_ingest_u16.argtypes = None
    assert o.oracle_ingest_u16(p(np.ascontiguousarray(raw, np.uint16), U16P), p(out), w, h, ctypes.c_uint32(filters),
                               None if xt is None else p(xt, I8P), INT4(*cblack), black, maximum, p(wbv),
                               ctypes.byref(dm)) == 0
    return out, dm.value


def test_black_level_known_answers():
    img = np.array([[[0, 511, 512, 513], [16383, 16000, 600, 20000]]], np.uint16)
    assert black_level(img, black=512, maximum=15871).tolist() == [[[0, 0, 0, 1], [15871, 15488, 88, 15871]]]
    out = black_level(img, cblack=(10, 20, 30, 40), maximum=16383)
    assert out.tolist() == [[[0, 491, 482, 473], [16373, 15980, 570, 16383]]]
    pos = np.array([[1, 2], [3, 4]])
    out = black_level(img, cblack=(10, 0, 0, 0), pos=pos, maximum=16383)
    # pixel (0,0): pos 1; pixel (0,1): pos 2
    assert out.tolist() == [[[0, 510, 511, 512], [16371, 15998, 598, 16383]]]


def test_adjust_maximum_rule():
    o = oracle()
    o.oracle_adjust_maximum.restype = ctypes.c_float
    f = lambda d, m, t: o.oracle_adjust_maximum(ctypes.c_float(d), ctypes.c_float(m), ctypes.c_float(t))
    assert f(14000.0, 15871.0, 0.75) == 14000.0      # within (0.75 max, max): adopted
    assert f(10000.0, 15871.0, 0.75) == 15871.0      # below the threshold
    assert f(33000.0, 15871.0, 0.75) == 15871.0      # above the white level
    assert f(14000.0, 15871.0, 0.0) == 15871.0       # threshold ~0 disables
    assert f(12000.0, 15871.0, 1.0) == 12000.0       # threshold > 0.99999 -> 0.75


@pytest.mark.parametrize("filters,xtrans", [(RGGB, None), (9, scene.XTRANS)])
def test_ingest_matches_reference_white_balance(filters, xtrans):
    h, w = 90, 132
    if xtrans is None:
        cfa = scene.bayer_scene(h, w, 3)
    else:
        cfa = scene.xtrans_scene(h, w, 3, lateral=None)
    raw = scene.quantize_u16(cfa, filters if xtrans is None else 0, xtrans)
    got, dmax = ingest(raw, filters, xtrans)
    img4 = black_level(place_native_u16(raw, filters, xtrans), black=512, maximum=16383)
    wb = np.array([2.1, 1.0, 1.5, 1.0], np.float32)
    want3 = np.zeros((h, w, 3), np.float32)
    xt = (scene.XTRANS if xtrans is None else xtrans).reshape(-1).astype(np.int8).copy()
    assert ref().ref_white_balance(p(img4, U16P), p(want3), w, h, p(wb), ctypes.c_uint32(filters), p(xt, I8P)) == 0
    assert np.array_equal(got, want3.sum(axis=2))
    # data maximum over each pixel's fcol_bayer channel (core/libraw_wrapper.cpp:1051)
    fc = scene.fcol_bayer(filters, h, w)
    assert dmax == float(np.take_along_axis(want3, fc[..., None].astype(np.int64), axis=2).max())
