"""Output epilogue restated in oracle/epilogue.c: the colour matrix and the
colour-space gamma checked bit for bit against the reference's own
convert_color_space / gamma_correct (oracle/_ref), and the 8/16-bit output
against closed forms (it lives in the pybind11 layer, which is not built)."""
import ctypes

import numpy as np
import pytest

from _common import p, oracle, ref
from libraw_enhanced_amd import scene

MATRICES = [np.eye(3, 4, dtype=np.float32),
            np.array([[1.62, -0.49, -0.13, 0.0], [-0.18, 1.35, -0.17, 0.0], [0.02, -0.31, 1.29, 0.0]],
                     np.float32),
            np.array([[0.41, 0.36, 0.18, 0.01], [0.21, 0.72, 0.07, -0.02], [0.02, 0.12, 0.95, 0.03]],
                     np.float32)]


def _img(h=37, w=53, seed=5):
    rgb = scene.render(h, w, scene.BASE_SEED + seed) * 1.3 - 0.1   # negatives and highlights
    rgb[0, 0] = (0.0, 0.0031308, 0.0031309)
    rgb[0, 1] = (0.018053968, 0.018053970, 1.0)
    return np.ascontiguousarray(rgb, np.float32)


def oracle_color(img, t):
    h, w, _ = img.shape
    out = np.empty_like(img)
    tt = np.ascontiguousarray(t, np.float32).reshape(12)
    assert oracle().oracle_convert_color_space(p(img), p(out), w, h, p(tt)) == 0
    return out


def oracle_gamma(img, gp, gs, cs):
    h, w, _ = img.shape
    out = np.empty_like(img)
    assert oracle().oracle_gamma_correct(p(img), p(out), w, h, ctypes.c_float(gp), ctypes.c_float(gs), cs) == 0
    return out


def oracle_quantize(a, bits):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint8 if bits == 8 else np.uint16)
    assert oracle().oracle_quantize(p(a), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.size), bits) == 0
    return out


@pytest.mark.parametrize("k", range(len(MATRICES)))
def test_color_space_matches_reference(k):
    r = ref()
    if r is None:
        pytest.skip("reference build not available")
    img = _img()
    h, w, _ = img.shape
    want = np.empty_like(img)
    tt = np.ascontiguousarray(MATRICES[k]).reshape(12)
    assert r.ref_convert_color_space(p(img), p(want), w, h, p(tt)) == 0
    assert np.array_equal(oracle_color(img, MATRICES[k]), want)


@pytest.mark.parametrize("cs,gp,gs", [(0, 2.222, 4.5), (0, 2.2, 0.0), (1, 0, 0), (2, 0, 0), (3, 0, 0),
                                      (4, 0, 0), (5, 1.8, 0), (6, 2.4, 12.92), (7, 0, 0), (8, 0, 0)])
def test_gamma_matches_reference(cs, gp, gs):
    r = ref()
    if r is None:
        pytest.skip("reference build not available")
    img = _img(seed=cs)
    h, w, _ = img.shape
    want = np.empty_like(img)
    assert r.ref_gamma_correct(p(img), p(want), w, h, ctypes.c_float(gp), ctypes.c_float(gs), cs) == 0
    assert np.array_equal(oracle_gamma(img, gp, gs, cs), want)


def test_gamma_closed_forms():
    v = np.array([[[0.0, 0.001, 1.0]]], np.float32)
    srgb = oracle_gamma(v, 0, 0, 1)
    assert srgb[0, 0, 0] == 0.0
    assert srgb[0, 0, 1] == np.float32(12.92) * np.float32(0.001)      # linear toe
    assert abs(srgb[0, 0, 2] - 1.0) < 1e-6                              # 1.055 - 0.055
    adobe = oracle_gamma(v, 0, 0, 2)
    assert adobe[0, 0, 0] == 0.0 and adobe[0, 0, 2] == 1.0
    assert abs(adobe[0, 0, 1] - 0.001 ** (1 / 2.222)) < 1e-6
    slope = oracle_gamma(np.array([[[0.1, 0.3, -1.0]]], np.float32), 2.0, 4.0, 0)
    assert slope[0, 0, 0] == np.float32(0.1) * np.float32(4.0)         # below 1/slope: linear
    assert abs(slope[0, 0, 1] - 0.3 ** 0.5) < 1e-6 and slope[0, 0, 2] == 0.0


def test_quantize_closed_forms():
    v = np.array([-1.0, 0.0, 0.001, 0.5, 0.998, 1.0, 2.0, np.float32(127.5 / 255)], np.float32)
    assert oracle_quantize(v, 8).tolist() == [0, 0, 0, 128, 254, 255, 255, 128]
    assert oracle_quantize(v, 16).tolist() == [0, 0, 66, 32768, 65404, 65535, 65535, 32768]
    ramp = np.linspace(-0.1, 1.1, 10001, dtype=np.float32)
    q = oracle_quantize(ramp, 8).astype(int)
    assert np.all(np.diff(q) >= 0) and q.min() == 0 and q.max() == 255
    want = np.floor(np.clip(ramp.astype(np.float64), 0, 1) * 255 + 0.5).astype(int)
    assert np.count_nonzero(q != want) <= 2        # float rounding at exact .5 ties only
