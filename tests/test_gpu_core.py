"""The pybind11 module _core on the GPU: each standalone method against the
oracle (tone map, micro-contrast and highlight recovery bit-exact; defringe,
standalone semantics with distinct buffers, within max-abs 2e-4), plus the
array conversions pybind11's array_t<float> does (float64 and non-contiguous
input) and the input left untouched."""
import numpy as np
import pytest

from _common import oracle_tone_mapping, psnr
from libraw_enhanced_amd import scene
from test_defringe_cpu import oracle_defringe
from test_highlight_cpu import magenta_scene, recover
from test_micro_contrast_cpu import _oracle_mc

pytestmark = pytest.mark.gpu
core = pytest.importorskip("libraw_enhanced_amd._core", reason="_core not built")


@pytest.fixture(scope="module")
def w():
    return core.LibRawWrapper()


@pytest.fixture(scope="module")
def img():
    return np.ascontiguousarray(scene.render(240, 320, scene.BASE_SEED + 61) * 1.3, np.float32)


def test_tone_mapping(w, img):
    before = img.copy()
    assert np.array_equal(w.tone_mapping(img, 1.2), oracle_tone_mapping(img, 1.2))
    assert np.array_equal(w.tone_mapping(img), oracle_tone_mapping(img, 1.0))
    assert np.array_equal(img, before)


def test_defringe_standalone(w, img):
    for kw in (dict(), dict(radius=6.0, strength=3.0, defringe_green=True)):
        got = w.defringe(img, **kw)
        want = oracle_defringe(img, False, radius=kw.get("radius", 10.0), strength=kw.get("strength", 10.0),
                               green=kw.get("defringe_green", False))
        assert float(np.abs(got - want).max()) <= 2e-4 and psnr(got, want) >= 70.0, kw


def test_micro_contrast(w, img):
    assert np.array_equal(w.enhance_micro_contrast(img, 0.7, 4.0, 0.1), _oracle_mc(img, 0.7, 4.0, 0.1))
    # threshold < 0: maximum / data_maximum of the loaded raw, 1 when none is loaded
    assert np.array_equal(w.enhance_micro_contrast(img), _oracle_mc(img, 1.0, 8.0, np.float32(0.06)))


def test_recover_highlights(w):
    hl = magenta_scene(64, 96, 9, 1.5)
    want, _, _, _ = recover(hl, 1.0)
    assert np.array_equal(w.recover_highlights(hl), want)


def test_conversions(w, img):
    # float64 and non-contiguous inputs are converted as pybind11's array_t<float> converts them
    want = oracle_tone_mapping(img, 1.0)
    assert np.array_equal(w.tone_mapping(img.astype(np.float64)), want)
    wide = np.zeros((240, 640, 3), np.float32)
    wide[:, ::2] = img
    assert np.array_equal(w.tone_mapping(wide[:, ::2]), want)
