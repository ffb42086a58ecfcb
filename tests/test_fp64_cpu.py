"""The FP64 evaluations of the float stages (oracle/fp64.h): the oracle's own
formulas in double.  Pinned by closed forms that hold exactly in any precision
and by their distance to the float oracle, which must be of the order of float
rounding (a plausible slip in the double build, such as an allocation sized for
floats, a float libm call left in, or a constant changed, moves them far
apart)."""
import numpy as np
import pytest

from _common import (fp64_axial_ca, fp64_defringe, fp64_detail_tonemap, fp64_lateral_ca, max_abs, oracle,
                     oracle_amaze, oracle_detail_tonemap)
from libraw_enhanced_amd import scene
from test_ca_cpu import axial, ca_scene, lateral
from test_defringe_cpu import oracle_defringe


@pytest.fixture(scope="module")
def img():
    return ca_scene(240, 360)


def test_detail_tonemap_fp64_near_float(img):
    a, b = oracle_detail_tonemap(img * 1.6), fp64_detail_tonemap(img * 1.6)
    err = max_abs(a, b)
    print(f"detail TM float vs fp64: {err:.3g}")
    assert 0 < err <= 2e-5


def test_detail_tonemap_fp64_constant_image():
    # var = 0 -> a = 0, b = mean = c, base = c: every channel scaled by the
    # smoothstep-blended ACES gain of c (core/libraw_wrapper.cpp:1398-1471)
    c = 0.8
    out = fp64_detail_tonemap(np.full((40, 50, 3), c))
    aces = (c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14)
    # the constants are the float literals of the reference, held as doubles
    f = lambda x: float(np.float32(x))
    aces = (c * (f(2.51) * c + f(0.03))) / (c * (f(2.43) * c + f(0.59)) + f(0.14))
    w = min(max((c - f(0.55)) / (f(0.90) - f(0.55)), 0.0), 1.0)
    w = w * w * (3 - 2 * w)
    gain = 1 + (min(max(aces / c, 0.0), 1.0) - 1) * w
    assert np.allclose(out, c * gain, rtol=0, atol=1e-12)


@pytest.mark.parametrize("inplace", [False, True])
def test_defringe_fp64_near_float(img, inplace):
    src = img * 1.8
    a, b = oracle_defringe(src, inplace), fp64_defringe(src, inplace)
    d = np.abs(a - b)
    frac = float(np.mean(np.any(d > 1e-5, axis=2)))
    print(f"defringe float vs fp64 (inplace={inplace}): max-abs {d.max():.3g}, {frac:.2%} px beyond 1e-5")
    assert d.max() > 0 and np.percentile(d, 99.9) <= 1e-5 and frac < 0.005


def test_defringe_fp64_grey_is_identity():
    # R = G = B: no chroma excess anywhere, every mask is 0 and the image passes through
    g = scene.render(64, 80, 5)[..., 1]
    grey = np.repeat(g[..., None], 3, axis=2)
    assert np.array_equal(fp64_defringe(grey, False), grey.astype(np.float64))


def test_axial_fp64(img):
    a, b = axial(oracle(), "oracle_axial_ca", img), fp64_axial_ca(img)
    err = max_abs(a, b)
    print(f"axial float vs fp64: {err:.3g}")
    assert 0 < err <= 1e-5
    # strength 0: R and B pass through the normalisation only
    z = fp64_axial_ca(img, strength=0.0)
    assert np.allclose(z, img.astype(np.float64), rtol=0, atol=1e-12)


def test_lateral_fp64(img):
    a, b = lateral(oracle(), "oracle_lateral_ca", img), fp64_lateral_ca(img)
    err = max_abs(a, b)
    print(f"lateral float vs fp64: {err:.3g}")
    assert 0 < err <= 1e-4
    g = np.repeat(img[..., 1:2], 3, axis=2)
    assert np.allclose(fp64_lateral_ca(g), g.astype(np.float64), rtol=0, atol=1e-12)
