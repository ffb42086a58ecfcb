"""Axial and lateral CA: the CPU restatement against the reference build
(bit-exact) and properties the mathematics fixes."""
import ctypes

import numpy as np
import pytest

from _common import oracle, oracle_amaze, p, ref
from libraw_enhanced_amd import scene


def axial(lib, fn, img, inplace=False, radius=6, eps=1e-4, strength=0.3):
    h, w, _ = img.shape
    a = np.ascontiguousarray(img, np.float32).copy()
    out = a if inplace else np.zeros_like(a)
    assert getattr(lib, fn)(p(a), p(out), w, h, radius, ctypes.c_float(eps), ctypes.c_float(strength)) == 0
    return out


def lateral(lib, fn, img, inplace=False, cell=96, iters=3, shift=6.0, conf=0.02, levels=3):
    h, w, _ = img.shape
    a = np.ascontiguousarray(img, np.float32).copy()
    out = a if inplace else np.zeros_like(a)
    assert getattr(lib, fn)(p(a), p(out), w, h, cell, iters, ctypes.c_float(shift), ctypes.c_float(conf), levels) == 0
    return out


def ca_scene(h, w, seed=scene.BASE_SEED + 3):
    rgb = scene.lateral_warp(scene.render(h, w, seed), 1.0008, 0.9993)
    return oracle_amaze(scene.cfa_to_rgb3(scene.mosaic(rgb, seed)))


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("kw", [dict(), dict(radius=3, eps=1e-3, strength=1.0)])
def test_axial_bit_exact_vs_reference(inplace, kw):
    img = ca_scene(260, 390)
    r = ref()
    r.ref_set_threads(4)
    assert np.array_equal(axial(oracle(), "oracle_axial_ca", img, inplace, **kw),
                          axial(r, "ref_axial_ca", img, inplace, **kw))


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("kw", [dict(), dict(cell=48, iters=5, shift=3.0, levels=2)])
def test_lateral_bit_exact_vs_reference(inplace, kw):
    img = ca_scene(400, 600)
    r = ref()
    r.ref_set_threads(4)
    assert np.array_equal(lateral(oracle(), "oracle_lateral_ca", img, inplace, **kw),
                          lateral(r, "ref_lateral_ca", img, inplace, **kw))


def test_axial_strength_zero_is_identity_up_to_normalisation():
    img = ca_scene(120, 160)
    out = axial(oracle(), "oracle_axial_ca", img, strength=0.0)
    # w = 0: R*norm_inv*norm_ref, G passes through
    assert np.array_equal(out[..., 1], img[..., 1])
    assert np.abs(out - img).max() <= 1e-6 * max(1.0, float(img.max()))


def test_axial_constant_image_unchanged():
    img = np.full((50, 70, 3), 0.4, np.float32)
    img[..., 0] = 0.3
    out = axial(oracle(), "oracle_axial_ca", img, strength=1.0)
    assert np.abs(out - img).max() < 1e-6


def test_lateral_grey_image_is_identity():
    """R = G = B: zero error everywhere, every shift stays 0, the warp samples
    at integer positions and returns the input."""
    g = scene.render(200, 300, 9)[..., 1:2]
    img = np.repeat(g, 3, axis=2).copy()
    out = lateral(oracle(), "oracle_lateral_ca", img)
    assert np.array_equal(out, img)


def test_lateral_recovers_radial_shift():
    """Grey scene (brightness constancy holds between channels) with R and B
    radially pre-warped: after registration R and B move back toward G near
    the left and right edges, where the shift is largest."""
    h, w = 600, 900
    g = scene.render(h, w, scene.BASE_SEED + 8)[..., 1:2]
    clean = np.repeat(g, 3, axis=2).copy()
    warped = scene.lateral_warp(clean, 1.003, 0.997)
    out = lateral(oracle(), "oracle_lateral_ca", warped)
    band = (slice(40, h - 40), slice(0, 120))
    before = np.abs(warped[band][..., 0] - clean[band][..., 0]).mean()
    after = np.abs(out[band][..., 0] - clean[band][..., 0]).mean()
    assert after < 0.8 * before
