"""The micro-contrast oracle (core/cpu_accelerator.cpp:3135-3319) pinned on
the CPU.  The reference's own code for this block needs Apple's <simd/simd.h>
and is not built here, so the oracle is checked against what the mathematics
fixes: the Gaussian's shape and normalisation, a float64 evaluation of the
same local statistics (a slip in an index, a sign or a term moves the result
far beyond float rounding), and the cases that must return the input.

The reference also states the algorithm a second time, in its Metal shader
(core/metal/enhance_micro_contrast.metal): separable 9-tap passes with a
rounded weight table, variance as blur(I^2) - blur(I)^2, and one blend
in + (enhanced - in) w for both gates.  A float64 evaluation written from that
statement, structured as the shader is, must agree with the oracle (which
follows the CPU code), so a misreading of either statement fails here."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _common import F32P, oracle, p
from libraw_enhanced_amd import scene


def _kernel():
    k = np.zeros(81, np.float32)
    oracle().oracle_mc_kernel(p(k))
    return k.reshape(9, 9)


def _oracle_mc(img, threshold=0.5, strength=8.0, target=0.06, mask=None):
    h, w, _ = img.shape
    src = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(src)
    m = None if mask is None else p(np.ascontiguousarray(mask, np.float32))
    rc = oracle().oracle_micro_contrast(p(src), p(out), w, h, ctypes.c_float(threshold),
                                        ctypes.c_float(strength), ctypes.c_float(target), m)
    assert rc == 0
    return out


def _stats64(img, k):
    """Local mean and deviation with the same 9x9 weights, float64, edge-clamped."""
    h, w, _ = img.shape
    x = img.astype(np.float64)
    pad = np.pad(x, ((4, 4), (4, 4), (0, 0)), mode="edge")
    mean = np.zeros_like(x)
    msq = np.zeros_like(x)
    for ky in range(9):
        for kx in range(9):
            win = pad[ky:ky + h, kx:kx + w]
            mean += win * float(k[ky, kx])
            msq += win * win * float(k[ky, kx])
    sd = np.sqrt(np.maximum(msq - mean * mean, 0.0))
    return mean, sd


def _mc64(img, k, threshold, strength, target, mask):
    mean, sd = _stats64(img, k)
    cm = sd / sd.max()
    f = np.where(cm < target, strength * (target - cm) / target, 0.0)
    e = mean + (img - mean) * (1.0 + f)
    mx = e.max(axis=2, keepdims=True)
    e = np.where(mx > 1.0, e / mx, e)
    if mask is None:
        gate = (img[:, :, 1] >= threshold)[:, :, None]
        return np.where(gate, e, img)
    wt = np.clip(mask, 0.0, 1.0)[:, :, None]
    return img + (e - img) * wt


def test_kernel_is_the_normalised_gaussian():
    k = _kernel()
    assert np.allclose(k, k.T) and np.allclose(k, k[::-1, ::-1])
    assert abs(float(k.sum(dtype=np.float64)) - 1.0) < 1e-6
    y, x = np.mgrid[-4:5, -4:5]
    sigma = 5.0 / 3.7
    g = np.exp(-(x * x + y * y) / (2.0 * sigma * sigma))
    assert np.allclose(k, g / g.sum(), rtol=2e-6, atol=1e-9)
    assert k[4, 4] == k.max()


@pytest.mark.parametrize("use_mask", [False, True])
def test_oracle_matches_float64_statistics(use_mask):
    h, w = 61, 83
    img = scene.render(h, w, scene.BASE_SEED + 3) * 1.2
    mask = scene.render(h, w, scene.BASE_SEED + 4)[:, :, 1] if use_mask else None
    got = _oracle_mc(img, 0.4, 8.0, 0.06, mask)
    want = _mc64(img.astype(np.float64), _kernel(), 0.4, 8.0, 0.06, mask)
    assert np.abs(got - want).max() < 5e-5


def test_gates_that_return_the_input():
    h, w = 37, 52
    img = scene.render(h, w, scene.BASE_SEED + 5)
    assert np.array_equal(_oracle_mc(img, threshold=np.inf), img)      # no pixel passes the gate
    assert np.array_equal(_oracle_mc(img, mask=np.zeros((h, w))), img)  # zero weight everywhere
    # no gain: mean + (I - mean) returns the input to rounding (values below 1)
    same = _oracle_mc(img * 0.9, threshold=0.0, strength=0.0)
    assert np.abs(same - img * 0.9).max() < 1e-6
    # a flat image has no deviation anywhere (0 / 0 contrast: no gain)
    flat = np.full((h, w, 3), 0.3, np.float32)
    assert np.allclose(_oracle_mc(flat, threshold=0.0), flat, atol=1e-7)


def test_inplace_equals_out_of_place():
    h, w = 40, 33
    img = np.ascontiguousarray(scene.render(h, w, scene.BASE_SEED + 6), np.float32)
    want = _oracle_mc(img, 0.3)
    buf = img.copy()
    assert oracle().oracle_micro_contrast(p(buf), p(buf), w, h, ctypes.c_float(0.3), ctypes.c_float(8.0),
                                          ctypes.c_float(0.06), ctypes.cast(None, F32P)) == 0
    assert np.array_equal(buf, want)


# core/metal/enhance_micro_contrast.metal:13-23, the shader's 9-tap weights
METAL_WEIGHTS = np.array([0.003693, 0.025109, 0.098814, 0.224689, 0.295406, 0.224689, 0.098814, 0.025109,
                          0.003693])


def test_kernel_marginal_is_the_metal_table():
    # The CPU code's 2D kernel is separable, and its 1D factor is the shader's
    # table up to the table's own error: the table (which sums to 1.000016)
    # departs from the exact normalised Gaussian of sigma 5/3.7 by up to
    # 3.3e-5, while sigma 1.35 would already move the centre weight by 3e-4.
    k = _kernel().astype(np.float64)
    m = k.sum(axis=0)
    assert np.abs(k - np.outer(k.sum(axis=1), m)).max() < 1e-7  # float32 weights
    assert np.abs(m - METAL_WEIGHTS).max() < 4e-5


def _metal_statement(img, threshold, strength, target, mask, weights):
    """core/metal/enhance_micro_contrast.metal:26-165 in float64: horizontal then
    vertical clamp-to-edge passes over I and I*I with the given 1D weights, then
    the per-pixel kernel (contrast, gain, rescale, one blend for both gates)."""
    h, w, _ = img.shape
    x = img.astype(np.float64)

    def blur(v):
        hp = np.zeros_like(v)
        for k in range(9):
            hp += v[:, np.clip(np.arange(w) - 4 + k, 0, w - 1)] * weights[k]
        out = np.zeros_like(v)
        for k in range(9):
            out += hp[np.clip(np.arange(h) - 4 + k, 0, h - 1)] * weights[k]
        return out

    mean = blur(x)
    var = blur(x * x) - mean * mean
    std = np.sqrt(np.maximum(var, 0.0))
    contrast = std / std.max()
    factor = np.where(contrast < target, strength * (target - contrast) / target, 0.0)
    enhanced = mean + (x - mean) * (1.0 + factor)
    maxc = enhanced.max(axis=2, keepdims=True)
    enhanced = np.where(maxc > 1.0, enhanced / maxc, enhanced)
    if mask is None:
        wt = (x[:, :, 1] >= threshold).astype(np.float64)
    else:
        wt = np.clip(mask.astype(np.float64), 0.0, 1.0)
    return x + (enhanced - x) * wt[:, :, None]


@pytest.mark.parametrize("use_mask", [False, True])
@pytest.mark.parametrize("strength,target", [(8.0, 0.06), (8.0, 0.04), (3.0, 0.3)])
def test_oracle_matches_the_metal_statement(use_mask, strength, target):
    h, w = 57, 71
    img = scene.render(h, w, scene.BASE_SEED + 8) * 1.3
    mask = scene.render(h, w, scene.BASE_SEED + 9)[:, :, 0] if use_mask else None
    # the shader's structure with the CPU code's weights (the table itself is
    # pinned above): what is left to differ is the oracle's float rounding,
    # scaled by the gain (up to 1 + strength) and by the square root where the
    # variance nears 0; a slip in the formula moves outputs by the size of
    # the enhancement (above 1e-2 here)
    weights = _kernel().astype(np.float64).sum(axis=0)
    got = _oracle_mc(img, 0.45, strength, target, mask).astype(np.float64)
    want = _metal_statement(img, 0.45, strength, target, mask, weights)
    err = np.abs(got - want)
    assert err.max() < 2e-4, err.max()
    assert np.mean(err < 1e-5) > 0.98
    # the enhancement did act on this scene (the comparison is not vacuous)
    assert np.abs(got - img).max() > 1e-2
