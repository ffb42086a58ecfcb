"""The highlight-recovery oracle (core/libraw_wrapper.cpp:1095-1383) pinned on
the CPU.  The reference function sits in the LibRaw-bound wrapper and is not
built here, so each pass is checked against values fixed by hand with dyadic
inputs (exact in float), the smoothstep closed forms of the mask, and
invariants.  The planted snapshot reading of the LoCA pass shows that the
scenes the GPU tests use separate raster order from a parallel reading."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _common import oracle, p
from libraw_enhanced_amd import scene

LL3 = ctypes.c_longlong * 3
RASTER, SNAPSHOT, LOCA_OFF = 0, 1, 2


def recover(img, sat, order=RASTER, wb=None, mask=False):
    out = np.ascontiguousarray(img, np.float32).copy()
    h, w, _ = out.shape
    m = np.zeros((h, w), np.float32) if mask else None
    mv = ctypes.c_float(-1.0)
    cnt = LL3()
    wbp = None if wb is None else p(np.asarray(wb, np.float32))
    rc = oracle().oracle_recover_highlights(p(out), w, h, ctypes.c_float(sat), None if m is None else p(m), wbp,
                                            ctypes.byref(mv), cnt, order)
    assert rc == 0
    return out, m, mv.value, list(cnt)


def magenta_scene(h, w, seed, gain):
    img = scene.render(h, w, scene.BASE_SEED + seed) * gain
    img[..., 0] *= 1.15
    img[..., 2] *= 1.2
    return np.ascontiguousarray(img, np.float32)


def test_below_threshold_is_identity():
    img = scene.render(23, 41, scene.BASE_SEED) * 0.8
    out, _, mv, cnt = recover(img, 1.0)
    assert np.array_equal(out, img) and mv == 0.0 and cnt == [0, 0, 0]


def test_mask_closed_forms():
    g = np.float32([0.5, 0.95, 1.0, 1.3])
    grey = np.repeat(g[None, :, None], 3, axis=2)  # 1 x 4 neutral pixels
    _, m, _, _ = recover(grey, 2.0, mask=True)
    assert np.array_equal(m[0], np.float32([0, 0, 1, 1]))  # smoothstep 0 below 0.95, 1 from 1.0
    mid = np.full((1, 1, 3), 0.975, np.float32)
    _, m, _, _ = recover(mid, 2.0, mask=True)
    assert abs(float(m[0, 0]) - 0.5) < 2e-5  # half-way up the smoothstep
    # (1.5, 1, 1) is grey once the WB (1.5, 1, 1) is undone: neutrality 1
    px = np.float32([[[1.5, 1.0, 1.0]]])
    _, m_wb, _, _ = recover(px, 9.0, wb=(1.5, 1.0, 1.0), mask=True)
    assert m_wb[0, 0] == 1.0
    # without a WB its chroma is 1/3, below the looser lower edge 0.5: neutrality 1
    _, m_no, _, _ = recover(px, 9.0, mask=True)
    assert m_no[0, 0] == 1.0
    # chroma 0.5 between the WB edges (0.3, 0.6): 1 - smoothstep(2/3)
    _, m_r, _, _ = recover(np.float32([[[2.0, 1.0, 1.0]]]), 9.0, wb=(1.0, 1.0, 1.0), mask=True)
    t = (0.5 - 0.3) / 0.3
    assert abs(float(m_r[0, 0]) - (1 - t * t * (3 - 2 * t))) < 1e-5
    # chroma 2/3 between the edges without a WB (0.5, 0.7): 1 - smoothstep(5/6)
    _, m_s, _, _ = recover(np.float32([[[3.0, 1.0, 1.0]]]), 9.0, mask=True)
    t = (2 / 3 - 0.5) / 0.2
    assert abs(float(m_s[0, 0]) - (1 - t * t * (3 - 2 * t))) < 1e-5


def test_uniform_blown_region_turns_neutral():
    img = np.tile(np.float32([1.25, 0.75, 1.5]), (9, 11, 1))
    out, _, mv, cnt = recover(img, 1.0)
    # no ratio pixel: grf = gbf = 1, G = (R + B) / 2, sl = 0; every pixel blown,
    # the 8-neighbour means of equal values are exact, and R + B = 2 G skips LoCA
    assert np.array_equal(out, np.tile(np.float32([1.25, 1.375, 1.5]), (9, 11, 1)))
    assert mv == 1.5 and cnt == [99, 99, 0]


def test_partial_highlight_blend():
    out, _, mv, cnt = recover(np.float32([[[0.75, 0.5, 1.0]]]), 0.5)
    # sl = (0.75 - 0.5) / 0.5 = 0.5; G = 0.875; R = 0.75/2 + 0.875/2, B = 1/2 + 0.875/2
    assert np.array_equal(out[0, 0], np.float32([0.8125, 0.875, 0.9375]))
    assert mv == 1.0 and cnt == [1, 0, 0]


def test_ratio_means_enter_the_rewrite():
    # three ratio pixels with R/G = 1.25, B/G = 1.5, then a highlight pixel with
    # G below the threshold; R / grf = B / gbf = 0.5 makes every step exact
    row = np.float32([[0.625, 0.5, 0.75]] * 3 + [[0.625, 0.25, 0.75]])[None]
    out, _, _, cnt = recover(row, 0.5)
    assert np.array_equal(out[0, 3], np.float32([0.625, 0.5, 0.75]))
    assert cnt == [4, 0, 0]
    # with the ratio pixels gone (grf = gbf = 1) the same pixel is rewritten differently
    out1, _, _, _ = recover(row[:, 3:], 0.5)
    assert not np.array_equal(out1[0, 0], out[0, 3])


def test_blown_means_read_rewritten_neighbours():
    img = np.zeros((5, 5, 3), np.float32)
    vals = np.arange(25, dtype=np.float32).reshape(5, 5) / 32.0  # dyadic, below 1
    img[..., 0] = vals
    img[..., 1] = vals
    img[..., 2] = vals
    img[2, 2] = img[2, 3] = (1.5, 0.5, 1.5)  # blown after the rewrite: (1.5, 1.5, 1.5)
    out, _, _, cnt = recover(img, 1.0)
    post = img.copy()
    post[2, 2] = post[2, 3] = 1.5
    for (y, x) in [(2, 2), (2, 3)]:
        nb = [post[y + dy, x + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dy or dx]
        want = np.sum(np.float64(nb), axis=0) / 8  # exact: dyadic values
        assert np.array_equal(out[y, x], want.astype(np.float32))
    assert cnt[:2] == [2, 2]


def test_loca_known_answer():
    # 5x5, sat 1: the centre is blown, so it takes the mean of its 8 neighbours
    # (magenta, non-highlight): (0.9375, 0.5, 0.9375).  The 3x3 contrast is
    # 0.09375 (edge weight 1), no pixel of the 5x5 window has a G/RB ratio in
    # [0.7, 1.35] (local ratio 1, so desired G = RB), and l = 0.9375 < 1 gives
    # saturation weight 0 and alpha 0.3.
    img = np.full((5, 5, 3), 0.25, np.float32)
    hi, lo = np.float32([0.984375, 0.5, 0.984375]), np.float32([0.890625, 0.5, 0.890625])
    for k, (dy, dx) in enumerate([(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]):
        img[2 + dy, 2 + dx] = hi if k % 2 == 0 else lo
    img[2, 2] = (1.5, 0.5, 1.5)
    out, _, mv, cnt = recover(img, 1.0)
    f = np.float32
    rb, g = f(0.9375), f(0.5)
    alpha = f(0.3)
    boost = alpha * min(f(rb * f(1.0)) - g, (rb - g) * f(0.9))
    g1 = g + boost
    res = max(f(0), f(0.5) * (rb + rb) - g1)
    trim = min(res * alpha * f(0.45), f(0.5) * (rb + rb) * f(0.12))
    assert np.array_equal(out[2, 2], np.float32([rb - trim, g1, rb - trim]))
    assert mv == 1.5 and cnt == [1, 1, 1]
    snap, _, _, _ = recover(img, 1.0, order=SNAPSHOT)  # one candidate: no order to respect
    assert np.array_equal(out, snap)


@pytest.mark.parametrize("seed,gain,sat", [(0, 1.3, 0.9), (1, 1.2, 0.8), (2, 1.6, 1.0)])
def test_loca_invariants(seed, gain, sat):
    img = magenta_scene(160, 220, seed, gain)
    h, w, _ = img.shape
    pre, _, _, _ = recover(img, sat, order=LOCA_OFF)  # the state the LoCA pass starts from
    out, _, _, cnt = recover(img, sat)
    d = (out != pre).any(axis=2)
    assert d.sum() == cnt[2] > 0
    yy, xx = np.nonzero(d)
    assert ((yy >= 2) & (yy <= h - 3) & (xx >= 2) & (xx <= w - 3)).all()
    a, b = pre[d], out[d]
    rb = 0.5 * (a[:, 0] + a[:, 2])
    assert (rb > a[:, 1] * np.float32(1.04)).all()                  # only magenta pixels move
    assert (a.max(axis=1) >= np.float32(sat) * np.float32(0.9)).all()  # bright enough
    assert (b[:, 1] > a[:, 1]).all()                                 # G lifted
    assert (b[:, 0] <= a[:, 0]).all() and (b[:, 2] <= a[:, 2]).all()  # R, B trimmed
    assert (b[:, 1] <= rb * 1.08 * (1 + 1e-6)).all()                 # never past the clamped ratio
    assert (b >= 0).all()


@pytest.mark.parametrize("seed,gain,sat", [(0, 1.3, 0.9), (1, 1.2, 0.8), (2, 1.6, 1.0)])
def test_scenes_separate_raster_from_snapshot(seed, gain, sat):
    img = magenta_scene(200, 300, seed, gain)
    a, _, _, ca = recover(img, sat)
    b, _, _, cb = recover(img, sat, order=SNAPSHOT)
    assert ca[2] > 1000 and (a != b).any(axis=2).sum() > 100
    a2, _, _, _ = recover(img, sat)  # raster order is deterministic
    assert np.array_equal(a, a2)
