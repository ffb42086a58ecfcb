"""HIP highlight recovery vs the oracle (core/libraw_wrapper.cpp:1095-1383),
through the C ABI: image, soft mask and max_val bit-exact.  The magenta scenes
have thousands of LoCA-adjusted pixels whose raster-order result differs from a
parallel (snapshot) reading, so a kernel that ignores the order fails here."""
import ctypes

import numpy as np
import pytest

from libraw_enhanced_amd import scene
from test_highlight_cpu import LOCA_OFF, SNAPSHOT, magenta_scene, recover

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def acc():
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    a = HIPAccelerator(0)
    yield a
    a.close()


def _check(got, want, what):
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{what}: {bad} values differ, max-abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("h,w,seed,gain,sat,wb,mask", [
    (200, 300, 0, 1.3, 0.9, None, False),
    (200, 300, 1, 1.2, 0.8, (2.0, 1.0, 1.5), True),
    (333, 517, 2, 1.6, 1.0, None, True),
    (130, 700, 3, 1.3, 0.85, None, False),
    (67, 131, 4, 1.4, 0.9, (1.8, 1.0, 1.3), True),
    (1000, 1500, 5, 1.3, 0.9, (2.0, 1.0, 1.5), True),
])
def test_recover_highlights_bit_exact(acc, h, w, seed, gain, sat, wb, mask):
    print(f"recover_highlights {h}x{w} seed {seed} sat {sat} wb {wb} mask {mask}", flush=True)
    img = magenta_scene(h, w, seed, gain)
    got, gmask, gmax = acc.recover_highlights(img, sat, want_mask=mask, wb_rgb=wb)
    want, wmask, wmax, cnt = recover(img, sat, wb=wb, mask=mask)
    assert cnt[2] > 0 or h * w < 10000
    _check(got, want, "image")
    if mask:
        _check(gmask, wmask, "mask")
    assert gmax == wmax


@pytest.mark.parametrize("h,w", [(1, 1), (1, 9), (2, 5), (3, 3), (4, 6), (5, 5), (6, 130), (70, 4), (65, 200)])
def test_tiny_frames(acc, h, w):
    print(f"recover_highlights tiny {h}x{w}", flush=True)
    img = magenta_scene(h, w, h * 31 + w, 1.5)
    got, gmask, gmax = acc.recover_highlights(img, 0.85, want_mask=True)
    want, wmask, wmax, _ = recover(img, 0.85, mask=True)
    _check(got, want, "image")
    _check(gmask, wmask, "mask")
    assert gmax == wmax


def test_known_answer_images(acc):
    print("recover_highlights known answers", flush=True)
    imgs = []
    a = np.full((5, 5, 3), 0.25, np.float32)
    for k, (dy, dx) in enumerate([(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]):
        a[2 + dy, 2 + dx] = (0.984375, 0.5, 0.984375) if k % 2 == 0 else (0.890625, 0.5, 0.890625)
    a[2, 2] = (1.5, 0.5, 1.5)
    imgs.append((a, 1.0))
    imgs.append((np.tile(np.float32([1.25, 0.75, 1.5]), (9, 11, 1)), 1.0))
    imgs.append((np.float32([[0.625, 0.5, 0.75]] * 3 + [[0.625, 0.25, 0.75]])[None].copy(), 0.5))
    imgs.append((scene.render(23, 41, scene.BASE_SEED) * 0.8, 1.0))  # nothing to recover
    for img, sat in imgs:
        got, _, gmax = acc.recover_highlights(img, sat)
        want, _, wmax, _ = recover(img, sat)
        _check(got, want, f"{img.shape}")
        assert gmax == wmax


def test_order_matters_on_the_gpu_scene(acc):
    """The GPU result equals the raster order and not the snapshot reading."""
    img = magenta_scene(200, 300, 1, 1.2)
    got, _, _ = acc.recover_highlights(img, 0.8)
    snap, _, _, _ = recover(img, 0.8, order=SNAPSHOT)
    pre, _, _, _ = recover(img, 0.8, order=LOCA_OFF)
    assert (got != snap).any() and (got != pre).any()


def test_device_entry_in_place(acc):
    from libraw_enhanced_amd.accelerator import native
    print("recover_highlights device entry, in place", flush=True)
    h, w = 300, 410
    img = magenta_scene(h, w, 7, 1.3)
    want, wmask, wmax, _ = recover(img, 0.9, wb=(2.0, 1.0, 1.5), mask=True)
    d_img = acc.dev_from_host(img)
    d_mask = acc.dev_alloc(h * w * 4)
    d_max = acc.dev_alloc(4)
    wb = np.float32([2.0, 1.0, 1.5])
    lib = native()
    for _ in range(2):  # the second call starts from the first's scratch state
        d_img.upload(img)
        rc = lib.lre_hip_dev_recover_highlights(acc.handle, d_img.ptr, d_img.ptr, w, h, ctypes.c_float(0.9),
                                                d_mask.ptr, wb.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                d_max.ptr, None)
        assert rc == 0
        _check(d_img.download((h, w, 3)), want, "image")
        _check(d_mask.download((h, w)), wmask, "mask")
        assert d_max.download((1,))[0] == wmax


def test_standalone_wrapper(acc):
    from libraw_enhanced_amd.accelerator import LibRawWrapper
    img = magenta_scene(64, 96, 9, 1.5)
    want, _, _, _ = recover(img, 1.0)
    assert np.array_equal(LibRawWrapper(0).recover_highlights(img), want)
