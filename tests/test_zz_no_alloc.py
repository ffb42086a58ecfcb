"""lre_hip_dev_amaze_noalloc (tile arenas in the library's fixed device array,
no allocation) against lre_hip_dev_amaze (arenas in the context's scratch),
through the C ABI: bit-identical output on two calls in a row, the second on
a smaller image, so the second call runs on arenas the first one left dirty
and on a shorter tile queue."""
import ctypes as ct

import numpy as np
import pytest

from _common import RGGB
from libraw_enhanced_amd import scene

pytestmark = pytest.mark.gpu


def _amaze(acc, fn, d_cfa, h, w, filters, cam, maxv):
    from libraw_enhanced_amd.accelerator import native
    d_rgb = acc.dev_from_host(np.full((h, w, 3), np.nan, np.float32))
    c = np.asarray(cam, np.float32)
    rc = getattr(native(), fn)(acc.handle, d_cfa.ptr, d_rgb.ptr, w, h, filters,
                               c.ctypes.data_as(ct.POINTER(ct.c_float)), maxv, None)
    assert rc == 0, native().lre_hip_last_error()
    acc.synchronize()
    return d_rgb.download((h, w, 3))


def test_noalloc_amaze_matches_allocating_entry():
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    acc = HIPAccelerator(0)
    # 1000x1500: 96 tiles over the 32 fixed arenas; then 200x260 (4 tiles), GBRG, WB gain
    for h, w, filters, cam, maxv in [(1000, 1500, RGGB, (1, 1, 1, 1), 1.0),
                                     (200, 260, 0x49494949, (2.1, 1.0, 1.5, 1.0), 0.9)]:
        print(f"noalloc amaze {h}x{w} filters {filters:#x}", flush=True)
        cfa = scene.mosaic(scene.render(h, w, scene.BASE_SEED + h), scene.BASE_SEED + h, filters)
        d_cfa = acc.dev_from_host(cfa)
        want = _amaze(acc, "lre_hip_dev_amaze", d_cfa, h, w, filters, cam, maxv)
        got = _amaze(acc, "lre_hip_dev_amaze_noalloc", d_cfa, h, w, filters, cam, maxv)
        bad = int(np.count_nonzero(got != want))
        assert bad == 0 and not np.isnan(got).any(), f"{h}x{w}: {bad} values differ"
    # argument errors are reported, not run
    assert native().lre_hip_dev_amaze_noalloc(acc.handle, None, None, 8, 8, RGGB, None, 1.0, None) == -1
    acc.close()
