"""The AMaZE kernel clears only part of its per-tile scratch (the CFA plane,
the borders of the colour-difference planes, and the Nyquist flags after
stage 5).  The oracle runs the same minimal clear on scratch filled with NaN or
with a large value; its output must equal the fully cleared (clean-tile) run
for every tile geometry: first/interior/last tiles, partial tiles, tiles whose
border fill runs past the scratch arrays (heights and widths 128k + 1..15),
all four Bayer phases, and content that trips the Nyquist test everywhere."""
import numpy as np
import pytest

from _common import RGGB, oracle_amaze
from libraw_enhanced_amd import scene

POISON_NAN, POISON_BIG = 2, 3
SIZES = [64, 130, 141, 159, 200, 270, 333]
PHASES = [RGGB, 0x16161616, 0x61616161, 0x49494949]


def _content(kind, h, w, f, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == "noise":
        cfa = rng.random((h, w), dtype=np.float32)
    elif kind == "nyquist":
        yy, xx = np.mgrid[0:h, 0:w]
        cfa = (0.5 + 0.4 * np.cos(np.pi * 0.9 * (xx + yy)) + 0.05 * rng.random((h, w))).astype(np.float32)
    else:
        cfa = scene.mosaic(scene.render(h, w, seed), seed, f)
    return scene.cfa_to_rgb3(cfa, f)


@pytest.mark.parametrize("kind", ["scene", "noise", "nyquist"])
def test_minimal_clear_equals_clean_tiles(kind):
    n = 0
    for h in SIZES:
        for w in SIZES:
            f = PHASES[n % 4]
            raw3 = _content(kind, h, w, f, 100 + n)
            clean = oracle_amaze(raw3, f, (2.1, 1, 1.5, 1), 1.0, mode=0)
            assert np.isfinite(clean).all()
            for mode in (POISON_NAN, POISON_BIG):
                got = oracle_amaze(raw3, f, (2.1, 1, 1.5, 1), 1.0, mode=mode)
                assert np.array_equal(got, clean), (h, w, hex(f), mode)
            n += 1
