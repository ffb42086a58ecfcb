"""Defringe: the CPU restatement against the reference build (bit-exact, in and
out of place) and the reference's own behavioural invariants."""
import ctypes

import numpy as np
import pytest

import _defringe_cases as cases
from _common import oracle, oracle_amaze, p, ref
from libraw_enhanced_amd import scene


def _run(lib, fn, img, inplace, radius=10.0, edge=0.1, chroma=0.15, strength=10.0, green=False, gscale=0.3):
    h, w, _ = img.shape
    a = np.ascontiguousarray(img, np.float32).copy()
    out = a if inplace else np.zeros_like(a)
    rc = getattr(lib, fn)(p(a), p(out), w, h, ctypes.c_float(radius), ctypes.c_float(edge),
                          ctypes.c_float(chroma), ctypes.c_float(strength), int(green), ctypes.c_float(gscale))
    assert rc == 0
    return out


def oracle_defringe(img, inplace=False, **kw):
    return _run(oracle(), "oracle_defringe", img, inplace, **kw)


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("kw", [dict(), dict(green=True), dict(radius=6.0, strength=2.0, green=True, gscale=0.5),
                                dict(strength=1.0, chroma=0.3)])
def test_defringe_bit_exact_vs_reference(inplace, kw):
    r = ref()
    r.ref_set_threads(4)
    img = oracle_amaze(scene.cfa_to_rgb3(scene.bayer_scene(240, 352, scene.BASE_SEED + 11))) * 1.8
    want = _run(r, "ref_defringe", img, inplace, **kw)
    got = oracle_defringe(img, inplace, **kw)
    assert np.array_equal(got, want)
    assert np.any(got != img)  # the scene exercises the correction


def test_defringe_inplace_differs_from_standalone():
    img = oracle_amaze(scene.cfa_to_rgb3(scene.bayer_scene(240, 352, scene.BASE_SEED + 11))) * 1.8
    a = oracle_defringe(img, inplace=True)
    b = oracle_defringe(img, inplace=False)
    assert np.any(a != b)


def _standalone(impl):
    def f(img, radius=10.0, strength=10.0, defringe_green=False):
        return impl(img, False, radius=radius, strength=strength, green=defringe_green)
    return f


def test_reference_invariants_hold_for_oracle():
    assert cases.check_all(_standalone(oracle_defringe)) == []


def test_reference_invariants_hold_for_reference_build():
    r = ref()
    assert cases.check_all(_standalone(lambda img, inpl, **kw: _run(r, "ref_defringe", img, inpl, **kw))) == []
