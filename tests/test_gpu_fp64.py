"""Error of every float stage against an FP64 evaluation of the same stage
(oracle/fp64.h), for the HIP output and for the reference's own CPU code (the
build in oracle/_ref, else the float restatement): the HIP output must be no
further from the FP64 result than the reference is, give or take the stage
tolerance (max-abs 2e-4)."""
import numpy as np
import pytest

from _common import (ROOT, fp64_axial_ca, fp64_defringe, fp64_detail_tonemap, fp64_lateral_ca, max_abs, oracle,
                     oracle_amaze, oracle_detail_tonemap)
from libraw_enhanced_amd import scene

pytestmark = pytest.mark.gpu
TOL = 2e-4


@pytest.fixture(scope="module")
def acc():
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    a = HIPAccelerator(0)
    yield a
    a.close()


@pytest.fixture(scope="module")
def img():
    h, w = 600, 900
    rgb = scene.lateral_warp(scene.render(h, w, scene.BASE_SEED + 41), 1.0008, 0.9993)
    return oracle_amaze(scene.cfa_to_rgb3(scene.mosaic(rgb, scene.BASE_SEED + 41))) * 1.6


def _cpu(name):
    """The reference build's function if it was built, else the restatement's."""
    import os
    from _common import ref
    if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so")):
        r = ref()
        r.ref_set_threads(8)
        return r, "ref_" + name, "reference"
    return oracle(), "oracle_" + name, "restatement"


def _check(stage, hip, cpu, f64, kind):
    eh, ec = max_abs(hip, f64), max_abs(cpu, f64)
    print(f"{stage}: max-abs vs FP64: HIP {eh:.3g}, {kind} {ec:.3g}", flush=True)
    assert eh <= ec + TOL


@pytest.mark.parametrize("inplace", [True, False])
def test_defringe_vs_fp64(acc, img, inplace):
    from test_defringe_cpu import _run
    lib, fn, kind = _cpu("defringe")
    a = img.copy()
    hip = acc.defringe(a, a if inplace else None)
    cpu = _run(lib, fn, img, inplace)
    _check(f"defringe inplace={inplace}", hip, cpu, fp64_defringe(img, inplace), kind)


def test_detail_tonemap_vs_fp64(acc, img):
    lib, fn, kind = _cpu("detail_tonemap")
    hip = acc.apply_detail_preserving_tonemap(img)
    cpu = img.copy()
    import ctypes
    getattr(lib, fn)(cpu.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), img.shape[1], img.shape[0])
    _check("detail TM", hip, cpu, fp64_detail_tonemap(img), kind)


def test_axial_vs_fp64(acc, img):
    from test_ca_cpu import axial
    lib, fn, kind = _cpu("axial_ca")
    hip = acc.ca_axial_cleanup(img, 6, 1e-4, 0.3)
    _check("axial CA", hip, axial(lib, fn, img), fp64_axial_ca(img), kind)


def test_lateral_vs_fp64(acc, img):
    from test_ca_cpu import lateral
    lib, fn, kind = _cpu("lateral_ca")
    hip = acc.ca_register_lateral(img, 96, 3, 6.0, 0.02, 3)
    _check("lateral CA", hip, lateral(lib, fn, img), fp64_lateral_ca(img), kind)
