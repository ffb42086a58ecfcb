"""The measured chains at their full sizes, and the X-Trans chain of the
lateral-CA scene, against the oracle chain on the GPU.

* 24 MP Bayer (4000 x 6000): AMaZE -> defringe (in place) -> detail tone map,
  the chain the throughput figure is quoted on;
* 26 MP X-Trans (4160 x 6240), pre-warped with radial lateral CA (R 1.0008,
  B 0.9993): X-Trans 3-pass -> lateral CA -> axial CA, both CA stages in place
  as the develop pipeline calls them (core/libraw_wrapper.cpp:1776,1794);
* the same X-Trans chain at 1040 x 1560.

The demosaics must be bit-exact; every later stage output within max-abs 2e-4
and PSNR 70 dB of the oracle chain.  Each float stage is also compared with
an FP64 evaluation of the chain (oracle/fp64.h) from the same demosaic
output: the HIP chain must be no further from it than the oracle chain is,
give or take the tolerance.
"""
import numpy as np
import pytest

from _common import (fp64_axial_ca, fp64_defringe, fp64_detail_tonemap, fp64_lateral_ca, max_abs, oracle,
                     oracle_amaze, oracle_detail_tonemap, psnr)
from libraw_enhanced_amd import scene

pytestmark = pytest.mark.gpu

TOL, MIN_PSNR = 2e-4, 70.0


@pytest.fixture(scope="module")
def acc():
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    a = HIPAccelerator(0)
    yield a
    a.close()


def _report(name, got, want):
    err = float(np.abs(got - want).max())
    q = psnr(got, want)
    ndiff = int(np.count_nonzero(np.any(got != want, axis=2)))
    print(f"{name}: max-abs {err:.3g}, PSNR {q:.1f} dB, {ndiff} px differ", flush=True)
    return err, q


def _fp64(name, got, want, f64):
    eh, eo = max_abs(got, f64), max_abs(want, f64)
    print(f"{name}: max-abs vs FP64 chain: HIP {eh:.3g}, oracle {eo:.3g}", flush=True)
    assert eh <= eo + TOL


def _run_stages(acc, chain, d_cfa):
    outs = []
    for s in chain.stages:
        chain.run_stage(s, d_cfa)
        acc.synchronize()
        outs.append(chain.result())
    return outs


def test_chain24_full_size(acc):
    from libraw_enhanced_amd.pipeline import DevelopChain
    from test_gpu_develop import _oracle_defringe
    h, w = 4000, 6000
    print(f"chain24 {w}x{h}: generating", flush=True)
    cfa = scene.bayer_scene(h, w, scene.BASE_SEED)
    d_cfa = acc.dev_from_host(cfa)
    chain = DevelopChain(acc, h, w, ["amaze", "defringe", "detail_tm"])
    g_amaze, g_def, g_tm = _run_stages(acc, chain, d_cfa)
    print("chain24: oracle chain", flush=True)
    o_amaze = oracle_amaze(scene.cfa_to_rgb3(cfa))
    bad = int(np.count_nonzero(g_amaze != o_amaze))
    assert bad == 0, f"AMaZE: {bad} values differ"
    o_def = _oracle_defringe(o_amaze, True)
    err, q = _report("chain24 defringe", g_def, o_def)
    assert err <= TOL and q >= MIN_PSNR
    o_tm = oracle_detail_tonemap(o_def)
    err, q = _report("chain24 detail TM", g_tm, o_tm)
    assert err <= TOL and q >= MIN_PSNR
    print("chain24: FP64 chain", flush=True)
    f_def = fp64_defringe(o_amaze, True)
    _fp64("chain24 defringe", g_def, o_def, f_def)
    _fp64("chain24 detail TM", g_tm, o_tm, fp64_detail_tonemap(f_def))


@pytest.mark.parametrize("h,w", [(1040, 1560), (4160, 6240)])
def test_xtrans_lateral_axial_chain(acc, h, w):
    from libraw_enhanced_amd.pipeline import DevelopChain
    from test_ca_cpu import axial, lateral
    from test_xtrans_cpu import oracle_xtrans
    print(f"xtrans chain {w}x{h}: generating (radial pre-warp)", flush=True)
    cfa = scene.xtrans_scene(h, w, scene.BASE_SEED, lateral=(1.0008, 0.9993))
    d_cfa = acc.dev_from_host(cfa)
    chain = DevelopChain(acc, h, w, ["xtrans3", "lateral_ca", "axial_ca"], filters=9, xtrans=scene.XTRANS)
    g_x, g_lat, g_ax = _run_stages(acc, chain, d_cfa)
    print(f"xtrans chain {w}x{h}: oracle chain", flush=True)
    o_x = oracle_xtrans(scene.cfa_to_rgb3(cfa, 0, scene.XTRANS))
    bad = int(np.count_nonzero(np.any(g_x != o_x, axis=2)))
    assert bad == 0, f"X-Trans: {bad} px differ"
    o_lat = lateral(oracle(), "oracle_lateral_ca", o_x, inplace=True)
    err, q = _report(f"xtrans {w}x{h} lateral CA", g_lat, o_lat)
    assert err <= TOL and q >= MIN_PSNR
    o_ax = axial(oracle(), "oracle_axial_ca", o_lat, inplace=True)
    err, q = _report(f"xtrans {w}x{h} axial CA", g_ax, o_ax)
    assert err <= TOL and q >= MIN_PSNR
    f_lat = fp64_lateral_ca(o_x, inplace=True)
    _fp64(f"xtrans {w}x{h} lateral CA", g_lat, o_lat, f_lat)
    _fp64(f"xtrans {w}x{h} axial CA", g_ax, o_ax, fp64_axial_ca(f_lat, inplace=True))
