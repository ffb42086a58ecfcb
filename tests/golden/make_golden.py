"""Writes the golden fixtures of the develop path from the reference's own CPU
code (oracle/_ref/libref_cpu.so, built from the reference sources by
oracle/ref_build.sh) run with one OpenMP thread.

Each fixture holds the input and the reference's output; the inputs come from
the seeded scene generator.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from libraw_enhanced_amd import scene  # noqa: E402

F32P = ctypes.POINTER(ctypes.c_float)
U16P = ctypes.POINTER(ctypes.c_uint16)
I8P = ctypes.POINTER(ctypes.c_int8)


def P(a, t=F32P):
    return a.ctypes.data_as(t)


def main():
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so"))
    ref.ref_set_threads(1)
    out = {}
    # AMaZE: (name, h, w, filters, cam_mul, maximum, seed)
    cases = [("amaze_rggb_64x96", 64, 96, 0x94949494, (1, 1, 1, 1), 1.0, 1),
             ("amaze_bggr_gain_80x112", 80, 112, 0x16161616, (2.1, 1, 1.5, 1), 2.0, 2),
             ("amaze_grbg_144x176", 144, 176, 0x61616161, (1, 1, 1, 1), 1.0, 3)]
    for name, h, w, filters, cm, mx, k in cases:
        rgb = scene.render(h, w, scene.BASE_SEED + k) * mx
        cfa = scene.mosaic(rgb, scene.BASE_SEED + k, filters, None, 4000.0 / mx, 0.002 * mx)
        raw3 = scene.cfa_to_rgb3(cfa, filters)
        res = np.zeros_like(raw3)
        cmv = np.asarray(cm, np.float32)
        assert ref.ref_amaze(P(raw3), P(res), w, h, ctypes.c_uint32(filters), P(cmv), ctypes.c_float(mx)) == 0
        out[name + "/cfa"] = cfa
        out[name + "/params"] = np.array([filters, *cm, mx], np.float64)
        out[name + "/out"] = res
    # detail tone map (in place)
    rgb = scene.render(40, 56, scene.BASE_SEED + 4) * 1.6
    res = rgb.copy()
    ref.ref_detail_tonemap(P(res), 56, 40)
    out["detail_tm_40x56/in"], out["detail_tm_40x56/out"] = rgb, res
    # ACES tone mapping
    rgb = scene.render(16, 24, scene.BASE_SEED + 5) * 3.0
    res = np.zeros_like(rgb)
    ref.ref_tone_mapping(P(rgb), P(res), 24, 16, ctypes.c_float(1.3))
    out["tone_16x24/in"], out["tone_16x24/out"] = rgb, res
    # white balance, Bayer and X-Trans, u16 x 4 input
    for name, filters, xt in (("wb_bayer_24x36", 0x94949494, None), ("wb_xtrans_24x36", 9, scene.XTRANS)):
        cfa = scene.mosaic(scene.render(24, 36, 7), 7, filters if xt is None else 0, xt)
        u16 = scene.quantize_u16(cfa, filters if xt is None else 0, xt)
        img4 = np.zeros((24, 36, 4), np.uint16)
        r = np.arange(24)[:, None]
        c = np.arange(36)[None, :]
        native = (np.asarray(xt)[r % 6, c % 6] if xt is not None
                  else (filters >> ((((r << 1) & 14) | (c & 1)) << 1)) & 3)
        np.put_along_axis(img4, native[..., None].astype(np.int64), u16[..., None], axis=2)
        wb = np.array([2.1, 1.0, 1.5, 1.0], np.float32)
        res = np.zeros((24, 36, 3), np.float32)
        xtv = (xt if xt is not None else scene.XTRANS).reshape(-1).astype(np.int8)
        assert ref.ref_white_balance(P(img4, U16P), P(res), 36, 24, P(wb), ctypes.c_uint32(filters), P(xtv, I8P)) == 0
        out[name + "/img4"], out[name + "/out"] = img4, res
    # defringe, in and out of place, on an AMaZE output (the AMaZE fixture above)
    img = out["amaze_rggb_64x96/out"] * 1.8
    for name, inplace, params in (("defringe_oop_64x96", False, (10.0, 0.1, 0.15, 10.0, 0, 0.3)),
                                  ("defringe_inplace_green_64x96", True, (6.0, 0.1, 0.15, 2.0, 1, 0.5))):
        a = img.copy()
        res = a if inplace else np.zeros_like(a)
        r_, e_, c_, s_, g_, gs_ = params
        assert ref.ref_defringe(P(a), P(res), 96, 64, ctypes.c_float(r_), ctypes.c_float(e_), ctypes.c_float(c_),
                                ctypes.c_float(s_), g_, ctypes.c_float(gs_)) == 0
        out[name + "/in"], out[name + "/out"] = img, res
        out[name + "/params"] = np.array([r_, e_, c_, s_, g_, gs_, int(inplace)], np.float64)
    # chromatic aberration on a radially pre-warped scene
    rgb = scene.lateral_warp(scene.render(96, 128, scene.BASE_SEED + 6), 1.003, 0.997)
    res = np.zeros_like(rgb)
    assert ref.ref_axial_ca(P(rgb), P(res), 128, 96, 6, ctypes.c_float(1e-4), ctypes.c_float(0.3)) == 0
    out["axial_96x128/in"], out["axial_96x128/out"] = rgb, res
    res = np.zeros_like(rgb)
    assert ref.ref_lateral_ca(P(rgb), P(res), 128, 96, 32, 3, ctypes.c_float(6.0), ctypes.c_float(0.02), 2) == 0
    out["lateral_96x128/in"], out["lateral_96x128/out"] = rgb, res
    # X-Trans 3-pass on a single tile
    xt = scene.XTRANS.reshape(-1).astype(np.int8).copy()
    cfa = scene.xtrans_scene(96, 120, scene.BASE_SEED + 7)
    raw3 = scene.cfa_to_rgb3(cfa, 0, scene.XTRANS)
    res = np.zeros_like(raw3)
    assert ref.ref_xtrans3(P(raw3), P(res), 120, 96, P(xt, I8P), ctypes.c_float(1.0)) == 0
    out["xtrans_96x120/cfa"], out["xtrans_96x120/out"] = cfa, res
    path = os.path.join(HERE, "develop_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
