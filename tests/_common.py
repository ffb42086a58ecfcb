"""Shared helpers of the test suite: loaders for the CPU restatement (the
oracle, the parity checker), the reference build in oracle/_ref, and metrics."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32P = ctypes.POINTER(ctypes.c_float)
U16P = ctypes.POINTER(ctypes.c_uint16)
I8P = ctypes.POINTER(ctypes.c_int8)
RGGB = 0x94949494
TILES_CLEAN, TILES_SEQUENTIAL = 0, 1


def p(a, t=F32P):
    return a.ctypes.data_as(t)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liblre_oracle.so")
        if not os.path.exists(path):
            pytest.skip("oracle not built (run __graft_entry__.build())")
        _oracle = ctypes.CDLL(path)
    return _oracle


def ref():
    """The reference's own CPU code, built by oracle/ref_build.sh (absent where
    the reference tree was not available at build time)."""
    global _ref
    if _ref is None:
        path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so")
        if not os.path.exists(path):
            pytest.skip("reference build oracle/_ref not present")
        _ref = ctypes.CDLL(path)
    return _ref


def oracle_amaze(raw3, filters=RGGB, cam_mul=(1, 1, 1, 1), maximum=1.0, mode=TILES_CLEAN):
    h, w, _ = raw3.shape
    out = np.zeros_like(raw3)
    cm = np.asarray(cam_mul, np.float32)
    rc = oracle().oracle_amaze(p(raw3), p(out), w, h, ctypes.c_uint32(filters), p(cm),
                               ctypes.c_float(maximum), mode)
    assert rc == 0
    return out


def ref_amaze(raw3, filters=RGGB, cam_mul=(1, 1, 1, 1), maximum=1.0, threads=1):
    h, w, _ = raw3.shape
    out = np.zeros_like(raw3)
    cm = np.asarray(cam_mul, np.float32)
    r = ref()
    r.ref_set_threads(threads)
    rc = r.ref_amaze(p(raw3), p(out), w, h, ctypes.c_uint32(filters), p(cm), ctypes.c_float(maximum))
    assert rc == 0
    return out


def oracle_detail_tonemap(rgb):
    out = np.ascontiguousarray(rgb, np.float32).copy()
    h, w, _ = out.shape
    assert oracle().oracle_detail_tonemap(p(out), w, h) == 0
    return out


def oracle_tone_mapping(rgb, after_scale):
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    out = np.empty_like(rgb)
    assert oracle().oracle_tone_mapping(p(rgb), p(out), w, h, ctypes.c_float(after_scale)) == 0
    return out


def oracle_white_balance(img4, wb, filters, xtrans=None, rgb=None):
    h, w, _ = img4.shape
    out = np.zeros((h, w, 3), np.float32) if rgb is None else rgb.copy()
    wbv = np.asarray(wb, np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(-1), np.int8)
    assert oracle().oracle_white_balance(p(img4, U16P), p(out), w, h, p(wbv), ctypes.c_uint32(filters),
                                         None if xt is None else p(xt, I8P)) == 0
    return out


def psnr(a, b, peak=1.0):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(peak * peak / mse)


def place_native_u16(u16_plane, filters=RGGB, xtrans=None):
    """(H, W) raw counts -> the reference's uint16 x 4 image, sample at its
    native LibRaw channel (0..3, second green = 3 for Bayer)."""
    h, w = u16_plane.shape
    img4 = np.zeros((h, w, 4), np.uint16)
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    if filters == 9:
        native = np.asarray(xtrans)[r % 6, c % 6]
    else:
        native = (filters >> ((((r << 1) & 14) | (c & 1)) << 1)) & 3
    np.put_along_axis(img4, native[..., None].astype(np.int64), u16_plane[..., None], axis=2)
    return img4


# ---- FP64 evaluations of the float stages (oracle/fp64.h: the oracle's own
# source compiled with double in place of float); arrays of float64 ----------
F64P = ctypes.POINTER(ctypes.c_double)


def _d(a):
    return np.ascontiguousarray(a, np.float64).copy()


def fp64_defringe(img, inplace, radius=10.0, edge=0.1, chroma=0.15, strength=10.0, green=False, gscale=0.3):
    h, w, _ = img.shape
    a = _d(img)
    out = a if inplace else np.zeros_like(a)
    D = ctypes.c_double
    assert oracle().oracle64_defringe(p(a, F64P), p(out, F64P), w, h, D(radius), D(edge), D(chroma), D(strength),
                                      int(green), D(gscale)) == 0
    return out


def fp64_detail_tonemap(img):
    a = _d(img)
    h, w, _ = a.shape
    assert oracle().oracle64_detail_tonemap(p(a, F64P), w, h) == 0
    return a


def fp64_axial_ca(img, inplace=False, radius=6, eps=1e-4, strength=0.3):
    h, w, _ = img.shape
    a = _d(img)
    out = a if inplace else np.zeros_like(a)
    D = ctypes.c_double
    assert oracle().oracle64_axial_ca(p(a, F64P), p(out, F64P), w, h, radius, D(eps), D(strength)) == 0
    return out


def fp64_lateral_ca(img, inplace=False, cell=96, iters=3, shift=6.0, conf=0.02, levels=3):
    h, w, _ = img.shape
    a = _d(img)
    out = a if inplace else np.zeros_like(a)
    D = ctypes.c_double
    assert oracle().oracle64_lateral_ca(p(a, F64P), p(out, F64P), w, h, cell, iters, D(shift), D(conf),
                                        levels) == 0
    return out


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))
