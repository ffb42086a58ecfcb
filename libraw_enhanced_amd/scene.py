"""Seeded synthetic scenes and CFA mosaics (ctypes front end of csrc/scene.c).

Every workload of the develop path is generated here: a smooth gradient,
200 hard-edged polygons, a zone plate up to 0.45 cycles/px, optional radial
lateral-CA pre-warp, CFA sampling and Poisson-Gaussian noise.  The C code owns
its PRNG, so the bytes are identical on every host.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._libs import load_library

BASE_SEED = 20261203
RGGB = 0x94949494
# Standard 6x6 X-Trans layout (0 = R, 1 = G, 2 = B); green is 3-periodic.
XTRANS = np.array(
    [[1, 1, 0, 1, 1, 2],
     [1, 1, 2, 1, 1, 0],
     [2, 0, 1, 0, 2, 1],
     [1, 1, 2, 1, 1, 0],
     [1, 1, 0, 1, 1, 2],
     [0, 2, 1, 2, 0, 1]], dtype=np.int8)

_f32p = ctypes.POINTER(ctypes.c_float)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_i8p = ctypes.POINTER(ctypes.c_int8)


def _lib():
    lib = load_library("liblre_scene.so")
    if not getattr(lib, "_lre_sigs", False):
        lib.lre_scene_render.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        lib.lre_scene_lateral_warp.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double]
        lib.lre_scene_mosaic.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                         _i8p, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        lib.lre_scene_quantize_u16.argtypes = [_f32p, _u16p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_uint32, _i8p, _f32p, ctypes.c_int,
                                               ctypes.c_int]
        lib.lre_scene_cfa_to_rgb3.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_uint32, _i8p]
        lib._lre_sigs = True
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


def render(height: int, width: int, seed: int = BASE_SEED) -> np.ndarray:
    """Clean scene, (H, W, 3) float32 in [0, 1]."""
    out = np.empty((height, width, 3), np.float32)
    _lib().lre_scene_render(_p(out, _f32p), width, height, seed)
    return out


def lateral_warp(rgb: np.ndarray, scale_r: float = 1.0008, scale_b: float = 0.9993) -> np.ndarray:
    """Radially scale R and B about the image centre (bilinear)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    out = np.empty_like(rgb)
    _lib().lre_scene_lateral_warp(_p(rgb, _f32p), _p(out, _f32p), w, h, scale_r, scale_b)
    return out


def mosaic(rgb: np.ndarray, seed: int = BASE_SEED, filters: int = RGGB, xtrans=None,
           gain_inv: float = 4000.0, read_sigma: float = 0.002) -> np.ndarray:
    """Sample through the CFA with Poisson-Gaussian noise -> (H, W) float32 plane."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = rgb.shape
    out = np.empty((h, w), np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(xtrans, np.int8)
    _lib().lre_scene_mosaic(_p(rgb, _f32p), _p(out, _f32p), w, h, filters,
                            None if xt is None else _p(xt, _i8p), gain_inv, read_sigma, seed)
    return out


def quantize_u16(cfa: np.ndarray, filters: int = RGGB, xtrans=None, wb=(2.1, 1.0, 1.5),
                 black: int = 512, white: int = 16383) -> np.ndarray:
    """14-bit raw counts: clamp(round(y / wb_c * (white - black)) + black, 0, white)."""
    cfa = np.ascontiguousarray(cfa, np.float32)
    h, w = cfa.shape
    out = np.empty((h, w), np.uint16)
    wbv = np.asarray(wb, np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(xtrans, np.int8)
    _lib().lre_scene_quantize_u16(_p(cfa, _f32p), _p(out, _u16p), w, h, filters,
                                  None if xt is None else _p(xt, _i8p), _p(wbv, _f32p),
                                  black, white)
    return out


def cfa_to_rgb3(cfa: np.ndarray, filters: int = RGGB, xtrans=None) -> np.ndarray:
    """Demosaic input layout of the reference: (H, W, 3), sample at its native channel."""
    cfa = np.ascontiguousarray(cfa, np.float32)
    h, w = cfa.shape
    out = np.empty((h, w, 3), np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(xtrans, np.int8)
    _lib().lre_scene_cfa_to_rgb3(_p(cfa, _f32p), _p(out, _f32p), w, h, filters,
                                 None if xt is None else _p(xt, _i8p))
    return out


def bayer_scene(height: int, width: int, seed: int = BASE_SEED, noise: bool = True) -> np.ndarray:
    """Noisy RGGB CFA plane of the seeded scene, float32 (H, W)."""
    rgb = render(height, width, seed)
    return mosaic(rgb, seed, RGGB, None, 4000.0 if noise else 0.0, 0.002)


def xtrans_scene(height: int, width: int, seed: int = BASE_SEED, noise: bool = True,
                 lateral=(1.0008, 0.9993)) -> np.ndarray:
    """Noisy X-Trans CFA plane, optionally pre-warped with radial lateral CA."""
    rgb = render(height, width, seed)
    if lateral is not None:
        rgb = lateral_warp(rgb, *lateral)
    return mosaic(rgb, seed, 0, XTRANS, 4000.0 if noise else 0.0, 0.002)


def fcol_bayer(filters: int, height: int, width: int) -> np.ndarray:
    """Per-site colour (0 R, 1 G, 2 B) of a Bayer `filters` word; G2 folds to 1."""
    r = np.arange(height)[:, None]
    c = np.arange(width)[None, :]
    native = (filters >> ((((r << 1) & 14) | (c & 1)) << 1)) & 3
    return np.where(native == 3, 1, native).astype(np.int8)


def fcol_xtrans(height: int, width: int, xtrans=XTRANS) -> np.ndarray:
    r = np.arange(height)[:, None] % 6
    c = np.arange(width)[None, :] % 6
    return np.asarray(xtrans)[r, c].astype(np.int8)


def env_seed(k: int = 0) -> int:
    return int(os.environ.get("LRE_SEED", BASE_SEED)) + k
