"""Device-resident develop chain on one image: the CFA plane and every
intermediate stay in HBM, stages are enqueued back to back on the context's
stream with no host synchronisation between them.

Stage names follow the reference's Accelerator methods; the default chain is
the 24 MP Bayer develop: AMaZE demosaic -> defringe -> detail tone map.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .accelerator import RGGB, HIPAccelerator, _check, native

_vp = ctypes.c_void_p
_f32p = ctypes.POINTER(ctypes.c_float)


class DevelopChain:
    """Enqueue a fixed stage list over one resident image."""

    def __init__(self, acc: HIPAccelerator, height: int, width: int, stages=None,
                 filters: int = RGGB, cam_mul=(1.0, 1.0, 1.0, 1.0), maximum_value: float = 1.0,
                 defringe_params=None, xtrans=None):
        self.acc = acc
        self.h, self.w = height, width
        self.filters = filters
        self.xtrans = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(36), np.int8)
        self.cam_mul = np.ascontiguousarray(cam_mul, np.float32)
        self.maximum_value = float(maximum_value)
        self.defringe_params = dict(radius=10.0, edge_threshold=0.1, chroma_threshold=0.15,
                                    strength=10.0, defringe_green=False, green_strength_scale=0.3)
        if defringe_params:
            self.defringe_params.update(defringe_params)
        n = height * width
        self.rgb = acc.dev_alloc(n * 12)
        self.stages = list(stages or available_chain())
        # the detail tone map writes into a second buffer, which then becomes the
        # image: with a separate output its guide comes straight from the input
        # pixels (no guide-plane pass), with the same result as in place
        self.rgb_alt = acc.dev_alloc(n * 12) if "detail_tm" in self.stages else None

    def run_stage(self, name: str, d_cfa, stream=None):
        lib, ctx, h, w = native(), self.acc.handle, self.h, self.w
        if name == "amaze":
            _check(lib.lre_hip_dev_amaze(ctx, d_cfa.ptr, self.rgb.ptr, w, h, self.filters,
                                         self.cam_mul.ctypes.data_as(_f32p), self.maximum_value,
                                         stream))
        elif name == "xtrans3":
            _check(lib.lre_hip_dev_xtrans3(ctx, d_cfa.ptr, self.rgb.ptr, w, h,
                                           self.xtrans.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                           self.maximum_value, stream))
        elif name == "lateral_ca":
            # in place, as the develop pipeline calls it (core/libraw_wrapper.cpp:1776)
            _check(lib.lre_hip_dev_lateral_ca(ctx, self.rgb.ptr, self.rgb.ptr, w, h, 96, 3, 6.0, 0.02, 3,
                                              stream))
        elif name == "axial_ca":
            _check(lib.lre_hip_dev_axial_ca(ctx, self.rgb.ptr, self.rgb.ptr, w, h, 6, 1e-4, 0.3, stream))
        elif name == "defringe":
            p = self.defringe_params
            # in place, as the develop pipeline calls it (core/libraw_wrapper.cpp:1804-1818)
            _check(lib.lre_hip_dev_defringe(ctx, self.rgb.ptr, self.rgb.ptr, w, h, p["radius"],
                                            p["edge_threshold"], p["chroma_threshold"],
                                            p["strength"], int(p["defringe_green"]),
                                            p["green_strength_scale"], stream))
        elif name == "detail_tm":
            if self.rgb_alt is None:
                self.rgb_alt = self.acc.dev_alloc(h * w * 12)
            _check(lib.lre_hip_dev_detail_tonemap(ctx, self.rgb.ptr, self.rgb_alt.ptr, w, h, stream))
            self.rgb, self.rgb_alt = self.rgb_alt, self.rgb
        elif name == "tone_mapping":
            _check(lib.lre_hip_dev_tone_mapping(ctx, self.rgb.ptr, self.rgb.ptr, w, h, 1.0, stream))
        else:
            raise ValueError(f"unknown stage {name}")

    def run(self, d_cfa, stream=None):
        for s in self.stages:
            self.run_stage(s, d_cfa, stream)

    def result(self) -> np.ndarray:
        return self.rgb.download((self.h, self.w, 3), np.float32)


CHAIN24 = ("amaze", "defringe", "detail_tm")


def available_chain():
    """The 24 MP Bayer chain: AMaZE -> defringe -> detail tone map.  Raises if
    the loaded library lacks one of its entry points (a timed chain must never
    silently run fewer stages)."""
    lib = native()
    for fn in ("lre_hip_dev_amaze", "lre_hip_dev_defringe", "lre_hip_dev_detail_tonemap"):
        if not hasattr(lib, fn):
            raise RuntimeError(f"liblre_hip.so lacks {fn}: rebuild it (python __graft_entry__.py)")
    return list(CHAIN24)
