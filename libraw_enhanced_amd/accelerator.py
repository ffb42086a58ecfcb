"""Host-side mirror of the reference's Accelerator / LibRawWrapper methods for
the develop path, running on the HIP backend through the C ABI of
liblre_hip.so (include/lre_hip.h).

Method names, argument meaning and defaults follow the reference:
  Accelerator methods          core/accelerator.h:247-375
  standalone numpy methods     core/python_bindings.cpp:324-351,
                               core/libraw_wrapper.cpp:3098-3214
There is no CPU fallback here: if the native library or the GPU is missing the
call raises, so a result always comes from the HIP kernels.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from ._libs import load_library

_f32p = ctypes.POINTER(ctypes.c_float)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_i8p = ctypes.POINTER(ctypes.c_int8)
_vp = ctypes.c_void_p

FILTERS_XTRANS = 9
RGGB = 0x94949494


class HIPError(RuntimeError):
    """A HIP entry point returned a non-zero status."""


class LibRawInvalidArgument(ValueError):
    """Mirror of the reference's LibRawInvalidArgument (python_bindings.cpp:503-505)."""


_SIGS = {
    "lre_hip_available": ([], ctypes.c_int),
    "lre_hip_create": ([ctypes.POINTER(_vp), ctypes.c_int], ctypes.c_int),
    "lre_hip_destroy": ([_vp], ctypes.c_int),
    "lre_hip_last_error": ([], ctypes.c_char_p),
    "lre_hip_synchronize": ([_vp], ctypes.c_int),
    "lre_hip_stream": ([_vp], _vp),
    "lre_hip_dev_malloc": ([_vp, ctypes.POINTER(_vp), ctypes.c_size_t], ctypes.c_int),
    "lre_hip_dev_free": ([_vp, _vp], ctypes.c_int),
    "lre_hip_memcpy_htod": ([_vp, _vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "lre_hip_memcpy_dtoh": ([_vp, _vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "lre_hip_amaze": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _f32p,
                       ctypes.c_float], ctypes.c_int),
    "lre_hip_xtrans3": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, _i8p, ctypes.c_float],
                        ctypes.c_int),
    "lre_hip_dev_xtrans3": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _i8p, ctypes.c_float, _vp],
                            ctypes.c_int),
    "lre_hip_tone_mapping": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float],
                             ctypes.c_int),
    "lre_hip_detail_tonemap": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "lre_hip_micro_contrast": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, _f32p], ctypes.c_int),
    "lre_hip_dev_micro_contrast": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, _vp, _vp], ctypes.c_int),
    "lre_hip_recover_highlights": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _f32p,
                                    _f32p, _f32p], ctypes.c_int),
    "lre_hip_dev_recover_highlights": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _f32p,
                                        _vp, _vp], ctypes.c_int),
    "lre_hip_defringe": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                          ctypes.c_float], ctypes.c_int),
    "lre_hip_dev_defringe": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                              ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                              ctypes.c_float, _vp], ctypes.c_int),
    "lre_hip_axial_ca": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "lre_hip_lateral_ca": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int], ctypes.c_int),
    "lre_hip_dev_axial_ca": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_float, ctypes.c_float, _vp], ctypes.c_int),
    "lre_hip_dev_lateral_ca": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, _vp],
                               ctypes.c_int),
    "lre_hip_black_level": ([_vp, _u16p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int], ctypes.c_int),
    "lre_hip_dev_ingest_u16": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _i8p,
                                ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, _f32p, _vp,
                                _vp], ctypes.c_int),
    "lre_hip_white_balance": ([_vp, _u16p, _f32p, ctypes.c_int, ctypes.c_int, _f32p,
                               ctypes.c_uint32, _i8p], ctypes.c_int),
    "lre_hip_dev_amaze": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _f32p,
                           ctypes.c_float, _vp], ctypes.c_int),
    "lre_hip_dev_amaze_noalloc": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _f32p,
                                   ctypes.c_float, _vp], ctypes.c_int),
    "lre_hip_dev_tone_mapping": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp],
                                 ctypes.c_int),
    "lre_hip_dev_detail_tonemap": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "lre_hip_dev_rgb3_to_cfa": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _i8p,
                                 _vp], ctypes.c_int),
    "lre_hip_convert_color_space": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, _f32p], ctypes.c_int),
    "lre_hip_gamma_correct": ([_vp, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                               ctypes.c_float, ctypes.c_int], ctypes.c_int),
    "lre_hip_quantize": ([_vp, _f32p, _vp, ctypes.c_size_t, ctypes.c_int], ctypes.c_int),
    "lre_hip_green_matching": ([_vp, _u16p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int],
                               ctypes.c_int),
    "lre_hip_dev_green_matching": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.c_int, _vp], ctypes.c_int),
    "lre_hip_remove_zeroes": ([_vp, _u16p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32], ctypes.c_int),
    "lre_hip_auto_wb": ([_vp, _u16p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _i8p, _f32p], ctypes.c_int),
    "lre_hip_dev_auto_wb": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _i8p, _vp, _vp],
                            ctypes.c_int),
    "lre_hip_scale_aber": ([_vp, _u16p, ctypes.c_int, ctypes.c_int, _f32p], ctypes.c_int),
    "lre_hip_dev_scale_aber": ([_vp, _vp, ctypes.c_int, ctypes.c_int, _f32p, _vp], ctypes.c_int),
    "lre_hip_dev_remove_zeroes": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _vp],
                                  ctypes.c_int),
    "lre_hip_dev_convert_color_space": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _f32p, _vp],
                                        ctypes.c_int),
    "lre_hip_dev_gamma_correct": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_int, _vp], ctypes.c_int),
    "lre_hip_dev_epilogue": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_float,
                              ctypes.c_float, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
}



class WbMeta(ctypes.Structure):
    """lre_wb_meta: the LibRaw fields scale_colors reads and writes."""
    _fields_ = [("user_mul", ctypes.c_float * 4), ("use_auto_wb", ctypes.c_int),
                ("use_camera_wb", ctypes.c_int), ("camerawb_daylight_fallback", ctypes.c_int),
                ("as_shot_wb_applied", ctypes.c_int), ("colors", ctypes.c_int), ("filters", ctypes.c_uint32),
                ("cam_mul", ctypes.c_float * 4), ("pre_mul", ctypes.c_float * 4),
                ("white", (ctypes.c_uint16 * 8) * 8), ("cblack", ctypes.c_uint * 10), ("black", ctypes.c_uint),
                ("maximum", ctypes.c_uint), ("warnings", ctypes.c_uint)]


_SIGS["lre_wb_needs_auto"] = ([ctypes.POINTER(WbMeta)], ctypes.c_int)
_SIGS["lre_scale_colors_meta"] = ([ctypes.POINTER(WbMeta), _f32p, _f32p], ctypes.c_int)

_lock = threading.Lock()


def native():
    """The loaded liblre_hip.so with its ctypes signatures."""
    lib = load_library("liblre_hip.so")
    with _lock:
        if not getattr(lib, "_lre_sigs", False):
            for name, (args, res) in _SIGS.items():
                fn = getattr(lib, name, None)
                if fn is None:
                    continue
                fn.argtypes = args
                fn.restype = res
            lib._lre_sigs = True
    return lib


def _check(rc: int):
    if rc != 0:
        msg = native().lre_hip_last_error().decode(errors="replace")
        if rc == -1:
            raise LibRawInvalidArgument(msg)
        raise HIPError(f"status {rc}: {msg}")


def _rgb(a, name="image"):
    a = np.asarray(a)
    if a.ndim != 3 or a.shape[2] != 3 or a.dtype != np.float32:
        raise LibRawInvalidArgument(f"{name} must be a (H, W, 3) float32 array")
    return np.ascontiguousarray(a)


def _p(a, t=_f32p):
    return a.ctypes.data_as(t)


class HIPAccelerator:
    """One HIP context (device, stream, scratch) — the drop-in backend."""

    def __init__(self, device: int = 0):
        lib = native()
        if not lib.lre_hip_available():
            raise HIPError("no HIP device available")
        h = _vp()
        _check(lib.lre_hip_create(ctypes.byref(h), device))
        self._h = h
        self.device = device

    # -- lifecycle --
    def close(self):
        if getattr(self, "_h", None):
            native().lre_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _check(native().lre_hip_synchronize(self._h))

    # -- device buffers --
    def dev_alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def dev_from_host(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        buf.upload(arr)
        return buf

    # -- Accelerator methods (host arrays in, new host arrays out) --
    def demosaic_bayer_amaze(self, raw_rgb, filters: int = RGGB, cam_mul=(1.0, 1.0, 1.0, 1.0),
                             maximum_value: float = 1.0):
        raw = _rgb(raw_rgb, "raw_buffer")
        h, w, _ = raw.shape
        out = np.empty_like(raw)
        cm = np.ascontiguousarray(cam_mul, np.float32)
        _check(native().lre_hip_amaze(self._h, _p(raw), _p(out), w, h, filters, _p(cm),
                                      maximum_value))
        return out

    def demosaic_xtrans_3pass(self, raw_rgb, xtrans, maximum_value: float = 1.0):
        raw = _rgb(raw_rgb, "raw_buffer")
        h, w, _ = raw.shape
        # a new zeroed buffer, as the reference's pipeline allocates: frames
        # with a side of 22 px or less have pixels no tile writes
        out = np.zeros_like(raw)
        xt = np.ascontiguousarray(np.asarray(xtrans).reshape(-1), np.int8)
        _check(native().lre_hip_xtrans3(self._h, _p(raw), _p(out), w, h, _p(xt, _i8p), maximum_value))
        return out

    def tone_mapping(self, rgb, after_scale: float = 1.0):
        src = _rgb(rgb)
        h, w, _ = src.shape
        out = np.empty_like(src)
        _check(native().lre_hip_tone_mapping(self._h, _p(src), _p(out), w, h, after_scale))
        return out

    def apply_detail_preserving_tonemap(self, rgb):
        src = _rgb(rgb)
        h, w, _ = src.shape
        out = np.empty_like(src)
        _check(native().lre_hip_detail_tonemap(self._h, _p(src), _p(out), w, h))
        return out

    def enhance_micro_contrast(self, rgb_input, threshold: float, strength: float = 8.0,
                               target_contrast: float = 0.06, mask=None):
        """Accelerator::enhance_micro_contrast (core/accelerator.h:338-341).  mask:
        (H, W) float32 weights in [0, 1] blending input and result; None gates the
        enhancement on green >= threshold."""
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, np.float32)
            if m.shape != (h, w):
                raise LibRawInvalidArgument(f"mask must have shape {(h, w)}, got {m.shape}")
        out = np.empty_like(src)
        _check(native().lre_hip_micro_contrast(self._h, _p(src), _p(out), w, h, threshold, strength,
                                               target_contrast, None if m is None else _p(m)))
        return out

    def recover_highlights(self, rgb, saturation_threshold: float, want_mask: bool = False, wb_rgb=None):
        """LibRawWrapper::Impl::recover_highlights (core/libraw_wrapper.cpp:1095-1383)
        on a copy of rgb.  Returns (image, mask or None, max_val): the soft
        highlight mask of the input is computed only when want_mask, wb_rgb (R, G,
        B) only changes that mask, and max_val is the largest channel of the
        highlight pixels before the rewrite."""
        src = _rgb(rgb)
        h, w, _ = src.shape
        out = np.empty_like(src)
        mask = np.empty((h, w), np.float32) if want_mask else None
        wb = None
        if wb_rgb is not None:
            wb = np.ascontiguousarray(wb_rgb, np.float32)
            if wb.size != 3:
                raise LibRawInvalidArgument("wb_rgb needs three values")
        mv = np.zeros(1, np.float32)
        _check(native().lre_hip_recover_highlights(self._h, _p(src), _p(out), w, h, saturation_threshold,
                                                   None if mask is None else _p(mask),
                                                   None if wb is None else _p(wb), _p(mv)))
        return out, mask, float(mv[0])

    def defringe(self, rgb_input, rgb_output=None, radius: float = 10.0,
                 edge_threshold: float = 0.1, chroma_threshold: float = 0.15,
                 strength: float = 10.0, defringe_green: bool = False,
                 green_strength_scale: float = 0.3):
        """Accelerator::defringe.  Passing the same array as rgb_input and
        rgb_output (or rgb_output=None with inplace semantics requested by
        passing rgb_output=rgb_input) runs the reference's in-place form."""
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        inplace = rgb_output is not None and rgb_output is rgb_input
        out = src if inplace else np.empty_like(src)
        _check(native().lre_hip_defringe(self._h, _p(src), _p(out), w, h, radius, edge_threshold,
                                         chroma_threshold, strength, int(defringe_green),
                                         green_strength_scale))
        return out

    def convert_color_space(self, rgb_input, transform):
        """transform: 3x4 (row-major) camera-to-output matrix (core/accelerator.h:310-313)."""
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        t = np.ascontiguousarray(np.asarray(transform, np.float32).reshape(12))
        out = np.empty_like(src)
        _check(native().lre_hip_convert_color_space(self._h, _p(src), _p(out), w, h, _p(t)))
        return out

    def gamma_correct(self, rgb_input, gamma_power: float = 0.0, gamma_slope: float = 0.0,
                      output_color_space: int = 1):
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        out = np.empty_like(src)
        _check(native().lre_hip_gamma_correct(self._h, _p(src), _p(out), w, h, gamma_power,
                                              gamma_slope, output_color_space))
        return out

    def quantize(self, rgb, bits: int = 8):
        """The 8/16-bit output array of the processed image (core/python_bindings.cpp:235-257)."""
        src = np.ascontiguousarray(rgb, np.float32)
        out = np.empty(src.shape, np.uint8 if bits == 8 else np.uint16)
        _check(native().lre_hip_quantize(self._h, _p(src), out.ctypes.data_as(_vp), src.size, bits))
        return out

    def ca_axial_cleanup(self, rgb_input, radius: int = 6, epsilon: float = 1e-4,
                         strength: float = 0.3):
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        out = np.empty_like(src)
        _check(native().lre_hip_axial_ca(self._h, _p(src), _p(out), w, h, radius, epsilon, strength))
        return out

    def ca_register_lateral(self, rgb_input, cell_size: int = 96, max_iterations: int = 3,
                            max_shift: float = 6.0, min_confidence: float = 0.02,
                            pyramid_levels: int = 3):
        src = _rgb(rgb_input)
        h, w, _ = src.shape
        out = np.empty_like(src)
        _check(native().lre_hip_lateral_ca(self._h, _p(src), _p(out), w, h, cell_size,
                                           max_iterations, max_shift, min_confidence,
                                           pyramid_levels))
        return out

    def apply_black_level_correction(self, img4, cblack=(0, 0, 0, 0), black: int = 0,
                                     maximum: int = 16383, pos_table=None):
        """In place on a (H, W, 4) uint16 image; returns it."""
        if img4.dtype != np.uint16 or img4.ndim != 3 or img4.shape[2] != 4 or not img4.flags.c_contiguous:
            raise LibRawInvalidArgument("raw_buffer must be a C-contiguous (H, W, 4) uint16 array")
        h, w, _ = img4.shape
        cb = (ctypes.c_int * 4)(*[int(v) for v in cblack])
        if pos_table is not None:
            pt = np.ascontiguousarray(pos_table, np.int32)
            pr, pc = pt.shape
            ptr = pt.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        else:
            pr = pc = 0
            ptr = None
        _check(native().lre_hip_black_level(self._h, _p(img4, _u16p), w, h, cb, ptr, pr, pc, black, maximum))
        return img4

    def apply_green_matching(self, img4, filters: int, maximum: int):
        """In place on a (H, W, 4) uint16 image (core/libraw_wrapper.cpp:496-585); returns it."""
        img4 = self._img4(img4)
        h, w, _ = img4.shape
        _check(native().lre_hip_green_matching(self._h, _p(img4, _u16p), w, h, filters, maximum))
        return img4

    def remove_zeroes(self, img4, filters: int):
        """In place on a (H, W, 4) uint16 image (core/libraw_wrapper.cpp:150-212); returns it."""
        img4 = self._img4(img4)
        h, w, _ = img4.shape
        _check(native().lre_hip_remove_zeroes(self._h, _p(img4, _u16p), w, h, filters))
        return img4

    def auto_wb(self, img4, filters: int, xtrans=None):
        """Automatic pre_mul of scale_colors (core/libraw_wrapper.cpp:621-663)."""
        img4 = self._img4(img4)
        h, w, _ = img4.shape
        xt = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(36), np.int8)
        pm = np.zeros(4, np.float32)
        _check(native().lre_hip_auto_wb(self._h, _p(img4, _u16p), w, h, filters,
                                        None if xt is None else _p(xt, _i8p), _p(pm)))
        return pm

    def scale_aber(self, img4, aber):
        """Legacy radial aberration resampling of scale_colors (:833-880), in place; returns img4."""
        img4 = self._img4(img4)
        h, w, _ = img4.shape
        ab = np.ascontiguousarray(aber, np.float32)
        _check(native().lre_hip_scale_aber(self._h, _p(img4, _u16p), w, h, _p(ab)))
        return img4

    @staticmethod
    def _img4(img4):
        if img4.dtype != np.uint16 or img4.ndim != 3 or img4.shape[2] != 4 or not img4.flags.c_contiguous:
            raise LibRawInvalidArgument("image must be a C-contiguous (H, W, 4) uint16 array")
        return img4

    def apply_white_balance(self, img4, wb, filters: int = RGGB, xtrans=None, rgb=None):
        img4 = np.ascontiguousarray(img4, np.uint16)
        if img4.ndim != 3 or img4.shape[2] != 4:
            raise LibRawInvalidArgument("raw_buffer must be (H, W, 4) uint16")
        h, w, _ = img4.shape
        out = np.zeros((h, w, 3), np.float32) if rgb is None else _rgb(rgb).copy()
        wbv = np.ascontiguousarray(wb, np.float32)
        xt = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(-1), np.int8)
        _check(native().lre_hip_white_balance(self._h, _p(img4, _u16p), _p(out), w, h, _p(wbv),
                                              filters, None if xt is None else _p(xt, _i8p)))
        return out


class DeviceBuffer:
    """A device allocation owned by a HIPAccelerator context."""

    def __init__(self, acc: HIPAccelerator, nbytes: int):
        self._acc = acc
        self.nbytes = int(nbytes)
        p = _vp()
        _check(native().lre_hip_dev_malloc(acc.handle, ctypes.byref(p), self.nbytes))
        self.ptr = p

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(native().lre_hip_memcpy_htod(self._acc.handle, self.ptr,
                                            arr.ctypes.data_as(_vp), arr.nbytes))

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(native().lre_hip_memcpy_dtoh(self._acc.handle, out.ctypes.data_as(_vp), self.ptr,
                                            out.nbytes))
        return out

    def free(self):
        if self.ptr:
            native().lre_hip_dev_free(self._acc.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class LibRawWrapper:
    """The standalone float32 methods of the reference's LibRawWrapper
    (core/python_bindings.cpp:324-351): input (H, W, 3) float32, new array out."""

    def __init__(self, device: int = 0):
        self._acc = HIPAccelerator(device)

    def tone_mapping(self, image, after_scale: float = 1.0):
        # core/libraw_wrapper.cpp:3098-3131
        return self._acc.tone_mapping(_rgb(image), after_scale)

    def enhance_micro_contrast(self, image, threshold: float = -1.0, strength: float = 8.0,
                               target_contrast: float = 0.06):
        # core/libraw_wrapper.cpp:3134-3173: threshold < 0 takes maximum /
        # data_maximum of the loaded raw, which falls back to 1 when there is
        # none (:3004-3011); no raw is ever loaded here
        if threshold < 0.0:
            threshold = 1.0
        return self._acc.enhance_micro_contrast(_rgb(image), threshold, strength, target_contrast)

    def recover_highlights(self, image, threshold: float = -1.0):
        # core/libraw_wrapper.cpp:3063-3095: a new array; threshold < 0 takes
        # maximum / data_maximum of the loaded raw, 1 when there is none
        # (:3004-3011); no raw is ever loaded here.  No mask, no white balance.
        if threshold < 0.0:
            threshold = 1.0
        return self._acc.recover_highlights(_rgb(image), threshold)[0]

    def defringe(self, image, radius: float = 10.0, strength: float = 10.0,
                 defringe_green: bool = False):
        # core/libraw_wrapper.cpp:3176-3214: copies the input and runs with
        # distinct buffers, edge 0.1, chroma 0.15, green scale 0.3
        return self._acc.defringe(_rgb(image), None, radius, 0.1, 0.15, strength,
                                  defringe_green, 0.3)


_WB_SCALARS = ("use_auto_wb", "use_camera_wb", "camerawb_daylight_fallback", "as_shot_wb_applied", "colors",
               "filters", "black", "maximum", "warnings")


def wb_meta(**fields) -> WbMeta:
    """An lre_wb_meta from keyword fields (arrays as sequences); pre_mul
    defaults to 1, colors to 3, everything else to 0."""
    m = WbMeta()
    m.colors = 3
    for i in range(4):
        m.pre_mul[i] = 1.0
    for k, v in fields.items():
        if k in ("user_mul", "cam_mul", "pre_mul", "cblack"):
            arr = getattr(m, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k == "white":
            w = np.asarray(v, np.uint16).reshape(8, 8)
            for r in range(8):
                for c in range(8):
                    m.white[r][c] = int(w[r, c])
        elif k in _WB_SCALARS:
            setattr(m, k, v)
        else:
            raise LibRawInvalidArgument(f"unknown white-balance field {k!r}")
    return m


def wb_needs_auto(meta: WbMeta) -> bool:
    """True when scale_colors takes the image-wide multipliers of auto_wb
    (core/libraw_wrapper.cpp:621-630)."""
    return bool(native().lre_wb_needs_auto(ctypes.byref(meta)))


def scale_colors_multipliers(meta: WbMeta, auto_pre_mul=None) -> np.ndarray:
    """scale_colors' multiplier logic (core/libraw_wrapper.cpp:612-862): updates
    meta's pre_mul, maximum, cblack and warnings in place and returns the four
    scale_mul values.  auto_pre_mul (HIPAccelerator.auto_wb's result) is
    required when wb_needs_auto(meta)."""
    out = np.zeros(4, np.float32)
    a = None if auto_pre_mul is None else np.ascontiguousarray(auto_pre_mul, np.float32)
    if a is not None and a.size != 4:
        raise LibRawInvalidArgument("auto_pre_mul needs four values")
    _check(native().lre_scale_colors_meta(ctypes.byref(meta), None if a is None else _p(a), _p(out)))
    return out
