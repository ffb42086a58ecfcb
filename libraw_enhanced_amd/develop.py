"""Many images on one GPU: the batched develop and the request loop, through
the C ABI (include/lre_hip.h, lre_hip_batch_* and lre_hip_serve_*).

The stage functions and their parameters are the reference's Accelerator
methods (core/accelerator.h:279-367); the order of a stage list is the
caller's.  The develop order of the reference's pipeline after demosaic is
lateral CA -> axial CA -> defringe (core/libraw_wrapper.cpp:1776-1818); the
chains here put the tone map last.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from .accelerator import FILTERS_XTRANS, RGGB, HIPAccelerator, _check, native

_vp = ctypes.c_void_p
_f = ctypes.c_float
_i = ctypes.c_int

STAGE_DEMOSAIC = 1
STAGE_LATERAL_CA = 2
STAGE_AXIAL_CA = 3
STAGE_DEFRINGE = 4
STAGE_DETAIL_TM = 5
STAGE_TONE_MAPPING = 6
STAGE_COLOR_SPACE = 7
STAGE_GAMMA = 8
STAGE_HIGHLIGHT = 9
STAGE_MICRO_CONTRAST = 10
STAGE_REMOVE_ZEROES = 11
STAGE_GREEN_MATCHING = 12

STAGE_NAMES = {
    "demosaic": STAGE_DEMOSAIC, "lateral_ca": STAGE_LATERAL_CA, "axial_ca": STAGE_AXIAL_CA,
    "defringe": STAGE_DEFRINGE, "detail_tm": STAGE_DETAIL_TM, "tone_mapping": STAGE_TONE_MAPPING,
    "color_space": STAGE_COLOR_SPACE, "gamma": STAGE_GAMMA,
    "highlight": STAGE_HIGHLIGHT, "micro_contrast": STAGE_MICRO_CONTRAST,
    "remove_zeroes": STAGE_REMOVE_ZEROES, "green_matching": STAGE_GREEN_MATCHING,
}

# demosaic -> axial CA -> defringe -> detail tone map
BATCH_CHAIN = ("demosaic", "axial_ca", "defringe", "detail_tm")


class ChainParams(ctypes.Structure):
    _fields_ = [
        ("defringe_radius", _f), ("defringe_edge_threshold", _f),
        ("defringe_chroma_threshold", _f), ("defringe_strength", _f),
        ("defringe_green", _i), ("defringe_green_scale", _f),
        ("axial_radius", _i), ("axial_epsilon", _f), ("axial_strength", _f),
        ("lateral_cell", _i), ("lateral_iterations", _i), ("lateral_max_shift", _f),
        ("lateral_min_confidence", _f), ("lateral_levels", _i),
        ("tone_after_scale", _f),
        ("color_transform", _f * 12), ("gamma_power", _f), ("gamma_slope", _f),
        ("output_color_space", _i),
        ("highlight_threshold", _f), ("micro_strength", _f), ("micro_target_contrast", _f),
    ]

    @classmethod
    def defaults(cls, **overrides):
        p = cls()
        native().lre_chain_params_default(ctypes.byref(p))
        for k, v in overrides.items():
            setattr(p, k, v)
        return p


class Image(ctypes.Structure):
    _fields_ = [
        ("width", _i), ("height", _i), ("filters", ctypes.c_uint32),
        ("xtrans", ctypes.c_int8 * 36), ("cam_mul", _f * 4),
        ("d_cfa", _vp), ("maximum_value", _f),
        ("d_raw", _vp), ("black", _i), ("white", _i), ("cblack", _i * 4), ("wb", _f * 4),
        ("adjust_threshold", _f),
        ("d_rgb", _vp),
    ]


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("image", Image), ("h_raw", _vp), ("h_rgb", _vp)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", _i), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]


_SIGS = {
    "lre_chain_params_default": ([ctypes.POINTER(ChainParams)], None),
    "lre_hip_batch_create": ([ctypes.POINTER(_vp), _i, _i], _i),
    "lre_hip_batch_destroy": ([_vp], _i),
    "lre_hip_batch_run": ([_vp, ctypes.POINTER(Image), _i, ctypes.POINTER(_i), _i,
                           ctypes.POINTER(ChainParams)], _i),
    "lre_hip_batch_synchronize": ([_vp], _i),
    "lre_hip_batch_capture": ([_vp, ctypes.POINTER(Image), _i, ctypes.POINTER(_i), _i,
                               ctypes.POINTER(ChainParams)], _i),
    "lre_hip_batch_replay": ([_vp], _i),
    "lre_hip_batch_set_config": ([_vp, _i, _i, _i], _i),
    "lre_hip_batch_get_config": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)], _i),
    "lre_hip_batch_set_grouping": ([_vp, _i, ctypes.c_size_t], _i),
    "lre_hip_batch_get_grouping": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_size_t)], _i),
    "lre_hip_batch_set_tiles": ([_vp, _i], _i),
    "lre_hip_batch_get_tiles": ([_vp, ctypes.POINTER(_i)], _i),
    "lre_hip_batch_autotune": ([_vp, ctypes.POINTER(Image), _i, ctypes.POINTER(_i), _i,
                                ctypes.POINTER(ChainParams), _i, ctypes.POINTER(_f)], _i),
    "lre_hip_serve_create": ([ctypes.POINTER(_vp), _i, _i, ctypes.c_size_t, ctypes.POINTER(_i), _i,
                              ctypes.POINTER(ChainParams)], _i),
    "lre_hip_serve_submit": ([_vp, ctypes.POINTER(Request)], _i),
    "lre_hip_serve_poll": ([_vp, ctypes.POINTER(Completion), _i, _i], _i),
    "lre_hip_serve_now": ([_vp], ctypes.c_double),
    "lre_hip_serve_footprint": ([_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)], _i),
    "lre_hip_serve_destroy": ([_vp], _i),
    "lre_hip_host_alloc": ([ctypes.c_size_t], _vp),
    "lre_hip_host_free": ([_vp], _i),
}


def lib():
    L = native()
    if not getattr(L, "_lre_develop_sigs", False):
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        L._lre_develop_sigs = True
    return L


def stage_codes(stages) -> "ctypes.Array":
    codes = [STAGE_NAMES[s] if isinstance(s, str) else int(s) for s in stages]
    return (_i * max(len(codes), 1))(*codes)


def make_image(width, height, *, filters=RGGB, xtrans=None, cam_mul=(1, 1, 1, 1), d_cfa=None,
               maximum_value=1.0, d_raw=None, black=0, white=0, cblack=(0, 0, 0, 0),
               wb=(1, 1, 1, 1), adjust_threshold=0.75, d_rgb=None) -> Image:
    im = Image()
    im.width, im.height, im.filters = int(width), int(height), int(filters)
    if xtrans is not None:
        flat = np.asarray(xtrans, np.int8).reshape(36)
        for k in range(36):
            im.xtrans[k] = int(flat[k])
    for k in range(4):
        im.cam_mul[k] = float(cam_mul[k])
        im.cblack[k] = int(cblack[k])
        im.wb[k] = float(wb[k])
    im.d_cfa = _ptr(d_cfa)
    im.maximum_value = float(maximum_value)
    im.d_raw = _ptr(d_raw)
    im.black, im.white = int(black), int(white)
    im.adjust_threshold = float(adjust_threshold)
    im.d_rgb = _ptr(d_rgb)
    return im


def _ptr(buf):
    if buf is None:
        return None
    if hasattr(buf, "ptr"):
        return buf.ptr.value if isinstance(buf.ptr, ctypes.c_void_p) else buf.ptr
    return int(buf)


class BatchDeveloper:
    """Develop many resident images in one pass per stage (lre_hip_batch_*)."""

    def __init__(self, device: int = 0, lanes: int = 2):
        h = _vp()
        _check(lib().lre_hip_batch_create(ctypes.byref(h), device, lanes))
        self._h = h

    def _args(self, images, stages, params):
        arr = (Image * max(len(images), 1))(*images)
        st = stage_codes(stages)
        p = params if params is not None else ChainParams.defaults()
        return arr, len(images), st, len(stages), p

    def run(self, images, stages=BATCH_CHAIN, params=None):
        arr, n, st, ns, p = self._args(images, stages, params)
        _check(lib().lre_hip_batch_run(self._h, arr, n, st, ns, ctypes.byref(p)))

    def synchronize(self):
        _check(lib().lre_hip_batch_synchronize(self._h))

    def capture(self, images, stages=BATCH_CHAIN, params=None):
        arr, n, st, ns, p = self._args(images, stages, params)
        _check(lib().lre_hip_batch_capture(self._h, arr, n, st, ns, ctypes.byref(p)))

    def replay(self):
        _check(lib().lre_hip_batch_replay(self._h))

    def set_config(self, lanes, amaze_slots_per_cu=2, xtrans_slots_per_cu=4):
        _check(lib().lre_hip_batch_set_config(self._h, lanes, amaze_slots_per_cu, xtrans_slots_per_cu))

    def config(self):
        a, b, c = _i(), _i(), _i()
        _check(lib().lre_hip_batch_get_config(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"lanes": a.value, "amaze_slots_per_cu": b.value, "xtrans_slots_per_cu": c.value}

    GROUP_MODES = {False: 0, True: 1, "auto": 2}

    def set_grouping(self, group_mode="auto", post_budget_bytes=0):
        """Group mode (True, False or "auto": on when the mean image is below
        2 MP): each post-demosaic stage as one launch over a group of images
        whose scratch fits post_budget_bytes (0: 8 GiB, at most a quarter of the device)."""
        mode = self.GROUP_MODES[group_mode if isinstance(group_mode, str) else bool(group_mode)]
        _check(lib().lre_hip_batch_set_grouping(self._h, mode, int(post_budget_bytes)))

    def grouping(self):
        g, b = _i(), ctypes.c_size_t()
        _check(lib().lre_hip_batch_get_grouping(self._h, ctypes.byref(g), ctypes.byref(b)))
        mode = {0: False, 1: True, 2: "auto"}[g.value]
        return {"group_mode": mode, "post_budget_bytes": b.value}

    def set_tiles(self, defringe_blur_rows=0):
        """Defringe blur tile height: 48, 96 or 0 (by image size); same bits."""
        _check(lib().lre_hip_batch_set_tiles(self._h, int(defringe_blur_rows)))

    def tiles(self):
        r = _i()
        _check(lib().lre_hip_batch_get_tiles(self._h, ctypes.byref(r)))
        return {"defringe_blur_rows": r.value}

    def autotune(self, images, stages=BATCH_CHAIN, params=None, reps=1):
        arr, n, st, ns, p = self._args(images, stages, params)
        ms = _f()
        _check(lib().lre_hip_batch_autotune(self._h, arr, n, st, ns, ctypes.byref(p), reps,
                                            ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            lib().lre_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A numpy view of pinned host memory (lre_hip_host_alloc)."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().lre_hip_host_alloc(nbytes)
        if not self.ptr:
            raise MemoryError(f"pinned allocation of {nbytes} bytes failed")
        buf = (ctypes.c_char * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(buf, self.dtype).reshape(self.shape)

    def free(self):
        if self.ptr:
            self.array = None
            lib().lre_hip_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DevelopServer:
    """The request loop (lre_hip_serve_*): FIFO admission by device footprint,
    `lanes` requests in flight on their own streams."""

    def __init__(self, device=0, lanes=3, hbm_budget_bytes=64 << 30, stages=BATCH_CHAIN,
                 params=None):
        h = _vp()
        st = stage_codes(stages)
        p = params if params is not None else ChainParams.defaults()
        _check(lib().lre_hip_serve_create(ctypes.byref(h), device, lanes, int(hbm_budget_bytes), st,
                                          len(stages), ctypes.byref(p)))
        self._h = h
        self._keep = {}

    def now(self) -> float:
        return lib().lre_hip_serve_now(self._h)

    def submit(self, req_id: int, image: Image, h_raw_ptr: int, h_rgb_ptr: int, keep=None):
        r = Request()
        r.id = int(req_id)
        r.image = image
        r.h_raw = h_raw_ptr
        r.h_rgb = h_rgb_ptr
        self._keep[int(req_id)] = keep
        _check(lib().lre_hip_serve_submit(self._h, ctypes.byref(r)))

    def poll(self, max_n=64, timeout_ms=100):
        out = (Completion * max_n)()
        k = lib().lre_hip_serve_poll(self._h, out, max_n, timeout_ms)
        if k < 0:
            _check(k)
        res = []
        for c in out[:k]:
            self._keep.pop(int(c.id), None)
            res.append({"id": int(c.id), "status": int(c.status), "t_submit": c.t_submit,
                        "t_admit": c.t_admit, "t_done": c.t_done})
        return res

    def footprint(self):
        cur, peak = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().lre_hip_serve_footprint(self._h, ctypes.byref(cur), ctypes.byref(peak)))
        return cur.value, peak.value

    def close(self):
        if getattr(self, "_h", None):
            lib().lre_hip_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServeReport:
    rate: float
    requests: int
    duration_s: float
    mp_total: float
    latencies_s: list = field(default_factory=list)

    def summary(self):
        lat = np.sort(np.asarray(self.latencies_s))
        pct = (lambda q: float(lat[min(len(lat) - 1, int(math.ceil(q * len(lat))) - 1)]) if len(lat) else None)
        return {"rate_per_s": self.rate, "requests": self.requests,
                "wall_s": round(self.duration_s, 3), "mp_per_s": round(self.mp_total / self.duration_s, 2),
                "p50_ms": round(pct(0.50) * 1e3, 2) if len(lat) else None,
                "p99_ms": round(pct(0.99) * 1e3, 2) if len(lat) else None,
                "max_ms": round(float(lat[-1]) * 1e3, 2) if len(lat) else None}


def poisson_schedule(rate: float, n: int, seed: int) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process of the given rate."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, n))


def run_open_loop(server: DevelopServer, arrivals, make_request, on_done=None, poll_ms=2):
    """Submit request k at arrivals[k] seconds after start (open loop: arrivals do
    not wait for completions) and collect every completion.  make_request(k)
    returns (Image, h_raw_ptr, h_rgb_ptr, megapixels, keepalive)."""
    n = len(arrivals)
    t0 = time.perf_counter()
    lat, mp_total, done, k = [], 0.0, 0, 0
    mp = {}
    while done < n:
        now = time.perf_counter() - t0
        while k < n and arrivals[k] <= now:
            im, raw_ptr, rgb_ptr, mpx, keep = make_request(k)
            mp[k] = mpx
            server.submit(k, im, raw_ptr, rgb_ptr, keep)
            k += 1
        wait = poll_ms if k >= n else max(0, min(poll_ms, int((arrivals[k] - now) * 1e3)))
        for c in server.poll(timeout_ms=wait):
            if c["status"] != 0:
                raise RuntimeError(f"request {c['id']} failed with status {c['status']}")
            lat.append(c["t_done"] - c["t_submit"])
            mp_total += mp[c["id"]]
            done += 1
            if on_done:
                on_done(c)
    return time.perf_counter() - t0, lat, mp_total
