#!/usr/bin/env python3
"""Throughput of the develop chain on one MI355X.

Workload: one 24 MP (4000 x 6000) RGGB Bayer image, float32, from the seeded
scene generator (gradient + 200 polygons + zone plate, Poisson-Gaussian
noise), developed by AMaZE demosaic -> defringe (radius 10, strength 10, green
off) -> detail tone map.  The CFA plane is resident in HBM before timing; a
step is one pass of the chain over the image.

Prints one JSON line: MP/s (value), ms per step, the roofline of the dominant
kernel (HIP events on the stream the stages run on), and the reference's own
CPU code timed on this host's cores on the same image.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per pixel of each stage (fp32 operator I/O)
STAGE_BYTES_PER_PX = {"amaze": 4 + 12, "defringe": 12 + 12, "detail_tm": 12 + 12,
                      "tone_mapping": 12 + 12}


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                                 ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, ev, stream):
        assert self.lib.hipEventRecord(ev, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.lib.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)


def cpu_baseline(raw3, stages, threads):
    """The reference's CPU code (oracle/_ref, built from its sources) on the same
    image and stage list; falls back to the CPU restatement if absent."""
    f32p = ctypes.POINTER(ctypes.c_float)
    h, w, _ = raw3.shape
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so")
    kind = "reference" if os.path.exists(ref_path) else "port"
    lib = ctypes.CDLL(ref_path if kind == "reference" else os.path.join(ROOT, "oracle", "liblre_oracle.so"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    if kind == "reference":
        lib.ref_set_threads(threads)
    rgb = np.zeros_like(raw3)
    tmp = np.zeros_like(raw3)
    cm = np.ones(4, np.float32)
    t0 = time.perf_counter()
    for s in stages:
        if s == "amaze":
            if kind == "reference":
                lib.ref_amaze(raw3.ctypes.data_as(f32p), rgb.ctypes.data_as(f32p), w, h,
                              ctypes.c_uint32(0x94949494), cm.ctypes.data_as(f32p), ctypes.c_float(1.0))
            else:
                lib.oracle_amaze(raw3.ctypes.data_as(f32p), rgb.ctypes.data_as(f32p), w, h,
                                 ctypes.c_uint32(0x94949494), cm.ctypes.data_as(f32p),
                                 ctypes.c_float(1.0), 0)
        elif s == "defringe":
            fn = lib.ref_defringe if kind == "reference" else lib.oracle_defringe
            fn(rgb.ctypes.data_as(f32p), tmp.ctypes.data_as(f32p), w, h, ctypes.c_float(10.0),
               ctypes.c_float(0.1), ctypes.c_float(0.15), ctypes.c_float(10.0), 0, ctypes.c_float(0.3))
            rgb, tmp = tmp, rgb
        elif s == "detail_tm":
            fn = lib.ref_detail_tonemap if kind == "reference" else lib.oracle_detail_tonemap
            fn(rgb.ctypes.data_as(f32p), w, h)
    dt = time.perf_counter() - t0
    return {"value": round(h * w / 1e6 / dt, 3), "unit": "MP/s", "cores": threads, "kind": kind,
            "sample": f"one {w}x{h} RGGB image through {' -> '.join(stages)} ({dt:.2f} s)"}


def load_traffic(stage):
    path = os.path.join(ROOT, "profiles", f"traffic_{stage}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--height", type=int, default=4000)
    ap.add_argument("--width", type=int, default=6000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): images are independent, "
              "so N GPUs would each run this same single-GPU workload", file=sys.stderr)
        sys.exit(2)

    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    from libraw_enhanced_amd.pipeline import DevelopChain, available_chain

    h, w = args.height, args.width
    seed = scene.BASE_SEED
    print(f"bench: generating {w}x{h} RGGB scene (seed {seed})", file=sys.stderr, flush=True)
    cfa = scene.bayer_scene(h, w, seed)
    acc = HIPAccelerator(0)
    d_cfa = acc.dev_from_host(cfa)
    stages = available_chain()
    chain = DevelopChain(acc, h, w, stages)
    stream = native().lre_hip_stream(acc.handle)
    ev = HipEvents()

    for _ in range(args.warmup):
        chain.run(d_cfa)
    acc.synchronize()

    marks = [[ev.create() for _ in range(len(stages) + 1)] for _ in range(args.steps)]
    print(f"bench: timing {args.steps} steps of {' -> '.join(stages)}", file=sys.stderr, flush=True)
    acc.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev.record(marks[k][0], stream)
        for i, s in enumerate(stages):
            chain.run_stage(s, d_cfa)
            ev.record(marks[k][i + 1], stream)
    acc.synchronize()
    elapsed = time.perf_counter() - t0

    stage_ms = {s: float(np.mean([ev.elapsed_ms(marks[k][i], marks[k][i + 1])
                                  for k in range(args.steps)])) for i, s in enumerate(stages)}
    mp = h * w / 1e6
    ms_per_step = elapsed * 1e3 / args.steps
    value = mp * args.steps / elapsed
    dominant = max(stage_ms, key=stage_ms.get)
    alg_bytes = STAGE_BYTES_PER_PX[dominant] * h * w
    achieved = alg_bytes / (stage_ms[dominant] * 1e-3) / 1e9
    line = {
        "metric": "MP/s full develop chain (24 MP Bayer)",
        "value": round(value, 2),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene: gradient + 200 polygons + zone plate, Poisson-Gaussian noise)",
        "config": {"workload": f"{' -> '.join(stages)} on one {w}x{h} RGGB float32 mosaic",
                   "image_mp": round(mp, 2), "stages": stages,
                   "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()}},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(dominant),
                     "alg_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        print(f"bench: CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        raw3 = scene.cfa_to_rgb3(cfa)
        line["cpu_baseline"] = cpu_baseline(raw3, stages, threads)
    else:
        line["cpu_baseline"] = None
    acc.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
