#!/usr/bin/env python3
"""Throughput of the develop chain on one MI355X.

Default workload (--workload chain24): one 24 MP (4000 x 6000) RGGB Bayer
image, float32, from the seeded scene generator (gradient + 200 polygons +
zone plate, Poisson-Gaussian noise), developed by AMaZE demosaic -> defringe
(radius 10, strength 10, green off) -> detail tone map.  The CFA plane is
resident in HBM before timing; a step is one pass of the chain over the image.
Prints one JSON line: MP/s (value), ms per step, the rooflines of the dominant
kernel (HBM and VALU issue; HIP events on the stream the stages run on), the
reference's own CPU code timed on this host's cores on the same image, parity
of the 24 MP output against that CPU output and against an FP64 evaluation,
and the metric's second half: p50/p99 latency of the request loop at 4, 8 and
16 requests/s (--serve-duration seconds each) beside the plain per-image path.

Other workloads, one JSON line each:
  xtrans26  one 26 MP (4160 x 6240) X-Trans image, pre-warped with radial
            lateral CA: X-Trans 3-pass -> lateral CA -> axial CA
  batch64   64 images (16x 12 MP, 16x 24 MP, 12x 45 MP, 8x 61 MP Bayer, 12x
            40 MP X-Trans), float32: demosaic -> axial CA -> defringe -> detail
            tone map as one batch (multi-image tile queues + HIP graph), against
            the plain per-image path (host arrays in and out of every stage)
  serve     Poisson arrivals of uint16 requests (40% 24 MP, 30% 45 MP, 20%
            61 MP Bayer, 10% 40 MP X-Trans; black 512, white 16383, WB
            2.1/1.0/1.5) through the request loop; p50/p99 latency and MP/s
  grouping  the batch's two post-stage modes (one launch per stage over a
            group of images; each image's chain on one of several streams) on
            batches of 256 x 0.3 MP, 64 x 3.1 MP and 16 x 12 MP Bayer images,
            eager and replayed from a graph, 1-3 lanes
  micro24   micro-contrast with a highlight mask on one resident 24 MP RGB
            image (the stage the pipeline runs for highlight_mode >= 4)
  highlight24  highlight recovery with its mask and white balance on one
            resident 24 MP RGB image (highlight_mode >= 4), magenta-cast
            highlights so that the LoCA pass has work
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
STAGE_BYTES_PER_PX = {"amaze": 4 + 12, "xtrans3": 4 + 12, "defringe": 12 + 12, "detail_tm": 12 + 12,
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


def host_cpu_info():
    """Model name of the host CPU and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus_in_affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads(args):
    """Threads of the CPU leg: --cpu-threads, else OMP_NUM_THREADS (the CPU share
    the host allots this job), else every CPU in the affinity mask."""
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(raw3, stages, threads, runs=5):
    """The reference's CPU code (oracle/_ref, built from its sources) on the same
    image and stage list, called as the develop pipeline calls it (defringe and
    the tone map in place); the CPU restatement if the build is absent.  One
    warm-up pass, then the median of `runs` timed passes.  Returns the
    cpu_baseline object and the output of the last pass."""
    f32p = ctypes.POINTER(ctypes.c_float)
    h, w, _ = raw3.shape
    host = host_cpu_info()
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so")
    kind = "reference" if os.path.exists(ref_path) else "port"
    lib = ctypes.CDLL(ref_path if kind == "reference" else os.path.join(ROOT, "oracle", "liblre_oracle.so"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    if kind == "reference":
        lib.ref_set_threads(threads)
    rgb = np.zeros_like(raw3)
    cm = np.ones(4, np.float32)
    P = lambda a: a.ctypes.data_as(f32p)

    def one_pass():
        for s in stages:
            if s == "amaze":
                if kind == "reference":
                    lib.ref_amaze(P(raw3), P(rgb), w, h, ctypes.c_uint32(0x94949494), P(cm), ctypes.c_float(1.0))
                else:
                    lib.oracle_amaze(P(raw3), P(rgb), w, h, ctypes.c_uint32(0x94949494), P(cm),
                                     ctypes.c_float(1.0), 0)
            elif s == "defringe":
                # in place, as postprocess calls it (core/libraw_wrapper.cpp:1810)
                fn = lib.ref_defringe if kind == "reference" else lib.oracle_defringe
                fn(P(rgb), P(rgb), w, h, ctypes.c_float(10.0), ctypes.c_float(0.1), ctypes.c_float(0.15),
                   ctypes.c_float(10.0), 0, ctypes.c_float(0.3))
            elif s == "detail_tm":
                fn = lib.ref_detail_tonemap if kind == "reference" else lib.oracle_detail_tonemap
                fn(P(rgb), w, h)

    one_pass()  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        one_pass()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    line = {"value": round(h * w / 1e6 / dt, 3), "unit": "MP/s", "cores": threads, "kind": kind,
            "sample": f"one {w}x{h} RGGB image through {' -> '.join(stages)}: 1 warm-up pass, median of "
                      f"{runs} passes ({dt:.2f} s each, {threads} OpenMP threads)"}
    line.update(host)
    return line, rgb


def fp64_post_chain(rgb, stages, threads):
    """An FP64 evaluation of the post-demosaic stages (the restatement compiled
    with double in place of float, oracle/fp64.h) from a demosaic output; the
    checker for the parity block, not a timed leg."""
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liblre_oracle.so"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    f64p = ctypes.POINTER(ctypes.c_double)
    D = ctypes.c_double
    a = np.ascontiguousarray(rgb, np.float64).copy()
    h, w, _ = a.shape
    P = a.ctypes.data_as(f64p)
    for st in stages:
        if st == "defringe":
            assert lib.oracle64_defringe(P, P, w, h, D(10.0), D(0.1), D(0.15), D(10.0), 0, D(0.3)) == 0
        elif st == "detail_tm":
            assert lib.oracle64_detail_tonemap(P, w, h) == 0
    return a


def edge_band_mask(h, w, band):
    """True within `band` px of the frame."""
    m = np.zeros((h, w), bool)
    m[:band, :] = m[-band:, :] = True
    m[:, :band] = m[:, -band:] = True
    return m


def parity_block(got, want, band):
    """max-abs, PSNR and differing pixels of got against want, over the whole
    image and split by the band of edge tiles where the reference's output
    depends on its thread schedule."""
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    px_diff = np.any(got != want, axis=2)
    edge = edge_band_mask(*got.shape[:2], band)
    out = {}
    for name, sel in (("all", np.ones_like(edge)), ("inside_edge_band", edge), ("outside_edge_band", ~edge)):
        dd = d[sel]
        mse = float(np.mean(dd * dd))
        out[name] = {"max_abs": float(dd.max()), "psnr_db": (None if mse == 0 else round(10 * np.log10(1.0 / mse), 2)),
                     "px_differ": int(np.count_nonzero(px_diff[sel])), "px": int(sel.sum())}
    return out


def valu_roofline(kernel, stage, ms, px):
    """The VALU issue roofline of a demosaic kernel: its VALU instructions per
    launch (rocprofv3 SQ_INSTS_VALU on the same workload, profiles/
    valu_demosaic.json) over the launch time measured here, against one wave64
    instruction per CU per clock; plus the stage's FP32 operations per pixel
    (the oracle's op count, profiles/opcount_demosaic.json) as a rate."""
    try:
        with open(os.path.join(ROOT, "profiles", "valu_demosaic.json")) as f:
            v = json.load(f)
        with open(os.path.join(ROOT, "profiles", "opcount_demosaic.json")) as f:
            o = json.load(f)["stages"][stage]
    except (OSError, KeyError, ValueError):
        return None
    insts = v["kernels"][kernel]["SQ_INSTS_VALU"]
    peak = v["peak_wave_insts_per_s"]
    rate = insts / (ms * 1e-3)
    ops = o["fp32_arith_per_px"] * px
    return {"kernel": kernel, "bound": "valu", "achieved": round(rate / 1e9, 1), "peak": round(peak / 1e9, 1),
            "unit": "G wave64 VALU instructions/s", "frac": round(rate / peak, 4),
            "insts_per_launch": int(insts), "lane_insts_per_px": round(insts * 64 / px, 1),
            "fp32_ops_per_px": o["fp32_arith_per_px"], "fp32_tops": round(ops / (ms * 1e-3) / 1e12, 2),
            "fp32_frac_of_lane_issue": round(ops / (ms * 1e-3) / (peak * 64), 4)}


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
    ap.add_argument("--workload", default="chain24",
                    choices=["chain24", "xtrans26", "batch64", "grouping", "serve", "micro24", "highlight24",
                             "copy"])
    ap.add_argument("--rates", default="4,8,16", help="serve: arrival rates (requests/s)")
    ap.add_argument("--duration", type=float, default=20.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--serve-duration", type=float, default=15.0,
                    help="chain24: seconds of arrivals per rate of the served-loop leg (0 skips it)")
    ap.add_argument("--lanes", type=int, default=3, help="serve: requests in flight")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): images are independent, "
              "so N GPUs would each run this same single-GPU workload", file=sys.stderr)
        sys.exit(2)
    if args.workload != "chain24":
        run = {"xtrans26": bench_xtrans26, "batch64": bench_batch64, "grouping": bench_grouping,
               "serve": bench_serve,
               "micro24": bench_micro24, "highlight24": bench_highlight24, "copy": bench_copy}[args.workload]
        print(json.dumps(run(args)), flush=True)
        return

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
    # The roofline covers the largest single kernel of the step: amaze_kernel.
    # (The defringe stage takes about as long, but as nine kernels, none of
    # them above 0.5 ms on 24 MP.)
    dominant = "amaze" if "amaze" in stage_ms else max(stage_ms, key=stage_ms.get)
    alg_bytes = STAGE_BYTES_PER_PX[dominant] * h * w
    achieved = alg_bytes / (stage_ms[dominant] * 1e-3) / 1e9
    line = {
        "metric": "MP/s full develop chain (24 MP Bayer) + p99 latency at fixed QPS vs plain path",
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
        # AMaZE does not stream: its time goes to VALU issue
        "valu_roofline": valu_roofline("amaze_kernel", "amaze", stage_ms["amaze"], h * w),
    }
    gpu_out = chain.result()  # the last timed step's output
    if not args.no_cpu:
        threads = cpu_threads(args)
        print(f"bench: CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        raw3 = scene.cfa_to_rgb3(cfa)
        line["cpu_baseline"], cpu_out = cpu_baseline(raw3, stages, threads)
        # Parity at full size against that output.  The reference's AMaZE
        # output within its edge tiles depends on its thread schedule; the band
        # is one tile (160 px) widened by the later stages' reach (defringe
        # 10 + 10 + 14 + 10, detail tone map 12 + 12).
        band = 160 + 44 + 24
        line["parity"] = {"against": f"{line['cpu_baseline']['kind']} CPU output of the same image",
                          "edge_band_px": band, "tolerance": {"max_abs": 2e-4, "psnr_db": 70.0},
                          **parity_block(gpu_out, cpu_out, band)}
        # error of both outputs against an FP64 evaluation of defringe -> detail
        # TM from the HIP AMaZE output (bit-identical to the restated AMaZE)
        chain.run_stage("amaze", d_cfa)
        acc.synchronize()
        print("bench: FP64 evaluation of the post-demosaic stages", file=sys.stderr, flush=True)
        f64 = fp64_post_chain(chain.result(), stages[1:], threads)
        edge = edge_band_mask(h, w, band)
        def err(x):
            d = np.abs(x.astype(np.float64) - f64)
            return {"max_abs": float(d.max()), "max_abs_outside_edge_band": float(d[~edge].max())}
        line["parity"]["vs_fp64"] = {"against": "FP64 evaluation of " + " -> ".join(stages[1:]) +
                                                " from the HIP AMaZE output", "hip": err(gpu_out),
                                     line["cpu_baseline"]["kind"]: err(cpu_out)}
        del f64
    else:
        line["cpu_baseline"] = None
    acc.close()
    if args.serve_duration > 0:
        # the metric's second half: p50/p99 latency of the request loop at fixed
        # arrival rates, beside the plain per-image path
        rates = [float(r) for r in args.rates.split(",") if r]
        line["serve"] = serve_leg(rates, args.serve_duration, args.lanes, burst=False)
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# other workloads

def _log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _threads(args):
    return cpu_threads(args)


def _ref_lib(threads):
    """oracle/_ref (the reference's CPU code built from its sources), else the restatement."""
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_cpu.so")
    if os.path.exists(ref_path):
        lib = ctypes.CDLL(ref_path)
        lib.ref_set_threads(threads)
        return lib, "reference"
    os.environ["OMP_NUM_THREADS"] = str(threads)
    return ctypes.CDLL(os.path.join(ROOT, "oracle", "liblre_oracle.so")), "port"


def cpu_baseline_stages(raw3, filters, xtrans, stages, threads, label):
    """One image through `stages` on the CPU (reference names, in-place as the
    develop pipeline calls them); returns the cpu_baseline object."""
    f32p = ctypes.POINTER(ctypes.c_float)
    i8p = ctypes.POINTER(ctypes.c_int8)
    lib, kind = _ref_lib(threads)
    R = kind == "reference"
    h, w, _ = raw3.shape
    rgb = np.zeros_like(raw3)
    tmp = np.zeros_like(raw3)
    cm = np.ones(4, np.float32)
    xt = None if xtrans is None else np.ascontiguousarray(np.asarray(xtrans).reshape(36), np.int8)
    P = lambda a: a.ctypes.data_as(f32p)
    t0 = time.perf_counter()
    for s in stages:
        if s == "demosaic" and filters != 9:
            if R:
                lib.ref_amaze(P(raw3), P(rgb), w, h, ctypes.c_uint32(filters), P(cm), ctypes.c_float(1.0))
            else:
                lib.oracle_amaze(P(raw3), P(rgb), w, h, ctypes.c_uint32(filters), P(cm), ctypes.c_float(1.0), 0)
        elif s == "demosaic":
            fn = lib.ref_xtrans3 if R else lib.oracle_xtrans3
            args = (P(raw3), P(rgb), w, h, xt.ctypes.data_as(i8p), ctypes.c_float(1.0))
            fn(*args) if R else fn(*args, 0)
        elif s == "lateral_ca":
            fn = lib.ref_lateral_ca if R else lib.oracle_lateral_ca
            fn(P(rgb), P(tmp), w, h, 96, 3, ctypes.c_float(6.0), ctypes.c_float(0.02), 3)
            rgb, tmp = tmp, rgb
        elif s == "axial_ca":
            fn = lib.ref_axial_ca if R else lib.oracle_axial_ca
            fn(P(rgb), P(tmp), w, h, 6, ctypes.c_float(1e-4), ctypes.c_float(0.3))
            rgb, tmp = tmp, rgb
        elif s == "defringe":
            fn = lib.ref_defringe if R else lib.oracle_defringe
            fn(P(rgb), P(rgb), w, h, ctypes.c_float(10.0), ctypes.c_float(0.1), ctypes.c_float(0.15),
               ctypes.c_float(10.0), 0, ctypes.c_float(0.3))
        elif s == "detail_tm":
            fn = lib.ref_detail_tonemap if R else lib.oracle_detail_tonemap
            fn(P(rgb), w, h)
    dt = time.perf_counter() - t0
    return {"value": round(h * w / 1e6 / dt, 3), "unit": "MP/s", "cores": threads, "kind": kind,
            "sample": f"one {w}x{h} {label} image through {' -> '.join(stages)} ({dt:.2f} s)"}


def bench_xtrans26(args):
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    from libraw_enhanced_amd.pipeline import DevelopChain
    h, w = 4160, 6240
    _log(f"generating {w}x{h} X-Trans scene (radial pre-warp R 1.0008, B 0.9993)")
    cfa = scene.xtrans_scene(h, w, scene.BASE_SEED)
    acc = HIPAccelerator(0)
    d_cfa = acc.dev_from_host(cfa)
    stages = ["xtrans3", "lateral_ca", "axial_ca"]
    chain = DevelopChain(acc, h, w, stages, filters=9, xtrans=scene.XTRANS)
    stream = native().lre_hip_stream(acc.handle)
    ev = HipEvents()
    for _ in range(args.warmup):
        chain.run(d_cfa)
    acc.synchronize()
    marks = [[ev.create() for _ in range(len(stages) + 1)] for _ in range(args.steps)]
    _log(f"timing {args.steps} steps of {' -> '.join(stages)}")
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev.record(marks[k][0], stream)
        for i, s in enumerate(stages):
            chain.run_stage(s, d_cfa)
            ev.record(marks[k][i + 1], stream)
    acc.synchronize()
    elapsed = time.perf_counter() - t0
    stage_ms = {s: float(np.mean([ev.elapsed_ms(marks[k][i], marks[k][i + 1]) for k in range(args.steps)]))
                for i, s in enumerate(stages)}
    mp = h * w / 1e6
    # HBM roofline of the X-Trans kernel, the largest launch: 4 B CFA in, 12 B RGB out per pixel
    alg_bytes = STAGE_BYTES_PER_PX["xtrans3"] * h * w
    achieved = alg_bytes / (stage_ms["xtrans3"] * 1e-3) / 1e9
    line = {"metric": "MP/s X-Trans develop chain (26 MP)", "value": round(mp * args.steps / elapsed, 2),
            "unit": "MP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene, radial lateral CA pre-warp, Poisson-Gaussian noise)",
            "config": {"workload": f"{' -> '.join(stages)} on one {w}x{h} X-Trans float32 mosaic",
                       "image_mp": round(mp, 2), "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()}},
            "roofline": {"kernel": "xtrans3", "bound": "hbm", "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": load_traffic("xtrans"), "alg_bytes_per_launch": alg_bytes},
            "valu_roofline": valu_roofline("xtrans_kernel", "xtrans3", stage_ms["xtrans3"], h * w)}
    if not args.no_cpu:
        threads = _threads(args)
        _log(f"CPU baseline on {threads} threads")
        raw3 = scene.cfa_to_rgb3(cfa, 0, scene.XTRANS)
        line["cpu_baseline"] = cpu_baseline_stages(raw3, 9, scene.XTRANS, ["demosaic", "lateral_ca", "axial_ca"],
                                                   threads, "X-Trans")
    acc.close()
    return line


def bench_micro24(args):
    """Micro-contrast (core/cpu_accelerator.cpp:3135-3319) on one resident 24 MP
    RGB image with a highlight mask, as the develop pipeline calls it (strength 8,
    target contrast 0.04).  The reference's own code for this stage needs
    Apple's simd headers, so the CPU baseline is the restatement ("port")."""
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    h, w = args.height, args.width
    _log(f"generating {w}x{h} RGB scene and mask")
    img = np.ascontiguousarray(scene.render(h, w, scene.BASE_SEED) * 1.3, np.float32)
    t = np.clip((img[:, :, 1] - 0.95) / 0.05, 0.0, 1.0)
    mask = np.ascontiguousarray(t * t * (3.0 - 2.0 * t), np.float32)
    acc = HIPAccelerator(0)
    lib = native()
    d_in, d_mask = acc.dev_from_host(img), acc.dev_from_host(mask)
    d_out = acc.dev_alloc(img.nbytes)
    stream = lib.lre_hip_stream(acc.handle)
    args_c = (ctypes.c_float(1.3), ctypes.c_float(8.0), ctypes.c_float(0.04))

    def step():
        rc = lib.lre_hip_dev_micro_contrast(acc.handle, d_in.ptr, d_out.ptr, w, h, *args_c, d_mask.ptr, stream)
        assert rc == 0, rc
    for _ in range(args.warmup):
        step()
    acc.synchronize()
    ev = HipEvents()
    marks = [(ev.create(), ev.create()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in marks:
        ev.record(a, stream)
        step()
        ev.record(b, stream)
    acc.synchronize()
    elapsed = time.perf_counter() - t0
    ms = float(np.mean([ev.elapsed_ms(a, b) for a, b in marks]))
    mp = h * w / 1e6
    alg = 28 * h * w  # RGB in 12 + mask 4 + RGB out 12 bytes per pixel
    line = {"metric": "MP/s micro-contrast (24 MP)", "value": round(mp * args.steps / elapsed, 2), "unit": "MP/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene x 1.3, green smoothstep mask)",
            "config": {"workload": f"enhance_micro_contrast with mask on one {w}x{h} float32 RGB image",
                       "image_mp": round(mp, 2), "stage_ms": round(ms, 4)},
            "roofline": {"kernel": "micro_contrast", "bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                         "alg_bytes_per_launch": alg}}
    if not args.no_cpu:
        threads = _threads(args)
        _log(f"CPU baseline (restatement) on {threads} threads")
        os.environ["OMP_NUM_THREADS"] = str(threads)
        olib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liblre_oracle.so"))
        f32p = ctypes.POINTER(ctypes.c_float)
        out = np.empty_like(img)
        t1 = time.perf_counter()
        olib.oracle_micro_contrast(img.ctypes.data_as(f32p), out.ctypes.data_as(f32p), w, h,
                                   ctypes.c_float(1.3), ctypes.c_float(8.0), ctypes.c_float(0.04),
                                   mask.ctypes.data_as(f32p))
        dt = time.perf_counter() - t1
        line["cpu_baseline"] = {"value": round(mp / dt, 3), "unit": "MP/s", "cores": threads, "kind": "port",
                                "sample": f"the whole {w}x{h} image ({dt:.2f} s)"}
    acc.close()
    return line


def bench_highlight24(args):
    """Highlight recovery (core/libraw_wrapper.cpp:1095-1383) on one resident
    24 MP RGB image, as postprocess calls it for highlight_mode >= 4: with the
    soft mask and the effective white balance (2.0, 1.0, 1.5), saturation
    threshold 0.9.  The scene is scaled by 1.3 with R x 1.15 and B x 1.2, so
    part of it is blown and the LoCA pass adjusts pixels.  The reference
    function needs LibRaw, so the CPU baseline is the restatement ("port"),
    serial as the reference is."""
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    h, w = args.height, args.width
    _log(f"generating {w}x{h} magenta-cast RGB scene")
    img = scene.render(h, w, scene.BASE_SEED) * 1.3
    img[..., 0] *= 1.15
    img[..., 2] *= 1.2
    img = np.ascontiguousarray(img, np.float32)
    acc = HIPAccelerator(0)
    lib = native()
    d_in = acc.dev_from_host(img)
    d_out, d_mask, d_max = acc.dev_alloc(img.nbytes), acc.dev_alloc(h * w * 4), acc.dev_alloc(4)
    stream = lib.lre_hip_stream(acc.handle)
    f32p = ctypes.POINTER(ctypes.c_float)
    wb = np.float32([2.0, 1.0, 1.5])
    sat = ctypes.c_float(0.9)

    def step():
        rc = lib.lre_hip_dev_recover_highlights(acc.handle, d_in.ptr, d_out.ptr, w, h, sat, d_mask.ptr,
                                                wb.ctypes.data_as(f32p), d_max.ptr, stream)
        assert rc == 0, rc
    for _ in range(args.warmup):
        step()
    acc.synchronize()
    ev = HipEvents()
    marks = [(ev.create(), ev.create()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in marks:
        ev.record(a, stream)
        step()
        ev.record(b, stream)
    acc.synchronize()
    elapsed = time.perf_counter() - t0
    ms = float(np.mean([ev.elapsed_ms(a, b) for a, b in marks]))
    mp = h * w / 1e6
    alg = 28 * h * w  # RGB in 12 + mask 4 + RGB out 12 bytes per pixel
    line = {"metric": "MP/s highlight recovery (24 MP)", "value": round(mp * args.steps / elapsed, 2),
            "unit": "MP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene x 1.3, R x 1.15, B x 1.2)",
            "config": {"workload": f"recover_highlights with mask and WB on one {w}x{h} float32 RGB image",
                       "image_mp": round(mp, 2), "stage_ms": round(ms, 4), "saturation_threshold": 0.9},
            "roofline": {"kernel": "recover_highlights", "bound": "hbm",
                         "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                         "alg_bytes_per_launch": alg}}
    if not args.no_cpu:
        _log("CPU baseline (restatement, serial as the reference)")
        olib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liblre_oracle.so"))
        out, mask = img.copy(), np.empty((h, w), np.float32)
        mv, cnt = ctypes.c_float(0.0), (ctypes.c_longlong * 3)()
        t1 = time.perf_counter()
        olib.oracle_recover_highlights(out.ctypes.data_as(f32p), w, h, sat, mask.ctypes.data_as(f32p),
                                       wb.ctypes.data_as(f32p), ctypes.byref(mv), cnt, 0)
        dt = time.perf_counter() - t1
        line["config"].update({"highlight_px": cnt[0], "blown_px": cnt[1], "loca_px": cnt[2]})
        line["cpu_baseline"] = {"value": round(mp / dt, 3), "unit": "MP/s", "cores": 1, "kind": "port",
                                "sample": f"the whole {w}x{h} image ({dt:.2f} s)"}
        got = d_out.download((h, w, 3))
        line["config"]["equal_to_port"] = bool(np.array_equal(got, out))
    acc.close()
    return line


def bench_copy(args):
    """Device-to-device copies of a known size: the reference point that scales
    the FETCH_SIZE / WRITE_SIZE counters into bytes."""
    from libraw_enhanced_amd.accelerator import HIPAccelerator, native
    nbytes = 1 << 30
    acc = HIPAccelerator(0)
    a, b = acc.dev_alloc(nbytes), acc.dev_alloc(nbytes)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p]
    stream = native().lre_hip_stream(acc.handle)
    for _ in range(max(1, args.warmup)):
        assert hip.hipMemcpyAsync(b.ptr, a.ptr, nbytes, 3, stream) == 0  # device to device
    acc.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert hip.hipMemcpyAsync(b.ptr, a.ptr, nbytes, 3, stream) == 0
    acc.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    acc.close()
    return {"metric": "device-to-device copy", "value": round(2 * nbytes / dt / 1e9, 1), "unit": "GB/s",
            "bytes_per_copy": nbytes, "steps": args.steps, "ms_per_copy": round(dt * 1e3, 3)}


# (count, height, width, filters) of the 64-image batch
BATCH64 = [(16, 3000, 4000, "bayer"), (16, 4000, 6000, "bayer"), (12, 5504, 8256, "bayer"),
           (8, 6336, 9504, "bayer"), (12, 5152, 7728, "xtrans")]


def _plain_path(acc, raw3, filters, xtrans, stages):
    """The reference's call pattern on the HIP backend: host arrays into and out
    of every Accelerator method (the per-image plain path)."""
    if filters == 9:
        rgb = acc.demosaic_xtrans_3pass(raw3, xtrans, 1.0)
    else:
        rgb = acc.demosaic_bayer_amaze(raw3, filters, (1, 1, 1, 1), 1.0)
    for s in stages[1:]:
        if s == "axial_ca":
            rgb = acc.ca_axial_cleanup(rgb)
        elif s == "defringe":
            rgb = acc.defringe(rgb, rgb)
        elif s == "detail_tm":
            rgb = acc.apply_detail_preserving_tonemap(rgb)
    return rgb


def bench_batch64(args):
    from libraw_enhanced_amd import develop as dv
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    acc = HIPAccelerator(0)
    stages = list(dv.BATCH_CHAIN)
    images, bufs, total_mp, k = [], [], 0.0, 0
    _log("generating 64 images (2080.8 MP) and uploading them")
    t_gen = time.perf_counter()
    samples = {}
    for count, h, w, kind in BATCH64:
        for _ in range(count):
            seed = scene.BASE_SEED + k
            if kind == "bayer":
                cfa, filters, xt = scene.bayer_scene(h, w, seed), scene.RGGB, None
            else:
                cfa, filters, xt = scene.xtrans_scene(h, w, seed, lateral=None), 9, scene.XTRANS
            d_cfa = acc.dev_from_host(cfa)
            d_rgb = acc.dev_alloc(h * w * 12)
            bufs.append((d_cfa, d_rgb))
            images.append(dv.make_image(w, h, filters=filters, xtrans=xt, d_cfa=d_cfa, maximum_value=1.0,
                                        d_rgb=d_rgb))
            samples.setdefault((h, w, kind), (cfa, filters, xt))
            total_mp += h * w / 1e6
            k += 1
    _log(f"generated in {time.perf_counter() - t_gen:.1f} s; autotuning on one image per size class")
    b = dv.BatchDeveloper(0, 2)
    tune_set = [images[sum(c for c, *_ in BATCH64[:i])] for i in range(len(BATCH64))]
    tune_ms = b.autotune(tune_set, stages, reps=1)
    cfg = {**b.config(), **b.grouping(), **b.tiles()}
    _log(f"autotune: {cfg} ({tune_ms:.1f} ms for the 5-image sample)")
    # eager batch in both post-stage modes (group: one launch per stage over
    # the batch; lanes: per image over streams), each after a warm-up run
    eager = {}
    picked = cfg["group_mode"]
    for mode in (not picked, picked):
        b.set_grouping(mode)
        b.run(images, stages)
        b.synchronize()
        t = time.perf_counter()
        b.run(images, stages)
        b.synchronize()
        eager["group" if mode else "lanes"] = round(time.perf_counter() - t, 3)
    eager_s = eager["group" if picked else "lanes"]
    _log(f"eager batch: {eager}")
    b.capture(images, stages)
    for _ in range(args.warmup):
        b.replay()
    b.synchronize()
    steps = max(1, args.steps)
    _log(f"timing {steps} graph replays of the 64-image batch")
    t0 = time.perf_counter()
    for _ in range(steps):
        b.replay()
    b.synchronize()
    elapsed = time.perf_counter() - t0
    value = total_mp * steps / elapsed
    # the plain per-image path on one image of each size class
    _log("timing the plain per-image path on one image per size class")
    plain_mp, plain_s = 0.0, 0.0
    for (h, w, kind), (cfa, filters, xt) in samples.items():
        raw3 = scene.cfa_to_rgb3(cfa, filters if xt is None else 0, xt)
        _plain_path(acc, raw3[:64, :64].copy(), filters, xt, stages)  # warm
        t = time.perf_counter()
        _plain_path(acc, raw3, filters, xt, stages)
        plain_s += time.perf_counter() - t
        plain_mp += h * w / 1e6
    plain = plain_mp / plain_s
    line = {"metric": "MP/s batched develop (64 uneven images)", "value": round(value, 2), "unit": "MP/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed * 1e3 / steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scenes base..base+63, Poisson-Gaussian noise)",
            "config": {"workload": "demosaic -> axial_ca -> defringe -> detail_tm over 16x12 MP + 16x24 MP + "
                                   "12x45.4 MP + 8x60.2 MP RGGB + 12x39.8 MP X-Trans, float32, one batch",
                       "total_mp": round(total_mp, 1), "autotuned": cfg, "eager_batch_s": eager_s,
                       "eager_batch_s_by_mode": eager,
                       "plain_per_image_mp_s": round(plain, 2), "batched_vs_plain": round(value / plain, 2),
                       "plain_sample_mp": round(plain_mp, 1)}}
    if not args.no_cpu:
        threads = _threads(args)
        _log(f"CPU baseline on {threads} threads (one 12 MP Bayer image)")
        cfa, filters, xt = samples[(3000, 4000, "bayer")]
        line["cpu_baseline"] = cpu_baseline_stages(scene.cfa_to_rgb3(cfa), filters, None, stages, threads, "RGGB")
    b.close()
    acc.close()
    return line


# request mix of the served loop: (share, height, width, kind)
SERVE_MIX = [(0.4, 4000, 6000, "bayer"), (0.3, 5504, 8256, "bayer"), (0.2, 6336, 9504, "bayer"),
             (0.1, 5152, 7728, "xtrans")]
WB = (2.1, 1.0, 1.5, 1.0)


def serve_leg(rates, duration, lanes, burst=True):
    """The request loop at fixed arrival rates: Poisson arrivals of uint16
    requests (SERVE_MIX) for `duration` seconds per rate, p50/p99 latency per
    rate; a closed burst for the capacity; and the plain per-image path (host
    arrays in and out of every Accelerator method) on one request per class."""
    from libraw_enhanced_amd import develop as dv
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    _log("serve: generating uint16 requests (2 scenes per size class) into pinned memory")
    inputs, outputs = {}, {}
    for ci, (_, h, w, kind) in enumerate(SERVE_MIX):
        for v in range(2):
            seed = scene.BASE_SEED + 1000 + 10 * ci + v
            if kind == "bayer":
                cfa = scene.bayer_scene(h, w, seed)
                raw = scene.quantize_u16(cfa, scene.RGGB, None)
            else:
                cfa = scene.xtrans_scene(h, w, seed, lateral=None)
                raw = scene.quantize_u16(cfa, 0, scene.XTRANS)
            pa = dv.PinnedArray(raw.shape, np.uint16)
            pa.array[...] = raw
            inputs[(ci, v)] = pa
        # one pinned output per class (the benchmark does not read the results)
        outputs[ci] = dv.PinnedArray((h, w, 3), np.float32)
    shares = np.cumsum([m[0] for m in SERVE_MIX])

    def image_for(ci):
        _, h, w, kind = SERVE_MIX[ci]
        if kind == "bayer":
            return dv.make_image(w, h, filters=scene.RGGB, cam_mul=WB, black=512, white=16383, wb=WB)
        return dv.make_image(w, h, filters=9, xtrans=scene.XTRANS, cam_mul=WB, black=512, white=16383, wb=WB)

    acc = HIPAccelerator(0)  # device init before the server threads
    server = dv.DevelopServer(0, lanes=lanes, hbm_budget_bytes=96 << 30)
    # warm-up: one request of each class (pool and code objects)
    for ci in range(len(SERVE_MIX)):
        server.submit(10_000 + ci, image_for(ci), inputs[(ci, 0)].ptr, outputs[ci].ptr)
    got = 0
    while got < len(SERVE_MIX):
        got += len(server.poll(timeout_ms=1000))
    results = []
    for rate in rates:
        n = max(1, int(round(rate * duration)))
        arr = dv.poisson_schedule(rate, n, seed=int(rate * 1000))
        rng = np.random.Generator(np.random.PCG64(int(rate * 7)))
        classes = np.searchsorted(shares, rng.random(n), side="right").clip(0, len(SERVE_MIX) - 1)
        variants = rng.integers(0, 2, n)

        def make(k):
            ci = int(classes[k])
            _, h, w, _ = SERVE_MIX[ci]
            return image_for(ci), inputs[(ci, int(variants[k]))].ptr, outputs[ci].ptr, h * w / 1e6, None

        _log(f"serve: rate {rate}/s, {n} requests over {arr[-1]:.1f} s")
        progress = {"t": time.perf_counter(), "k": 0}

        def on_done(c):
            progress["k"] += 1
            if time.perf_counter() - progress["t"] > 30:
                progress["t"] = time.perf_counter()
                _log(f"serve: rate {rate}/s, {progress['k']} of {n} requests done")

        wall, lat, mp_total = dv.run_open_loop(server, arr, make, on_done=on_done)
        rep = dv.ServeReport(rate=rate, requests=n, duration_s=wall, mp_total=mp_total, latencies_s=lat)
        summ = rep.summary()
        summ["offered_mp_s"] = round(mp_total / arr[-1], 2)
        results.append(summ)
        _log(f"serve: rate {rate}/s: {summ}")
    capacity = n_burst = None
    if burst:
        # capacity: a closed burst of requests, all queued at once
        n_burst = max(len(SERVE_MIX), int(round(2 * duration)))
        rng = np.random.Generator(np.random.PCG64(99))
        classes = np.searchsorted(shares, rng.random(n_burst), side="right").clip(0, len(SERVE_MIX) - 1)

        def make_burst(k):
            ci = int(classes[k])
            _, h, w, _ = SERVE_MIX[ci]
            return image_for(ci), inputs[(ci, k % 2)].ptr, outputs[ci].ptr, h * w / 1e6, None

        _log(f"serve: capacity, burst of {n_burst} requests")
        wall, lat, mp_burst = dv.run_open_loop(server, np.zeros(n_burst), make_burst)
        capacity = mp_burst / wall
    _, peak = server.footprint()
    server.close()
    # the plain per-image path: one request of each class, host arrays through every stage
    _log("serve: timing the plain per-image path (one request per class)")
    plain_lat = {}
    for ci, (_, h, w, kind) in enumerate(SERVE_MIX):
        raw = inputs[(ci, 0)].array
        t = time.perf_counter()
        img4 = np.zeros((h, w, 4), np.uint16)
        if kind == "bayer":  # LibRaw's 4-channel layout: the native channel (G2 = 3)
            r, c = np.arange(h)[:, None], np.arange(w)[None, :]
            f = (scene.RGGB >> ((((r << 1) & 14) | (c & 1)) << 1)) & 3
        else:
            f = scene.fcol_xtrans(h, w)
        np.put_along_axis(img4, f[..., None].astype(np.int64), raw[..., None], axis=2)
        img4 = acc.apply_black_level_correction(img4, (0, 0, 0, 0), 512, 16383, None)
        rgb = acc.apply_white_balance(img4, WB, scene.RGGB if kind == "bayer" else 9,
                                      None if kind == "bayer" else scene.XTRANS)
        if kind == "bayer":
            rgb = acc.demosaic_bayer_amaze(rgb, scene.RGGB, WB, 15871.0)
        else:
            rgb = acc.demosaic_xtrans_3pass(rgb, scene.XTRANS, 15871.0)
        rgb = acc.ca_axial_cleanup(rgb)
        rgb = acc.defringe(rgb, rgb)
        rgb = acc.apply_detail_preserving_tonemap(rgb)
        plain_lat[ci] = time.perf_counter() - t
    acc.close()
    for pa in list(inputs.values()) + list(outputs.values()):
        pa.free()
    mean_mp = sum(m[0] * m[1] * m[2] / 1e6 for m in SERVE_MIX)
    mean_plain_s = sum(m[0] * plain_lat[ci] for ci, m in enumerate(SERVE_MIX))
    plain_mp_s = mean_mp / mean_plain_s
    out = {"workload": "ingest -> demosaic -> axial_ca -> defringe -> detail_tm, Poisson arrivals, "
                       "40% 24 MP / 30% 45.4 MP / 20% 60.2 MP Bayer / 10% 39.8 MP X-Trans",
           "rates": results, "lanes": lanes, "duration_s_per_rate": duration, "peak_device_bytes": int(peak),
           "plain_per_image_ms": {f"{m[1]}x{m[2]}{'x' if m[3] == 'xtrans' else ''}":
                                  round(plain_lat[ci] * 1e3, 1) for ci, m in enumerate(SERVE_MIX)},
           "plain_per_image_mp_s": round(plain_mp_s, 2),
           "plain_mean_request_ms": round(mean_plain_s * 1e3, 1)}
    if capacity is not None:
        out.update({"capacity_mp_s": round(capacity, 2), "capacity_requests": n_burst,
                    "served_capacity_vs_plain": round(capacity / plain_mp_s, 2)})
    return out


# (count, height, width) of the grouping batches
GROUPING = [(256, 480, 640), (64, 1536, 2048), (16, 3000, 4000)]


def bench_grouping(args):
    from libraw_enhanced_amd import develop as dv
    from libraw_enhanced_amd import scene
    from libraw_enhanced_amd.accelerator import HIPAccelerator
    acc = HIPAccelerator(0)
    stages = list(dv.BATCH_CHAIN)
    steps = max(1, args.steps)
    rows = []
    for n, h, w in GROUPING:
        _log(f"grouping: {n} x {w}x{h}")
        cfa = scene.bayer_scene(h, w, scene.BASE_SEED)
        bufs = [(acc.dev_from_host(cfa), acc.dev_alloc(h * w * 12)) for _ in range(n)]
        imgs = [dv.make_image(w, h, filters=scene.RGGB, d_cfa=c, maximum_value=1.0, d_rgb=r) for c, r in bufs]
        b = dv.BatchDeveloper(0, 1)
        row = {"images": n, "width": w, "height": h, "total_mp": round(n * h * w / 1e6, 1)}
        for mode in (False, True):
            for lanes in (1, 2, 3):
                b.set_grouping(mode)
                b.set_config(lanes, 2, 4)
                b.run(imgs, stages)
                b.synchronize()
                t = time.perf_counter()
                for _ in range(steps):
                    b.run(imgs, stages)
                b.synchronize()
                eager = (time.perf_counter() - t) / steps
                b.capture(imgs, stages)
                b.replay()
                b.synchronize()
                t = time.perf_counter()
                for _ in range(steps):
                    b.replay()
                b.synchronize()
                graph = (time.perf_counter() - t) / steps
                key = f"{'group' if mode else 'lanes'}_{lanes}"
                row[key] = {"eager_ms": round(eager * 1e3, 2), "graph_ms": round(graph * 1e3, 2)}
        best = {m: min(v["graph_ms"] for k, v in row.items() if k.startswith(m)) for m in ("group", "lanes")}
        row["best_graph_ms"] = best
        row["group_vs_lanes"] = round(best["lanes"] / best["group"], 2)
        rows.append(row)
        b.close()
        del imgs, bufs
    return {"metric": "batch post-stage modes (ms per batch)", "value": rows[0]["group_vs_lanes"],
            "unit": "x (lanes time / group time, 256 x 0.3 MP)", "n_gpus": 1, "steps": steps,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic (one seeded scene per size)",
            "config": {"workload": " -> ".join(stages) + " over batches of equal Bayer images", "batches": rows}}


def bench_serve(args):
    rates = [float(r) for r in args.rates.split(",") if r]
    leg = serve_leg(rates, args.duration, args.lanes)
    return {"metric": "p99 latency of the served develop loop", "value": leg["rates"][-1]["p99_ms"], "unit": "ms",
            "n_gpus": 1, "steps": sum(r["requests"] for r in leg["rates"]), "warmup": len(SERVE_MIX),
            "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "u16->f32", "data": "synthetic uint16 CFA (seeded scenes, black 512, white 16383)",
            "config": leg}


if __name__ == "__main__":
    main()
