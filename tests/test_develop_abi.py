"""The ctypes mirrors of the batch / serve structs match the C layout of
include/lre_hip.h (checked by compiling a probe with the host compiler), and
the Python helpers of the request loop behave (no GPU needed)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "lre_hip.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
    printf("lre_image %zu\nlre_chain_params %zu\nlre_request %zu\nlre_completion %zu\n",
           sizeof(lre_image), sizeof(lre_chain_params), sizeof(lre_request), sizeof(lre_completion));
    F(lre_image, filters) F(lre_image, xtrans) F(lre_image, cam_mul) F(lre_image, d_cfa)
    F(lre_image, maximum_value) F(lre_image, d_raw) F(lre_image, black) F(lre_image, cblack)
    F(lre_image, wb) F(lre_image, adjust_threshold) F(lre_image, d_rgb)
    F(lre_chain_params, axial_radius) F(lre_chain_params, lateral_levels)
    F(lre_chain_params, tone_after_scale) F(lre_chain_params, color_transform)
    F(lre_chain_params, gamma_power) F(lre_chain_params, output_color_space)
    F(lre_chain_params, highlight_threshold) F(lre_chain_params, micro_target_contrast)
    F(lre_request, image) F(lre_request, h_raw) F(lre_request, h_rgb)
    F(lre_completion, status) F(lre_completion, t_submit) F(lre_completion, t_done)
    return 0;
}
"""


def test_struct_layout_matches_header(tmp_path):
    from libraw_enhanced_amd import develop as dv
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    types = {"lre_image": dv.Image, "lre_chain_params": dv.ChainParams, "lre_request": dv.Request,
             "lre_completion": dv.Completion}
    for key, val in got.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(types[t], m).offset == int(val), key
        else:
            assert ctypes.sizeof(types[key]) == int(val), key


def test_stage_codes_and_schedule():
    from libraw_enhanced_amd import develop as dv
    codes = dv.stage_codes(dv.BATCH_CHAIN)
    assert list(codes) == [1, 3, 4, 5]
    t = dv.poisson_schedule(8.0, 4000, 1)
    assert np.all(np.diff(t) > 0)
    assert abs(4000 / t[-1] - 8.0) < 0.5        # mean rate
    assert np.array_equal(t, dv.poisson_schedule(8.0, 4000, 1))


def test_serve_report_percentiles():
    from libraw_enhanced_amd import develop as dv
    r = dv.ServeReport(rate=4, requests=100, duration_s=10.0, mp_total=500.0,
                       latencies_s=[i / 1000 for i in range(1, 101)])
    s = r.summary()
    assert s["p50_ms"] == 50.0 and s["p99_ms"] == 99.0 and s["mp_per_s"] == 50.0


def test_batch_entry_points_exported():
    from libraw_enhanced_amd import develop as dv
    lib = dv.lib()
    for name in dv._SIGS:
        assert hasattr(lib, name), name
