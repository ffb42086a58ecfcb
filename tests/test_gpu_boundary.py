"""The C++ HIPAccelerator on the GPU, inside the reference's facade pattern
(tests/boundary_shim.cpp: GPU backend first, the reference's CPUAccelerator
when it returns false, core/accelerator.cpp:141-157):

* on device 0 the HIP backend serves AMaZE, in-place defringe, axial CA and the
  tone map, with the oracle's results (bit-exact / within max-abs 2e-4);
* a backend bound to a device that does not exist fails to initialise, every
  call falls back, and the results are the reference CPU code's;
* a call the C ABI rejects (axial CA radius 17, above the kernel's 16) returns
  false on a working device, and the facade's CPU result is served instead."""
import ctypes

import numpy as np
import pytest

from _common import oracle_amaze, oracle_tone_mapping, psnr
from libraw_enhanced_amd import scene
from test_boundary_cpp import P, boundary, cpu_reference, facade_calls

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    h, w = 600, 900
    raw3 = scene.cfa_to_rgb3(scene.bayer_scene(h, w, scene.BASE_SEED + 81))
    img = np.ascontiguousarray(oracle_amaze(raw3) * 1.6, np.float32)
    return img, raw3


def test_hip_backend_serves_the_calls(data):
    from test_ca_cpu import axial
    from test_gpu_develop import _oracle_defringe
    from _common import oracle
    img, raw3 = data
    lib = boundary()
    f = lib.boundary_create(0)
    assert lib.boundary_hip_ok(f) == 1
    served, outs = facade_calls(lib, f, img, raw3)
    lib.boundary_destroy(f)
    assert set(served.values()) == {1}, served
    assert np.array_equal(outs["amaze"], oracle_amaze(raw3))
    want = _oracle_defringe(img, True)
    assert float(np.abs(outs["defringe"] - want).max()) <= 2e-4 and psnr(outs["defringe"], want) >= 70
    want = axial(oracle(), "oracle_axial_ca", img)
    assert float(np.abs(outs["axial"] - want).max()) <= 2e-4
    assert np.array_equal(outs["tone"], oracle_tone_mapping(img, 1.0))


def test_missing_device_falls_back_to_cpu(data):
    from _common import ref
    img, raw3 = data
    lib = boundary()
    ref().ref_set_threads(1)  # the reference's AMaZE edge tiles depend on the thread schedule
    f = lib.boundary_create(99)
    assert lib.boundary_hip_ok(f) == 0
    served, outs = facade_calls(lib, f, img, raw3)
    lib.boundary_destroy(f)
    assert set(served.values()) == {0}, served
    want = cpu_reference(img, raw3)
    for k in want:
        assert np.array_equal(outs[k], want[k]), k


def test_rejected_call_falls_back_on_a_working_device(data):
    from _common import ref
    img, _ = data
    h, w, _ = img.shape
    lib = boundary()
    f = lib.boundary_create(0)
    assert lib.boundary_hip_ok(f) == 1
    out = np.zeros_like(img)
    assert lib.boundary_axial(f, P(img), P(out), w, h, 17, 1e-4, 0.3) == 0  # HIP said false: CPU served
    lib.boundary_destroy(f)
    want = np.zeros_like(img)
    r = ref()
    r.ref_set_threads(4)
    r.ref_axial_ca(P(img), P(want), w, h, 17, ctypes.c_float(1e-4), ctypes.c_float(0.3))
    assert np.array_equal(out, want)
