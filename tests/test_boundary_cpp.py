"""The C++ HIPAccelerator (the backend class a maintainer adds next to the
reference's CPUAccelerator) compiles against the reference's own
core/accelerator.h and links against liblre_hip.so.  Runs only where the
reference tree is available (it is read, never copied)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CORE = "/root/reference/core"


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_CORE, "accelerator.h")),
                    reason="reference tree not present")
def test_hip_accelerator_builds_against_reference_header(tmp_path):
    lib_dir = os.path.join(ROOT, "libraw_enhanced_amd", "_lib")
    if not os.path.exists(os.path.join(lib_dir, "liblre_hip.so")):
        pytest.skip("liblre_hip.so not built")
    main = tmp_path / "main.cpp"
    main.write_text(
        '#include "hip_accelerator.h"\n'
        "#include <cstdio>\n"
        "int main() {\n"
        "  libraw_enhanced::HIPAccelerator acc;\n"
        "  bool ok = acc.initialize();\n"
        "  std::printf(\"initialized=%d available=%d\\n\", (int)ok, (int)acc.is_available());\n"
        "  float px[3] = {0.f, 0.f, 0.f};\n"
        "  libraw_enhanced::ImageBufferFloat in; in.image = reinterpret_cast<float (*)[3]>(px);\n"
        "  in.width = 1; in.height = 1;\n"
        "  return acc.tone_mapping(in, in, 1.0f) == ok ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "main"
    cmd = ["g++", "-std=c++17", "-O1", f"-I{REF_CORE}", f"-I{ROOT}/include",
           f"-I{ROOT}/libraw_enhanced_amd/csrc", str(main),
           f"{ROOT}/libraw_enhanced_amd/csrc/hip_accelerator.cpp",
           f"-L{lib_dir}", "-llre_hip", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "initialized=" in out.stdout


# ---- the facade around the backend (tests/boundary_shim.cpp, built by
# oracle/boundary_build.sh): the reference's selection, GPU backend first and
# the CPU backend when it returns false (core/accelerator.cpp:141-157) --------
import ctypes  # noqa: E402

import numpy as np  # noqa: E402

BOUNDARY = os.path.join(ROOT, "oracle", "_ref", "libhip_boundary.so")
F32P = ctypes.POINTER(ctypes.c_float)


def boundary():
    if not os.path.exists(BOUNDARY):
        pytest.skip("oracle/_ref/libhip_boundary.so not built (needs the reference tree at build time)")
    lib = ctypes.CDLL(BOUNDARY)
    lib.boundary_create.restype = ctypes.c_void_p
    lib.boundary_create.argtypes = [ctypes.c_int]
    for name in ("boundary_destroy", "boundary_hip_ok"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.boundary_amaze.argtypes = [ctypes.c_void_p, F32P, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, F32P,
                                   ctypes.c_float]
    lib.boundary_defringe.argtypes = [ctypes.c_void_p, F32P, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int]
    lib.boundary_axial.argtypes = [ctypes.c_void_p, F32P, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_float]
    lib.boundary_tone_mapping.argtypes = [ctypes.c_void_p, F32P, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    return lib


def P(a):
    return a.ctypes.data_as(F32P)


def facade_calls(lib, f, img, raw3):
    """amaze, in-place defringe, axial CA and the tone map through the facade:
    (which backend served each, outputs)."""
    h, w, _ = img.shape
    served, outs = {}, {}
    out = np.zeros_like(raw3)
    cm = np.ones(4, np.float32)
    served["amaze"] = lib.boundary_amaze(f, P(raw3), P(out), w, h, 0x94949494, P(cm), 1.0)
    outs["amaze"] = out
    a = img.copy()
    served["defringe"] = lib.boundary_defringe(f, P(a), P(a), w, h, 10.0, 10.0, 0)  # in place, as postprocess
    outs["defringe"] = a
    out = np.zeros_like(img)
    served["axial"] = lib.boundary_axial(f, P(img), P(out), w, h, 6, 1e-4, 0.3)
    outs["axial"] = out
    out = np.zeros_like(img)
    served["tone"] = lib.boundary_tone_mapping(f, P(img), P(out), w, h, 1.0)
    outs["tone"] = out
    return served, outs


def cpu_reference(img, raw3):
    from _common import ref
    r = ref()
    r.ref_set_threads(1)
    h, w, _ = img.shape
    outs = {}
    out = np.zeros_like(raw3)
    cm = np.ones(4, np.float32)
    r.ref_amaze(P(raw3), P(out), w, h, ctypes.c_uint32(0x94949494), P(cm), ctypes.c_float(1.0))
    outs["amaze"] = out
    a = img.copy()
    r.ref_defringe(P(a), P(a), w, h, ctypes.c_float(10.0), ctypes.c_float(0.1), ctypes.c_float(0.15),
                   ctypes.c_float(10.0), 0, ctypes.c_float(0.3))
    outs["defringe"] = a
    out = np.zeros_like(img)
    r.ref_axial_ca(P(img), P(out), w, h, 6, ctypes.c_float(1e-4), ctypes.c_float(0.3))
    outs["axial"] = out
    out = np.zeros_like(img)
    r.ref_tone_mapping(P(img), P(out), w, h, ctypes.c_float(1.0))
    outs["tone"] = out
    return outs


def test_facade_falls_back_to_cpu_without_a_device():
    """A backend that cannot initialise (no device here) returns false from
    every method, and the facade's CPU backend produces the reference's own
    results."""
    from libraw_enhanced_amd import scene
    lib = boundary()
    f = lib.boundary_create(0)
    if lib.boundary_hip_ok(f):
        lib.boundary_destroy(f)
        pytest.skip("a HIP device is present: covered by the GPU test")
    raw3 = scene.cfa_to_rgb3(scene.bayer_scene(200, 300, scene.BASE_SEED + 71))
    img = np.ascontiguousarray(scene.render(200, 300, scene.BASE_SEED + 72) * 1.5, np.float32)
    # one OpenMP thread for both copies of the CPU code (one libomp): the
    # reference's AMaZE edge tiles depend on the thread schedule
    from _common import ref
    ref().ref_set_threads(1)
    served, outs = facade_calls(lib, f, img, raw3)
    lib.boundary_destroy(f)
    assert set(served.values()) == {0}, served
    want = cpu_reference(img, raw3)
    for k in want:
        assert np.array_equal(outs[k], want[k]), k
