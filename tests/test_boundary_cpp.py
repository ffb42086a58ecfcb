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
