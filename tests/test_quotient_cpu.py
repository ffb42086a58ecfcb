"""The defringe kernels take smoothstep quotients n / d as
q = n y; r = n - d q; q += r y; r = n - d q; q += r y, with y = 1/d rounded to
nearest (defringe.hip: quot).  This test runs the same sequence in C (fused
multiply-adds through fmaf) against IEEE division, for every numerator whose
quotient lies in [1/2, 2), for each ramp width the kernels use: every literal
edge pair of defringe.hip and the parameter-dependent edges at the default
and two other settings.  Scaling n by a power of two scales the quotient and
every residual exactly while they stay normal, so these two binades stand for
every quotient the kernels use (2^-76 .. 1)."""
from __future__ import annotations

import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from _common import ROOT

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* mismatches of the corrected reciprocal quotient over n in [d/2, 2d) */
long check(float d, float y) {
    long bad = 0;
    const uint32_t lo = f2u(d * 0.5f), hi = f2u(d * 2.0f);
    for (uint32_t u = lo; u < hi; u++) {
        const float n = u2f(u);
        float q = n * y;
        float r = fmaf(-d, q, n);
        q = fmaf(r, y, q);
        r = fmaf(-d, q, n);
        q = fmaf(r, y, q);
        if (f2u(q) != f2u(n / d)) bad++;
    }
    return bad;
}
"""


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("quot")
    c, so = d / "quot.c", d / "quot.so"
    c.write_text(SRC)
    flags = ["-O2", "-ffp-contract=off", "-fPIC", "-shared"]
    if os.uname().machine == "x86_64":
        flags.append("-mfma")
    subprocess.run(["gcc", *flags, "-o", str(so), str(c), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    lib.check.restype = ctypes.c_long
    lib.check.argtypes = [ctypes.c_float, ctypes.c_float]
    return lib


def _mul(*xs):
    acc = np.float32(xs[0])
    for x in xs[1:]:
        acc = np.float32(acc * np.float32(x))
    return acc


def literal_edges():
    text = open(os.path.join(ROOT, "libraw_enhanced_amd", "csrc", "defringe.hip")).read()
    pairs = set(re.findall(r"sstepc\((\d[\d.]*)f, (\d[\d.]*)f,", text))
    assert len(pairs) >= 30
    return sorted((float(a), float(b)) for a, b in pairs)


def gate_edges(strength, et, ct):
    """The parameter-dependent edges as the defringe launcher forms them."""
    sv = np.float32(math.sqrt(max(strength, 1.0)))
    sens = np.float32(min(max(sv, np.float32(1.0)), np.float32(2.5)))
    i = np.float32(np.float32(1.0) / sens)
    mc = np.float32(np.float32(0.02) * i)
    rel = [(0.55, 1.35), (0.24, 0.52), (0.20, 0.90), (0.06, 0.22), (0.035, 0.16), (0.05, 0.20)]
    out = [(_mul(a, i), _mul(b, i)) for a, b in rel]
    out += [(_mul(et, a, i), _mul(et, b, i)) for a, b in [(0.35, 1.50), (0.10, 0.80)]]
    out += [(_mul(ct, a, i), _mul(ct, b, i)) for a, b in
            [(0.45, 1.60), (0.22, 0.95), (0.12, 0.65), (0.10, 0.55), (0.35, 1.20), (1.20, 2.40),
             (0.80, 2.50), (0.55, 1.65)]]
    out.append((mc, np.float32(mc * np.float32(5.0))))
    return out


def _width_and_recip(e0, e1):
    d = np.float32(np.float32(e1) - np.float32(e0))
    return d, np.float32(np.float32(1.0) / d)


def test_literal_edges_give_ieee_quotients(checker):
    for e0, e1 in literal_edges():
        d, y = _width_and_recip(e0, e1)
        assert checker.check(float(d), float(y)) == 0, (e0, e1)


@pytest.mark.parametrize("strength,et,ct", [(10.0, 0.1, 0.15), (2.0, 0.1, 0.15), (4.0, 0.25, 0.3)])
def test_parameter_edges_give_ieee_quotients(checker, strength, et, ct):
    for e0, e1 in gate_edges(strength, et, ct):
        d, y = _width_and_recip(e0, e1)
        assert 2.0 ** -20 <= d <= 2.0 ** 100
        assert checker.check(float(d), float(y)) == 0, (e0, e1)


def test_checker_catches_a_plain_reciprocal_product(checker, tmp_path):
    """A single product n * RN(1/d) is not the IEEE quotient for a typical d:
    the residual corrections are what make the sequence exact."""
    src = SRC.replace("float r = fmaf(-d, q, n);", "float r = 0.0f;").replace(
        "r = fmaf(-d, q, n);\n        q = fmaf(r, y, q);\n        if", "if")
    c, so = tmp_path / "plain.c", tmp_path / "plain.so"
    c.write_text(src)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(so), str(c), "-lm"],
                   check=True)
    lib = ctypes.CDLL(str(so))
    lib.check.restype = ctypes.c_long
    lib.check.argtypes = [ctypes.c_float, ctypes.c_float]
    d, y = _width_and_recip(0.010, 0.060)
    assert lib.check(float(d), float(y)) > 0
