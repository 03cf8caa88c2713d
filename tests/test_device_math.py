"""The compact FP64 elementary functions of the collision kernels (log_pos, exp_small,
sincos_turn, quot in picfoam_amd/csrc/device.h), compiled for the host with g++ from the
same source text, against long-double libm over their argument ranges: log within 2.5 ulp
on (1e-174, 1], exp within 1.5 ulp on [-352, 0], sin/cos of 2 pi u within 2e-16 absolute,
the quotient within 1 ulp even from a coarse reciprocal estimate."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "picfoam_amd", "csrc", "device.h")

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include <algorithm>
#define __device__
#define __forceinline__ inline
// host stand-ins of the AMDGPU builtins; the reciprocal estimate is deliberately coarse
static double __builtin_amdgcn_rcp(double d) { return (double)(1.0f / (float)d); }
static double __builtin_amdgcn_frexp_mant(double x) { int e; return std::frexp(x, &e); }
static int __builtin_amdgcn_frexp_exp(double x) { int e; std::frexp(x, &e); return e; }
static double __builtin_amdgcn_ldexp(double x, int n) { return std::ldexp(x, n); }
using std::fma; using std::rint;
@FUNCTIONS@
static double ulp(double x) { x = std::fabs(x); return std::nextafter(x, INFINITY) - x; }
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(0, 1);
    double el = 0, ee = 0, es = 0, eq = 0;
    const long double twopi = 6.283185307179586476925286766559L;
    for (int i = 0; i < 400000; ++i) {
        double x = std::exp(-400 * U(g));
        if (i % 3 == 0) x = 1e-4 + U(g);
        if (i % 5 == 0) x = 1.0 - 1e-12 * U(g);
        const long double tl = std::log((long double)x);
        if (tl != 0) el = std::max(el, (double)(std::fabs((long double)log_pos(x) - tl) / ulp((double)tl)));
        const double y = (i % 4 == 0) ? -6 * U(g) : -352 * U(g);
        const long double te = std::exp((long double)y);
        ee = std::max(ee, (double)(std::fabs((long double)exp_small(y) - te) / ulp((double)te)));
        const double u = (i % 7 == 0) ? (i % 4) * 0.25 : U(g);
        double s, c;
        sincos_turn(u, &s, &c);
        es = std::max(es, (double)std::fabs((long double)s - std::sin(twopi * u)));
        es = std::max(es, (double)std::fabs((long double)c - std::cos(twopi * u)));
        const double n = U(g) * 10 - 5, d = 0.5 + U(g) * 3;
        const long double tq = (long double)n / d;
        if (tq != 0) eq = std::max(eq, (double)(std::fabs((long double)quot(n, d) - tq) / ulp((double)tq)));
    }
    printf("%.6g %.6g %.6g %.6g %.17g\n", el, ee, es, eq, log_pos(1.0));
    return 0;
}
"""


def _functions():
    text = open(SRC).read()
    a = text.index("constexpr double kLn2Hi")
    b = text.index("// Philox4x32-10")
    # the double overloads only (the float ones call the library)
    return "\n".join(line for line in text[a:b].splitlines() if "float" not in line)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_compact_fp64_functions(tmp_path):
    src = tmp_path / "m.cpp"
    src.write_text(HARNESS.replace("@FUNCTIONS@", _functions()))
    exe = tmp_path / "m"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    el, ee, es, eq, l1 = (float(v) for v in out)
    assert el < 2.5, el
    assert ee < 1.5, ee
    assert es < 2e-16, es
    assert eq <= 1.0, eq
    assert l1 == 0.0
