"""The glibc-exact logf/sinf/expf of nemo_amd/csrc/glibc_math.h against the
host C library, over every float the callers can produce:

  logf(1 - b), b = float(u32) * 2^-32       (Box-Muller radius, RNG.cpp:30)
  sinf(a),     a = float(u32) * 1.4629181e-9 (Box-Muller angle, RNG.cpp:28)
  fmodf(x, y)  (exact by definition) against the C library's, mostly for
               y = float(2 pi), Kuramoto's phase wrap (Kuramoto.cpp:124)
  expf(x),     every float (HH.cpp:92-98 calls it on the membrane
               potential's functions; the special cases are selected, not
               branched, so they are checked too)
  x / c        HH's exponent divisions by 20, 80, 18 and 1 as a reciprocal
               product with one FMA correction (div_by_const), every float

The same header is compiled for the host here; the GPU tests run it on the
device against the same host results.
"""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "glibc_math.h"
using namespace nemo_math;
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
int main() {
    long bad_log = 0, bad_sin = 0, bad_exp = 0, n_log = 0, n_sin = 0, n_exp = 0;
    /* every distinct float(u32): exact below 2^24, then 2^23 values per binade */
    for (uint64_t u = 0; u < (1ull << 32); ) {
        float f = (float)(uint32_t)u;
        float b = f * 0.00000000023283064365386962890625f;
        float x = 1 - b;
        if (x > 0) { n_log++; if (bits(logf(x)) != bits(logf_exact(x))) bad_log++; }
        float a = f * 1.4629180792671596810513378043098e-9f;
        n_sin++; if (bits(sinf(a)) != bits(sinf_exact(a))) bad_sin++;
        uint64_t step = u < (1ull << 24) ? 1 : (1ull << (64 - __builtin_clzll(u) - 24));
        u += step;
    }
    /* every float */
    #pragma omp parallel for reduction(+:bad_exp, n_exp) schedule(static, 1 << 20)
    for (int64_t u = 0; u < (int64_t(1) << 32); u += 1) {
        uint32_t w = (uint32_t)u;
        float x; memcpy(&x, &w, 4);
        n_exp++; if (bits(expf(x)) != bits(expf_exact(x))) bad_exp++;
    }
    /* fmodf: by 2 pi rounded to float (Kuramoto's phase wrap) every float
     * below 1024 in magnitude, every 97th float beyond; other divisors,
     * subnormal ones included, every 4099th float */
    long bad_fmod = 0, n_fmod = 0;
    const float y0 = (float)(2.0f * M_PI);
    #pragma omp parallel for reduction(+:bad_fmod, n_fmod) schedule(dynamic, 1 << 16)
    for (int64_t u = 0; u < (int64_t(1) << 32); u += 1) {
        uint32_t w = (uint32_t)u;
        if ((w & 0x7fffffffu) >= 0x44800000u && u % 97) continue;
        float x; memcpy(&x, &w, 4);
        float a = fmodf(x, y0), b = fmodf_exact(x, y0);
        n_fmod++;
        if (bits(a) != bits(b) && !(a != a && b != b)) bad_fmod++;
    }
    const float ys[] = {1.0f, 3.0f, 1e-40f, -2.5f, 1e30f, 0.0f};
    for (float y : ys) {
        #pragma omp parallel for reduction(+:bad_fmod, n_fmod) schedule(dynamic, 1 << 12)
        for (int64_t u = 0; u < (int64_t(1) << 32); u += 4099) {
            uint32_t w = (uint32_t)u;
            float x; memcpy(&x, &w, 4);
            float a = fmodf(x, y), b = fmodf_exact(x, y);
            n_fmod++;
            if (bits(a) != bits(b) && !(a != a && b != b)) bad_fmod++;
        }
    }
    /* div_by_const against the correctly rounded division, every float x,
     * for HH's divisors: quotients must agree wherever |x| >= 2^-122, and
     * expf must agree on the two quotients everywhere */
    long div_far = 0, div_exp = 0, div_near = 0;
    const float cs[] = {1.0f, 20.0f, 80.0f, 18.0f};
    for (float c : cs) {
        const float inv = 1.0f / c;
        #pragma omp parallel for reduction(+:div_far, div_exp, div_near) schedule(static, 1 << 20)
        for (int64_t u = 0; u < (int64_t(1) << 32); u += 1) {
            uint32_t w = (uint32_t)u;
            float x; memcpy(&x, &w, 4);
            if (x != x) continue;
            const float a = x / c, b = div_by_const(x, c, inv);
            if (bits(a) == bits(b)) continue;
            if (fabsf(x) >= 0x1p-122f) div_far++; else div_near++;
            if (bits(expf(a)) != bits(expf(b))) div_exp++;
        }
    }
    printf("%ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld\n", bad_log, n_log, bad_sin, n_sin, bad_exp, n_exp, bad_fmod, n_fmod,
           div_far, div_exp, div_near);
    return 0;
}
"""


@pytest.fixture(scope="module")
def results():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "m.cpp")
        exe = os.path.join(d, "m")
        open(src, "w").write(PROGRAM)
        subprocess.run(["g++", "-O2", "-fopenmp", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "nemo_amd", "csrc"),
                        src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True, timeout=600).stdout
    return [int(x) for x in out.split()]


def test_logf_matches_host_over_box_muller_domain(results):
    bad, n = results[0], results[1]
    assert n > 1 << 24
    assert bad == 0


def test_sinf_matches_host_over_box_muller_domain(results):
    bad, n = results[2], results[3]
    assert n > 1 << 26
    assert bad == 0


def test_expf_matches_host_over_every_float(results):
    bad, n = results[4], results[5]
    assert n == 1 << 32
    assert bad == 0


def test_division_by_hh_constants_matches_for_expf(results):
    """HH's exponent divisions (x / 20, / 80, / 18, / 1) by reciprocal and one
    FMA correction: the correctly rounded quotient wherever |x| >= 2^-122,
    and the same expf everywhere.  Below 2^-122 the quotients do differ
    (div_near > 0), so the comparison is not vacuous."""
    div_far, div_exp, div_near = results[8], results[9], results[10]
    assert div_far == 0
    assert div_exp == 0
    assert div_near > 0


def test_fmodf_matches_host(results):
    bad, n = results[6], results[7]
    assert n > 1 << 31
    assert bad == 0
