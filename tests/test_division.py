"""The quotient rule of the reference-order GPU solve (gslex.hip, qdiv):
a/b as q = RN(a y), r = fma(-q, b, a), RN(q + r y) with y = RN(1/b), the
reciprocal of a divisor that does not change over the sweeps.  Markstein's
theorem makes this RN(a/b) for finite normal operands in round-to-nearest;
this test checks it against IEEE division on the CPU (glibc's fma is
correctly rounded) for random operands of many magnitudes, operands with few
decimal digits like the CLI's resistances and voltages, and divisors whose
significand is all ones -- a mistake in the rule (a dropped correction, a
sign) fails here at once."""
import os
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
static uint64_t s = 88172645463325252ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double u01(void) { return (xr() >> 11) * 0x1.0p-53; }
static double rnd(int lo, int hi) {
    double v = ldexp(1.0 + u01(), lo + (int)floor((hi - lo) * u01()));
    return (xr() & 1) ? -v : v;
}
static double qdiv(double a, double b, double y) {
    double q = a * y, r = fma(-q, b, a), q1 = fma(r, y, q);
    return a == 0.0 ? q : q1;
}
int main(int argc, char **argv) {
    long N = atol(argv[1]), bad = 0;
    for (long k = 0; k < N; ++k) {
        double a = rnd(-60, 60), b = rnd(-40, 40);
        if (k % 4 == 0) b = 1.0 + (double)(xr() % 1000000) * 1e-6;
        if (k % 4 == 1) b = (double)(1 + xr() % 5000) * 1e-3;
        if (k % 8 == 2) a = (double)(xr() % 1000) * 1e-3;
        if (k % 16 == 3) a = 0.0;
        if (qdiv(a, b, 1.0 / b) != a / b || signbit(qdiv(a, b, 1.0 / b)) != signbit(a / b)) ++bad;
    }
    for (int e = -20; e <= 20; ++e)
        for (long k = 0; k < N / 400; ++k) {
            double b = ldexp(2.0 - 0x1.0p-52, e) * (1 - (double)(k % 7) * 0x1.0p-52), a = rnd(-30, 30);
            if (qdiv(a, b, 1.0 / b) != a / b) ++bad;
        }
    printf("%ld\n", bad);
    return 0;
}
"""


def test_reciprocal_division_equals_ieee_division(tmp_path):
    src = tmp_path / "qdiv.c"
    src.write_text(SRC)
    exe = tmp_path / "qdiv"
    r = subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(src), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, check=True).stdout
    assert int(out) == 0


def test_a_wrong_rule_is_caught(tmp_path):
    """Without the fma correction the same check finds mismatches."""
    src = tmp_path / "q.c"
    src.write_text(SRC.replace("return a == 0.0 ? q : q1;", "(void)q1; return q;"))
    exe = tmp_path / "q"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(src), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "200000"], capture_output=True, text=True, check=True).stdout
    assert int(out) > 0
