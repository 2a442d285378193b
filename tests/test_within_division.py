"""The within kernels divide by degree_per_kilometer_longitude / _latitude
(src/include/util.h:72-96) through precomputed reciprocals: q = a y with
y = RN(1 / b), then two fma corrections q + (a - b q) y.  The result must be
the correctly rounded a / b (the reference divides), or distances and
decisions would change in the last bit.  This test restates the kernel's
sequence in C (fma from libm, IEEE binary64) and compares it with the
division on every entry of the reference's 900-value table and the latitude
constant, for a spread of dividends: f32 differences promoted to double
(every binade the coordinates produce), random doubles, arbitrary bit
patterns (subnormal, huge and non-finite dividends take the division).
The raster steps the hierarchy's pixel windows divide by go through the
same sequence and are checked the same way."""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
static double div_rcp(double a, double b, double y) {
    double q = a * y;
    q = fma(fma(-b, q, a), y, q);
    q = fma(fma(-b, q, a), y, q);
    const double m = fabs(a);
    if (!(m < 0x1p1000) || (m < 0x1p-900 && m != 0.0)) q = a / b;
    return q;
}
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
int main(int argc, char **argv) {
    int nb = argc - 1;
    long bad = 0, n = 0;
    for (int t = 0; t < nb; t++) {
        double b = strtod(argv[t + 1], NULL);
        volatile double one = 1.0;
        double y = one / b;
        for (int k = 0; k < 40000; k++) {
            double a;
            switch (k & 3) {
            case 0: { /* a float difference promoted */
                float f = (float)ldexp(1.0 + (xr() >> 11) * 0x1p-53, (int)(xr() % 70) - 60);
                a = (double)f; break; }
            case 1: a = ldexp(1.0 + (xr() >> 11) * 0x1p-53, (int)(xr() % 200) - 100); break;
            case 2: { /* any bit pattern: all binades, subnormals (they take the division), inf, nan */
                      uint64_t bits = xr(); double v;
                      __builtin_memcpy(&v, &bits, 8); a = v; break; }
            default: { float f = (float)(xr() % 4000) * 0.01f - (float)(xr() % 4000) * 0.01f; a = (double)f; }
            }
            if (xr() & 1) a = -a;
            double q = a / b;
            double g = div_rcp(a, b, y);
            if (!(g == q) && !(isnan(g) && isnan(q))) bad++;
            n++;
        }
    }
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


def test_reciprocal_division_is_the_rounded_quotient(tmp_path):
    src = tmp_path / "divr.c"
    src.write_text(SRC)
    exe = tmp_path / "divr"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    table = json.load(open(os.path.join(ROOT, "tests", "golden", "km_table.json")))["values"]
    divisors = [repr(float(v)) for v in table] + [repr(360.0 / 40076.0)]
    # raster steps (the within hierarchy's pixel windows divide by them):
    # 1.1e-5 * 2^k grid steps and arbitrary ones across 2^-20 .. 2^10
    rng = np.random.default_rng(5)
    steps = [1.1e-5 * 2.0 ** k for k in range(20)] + list(2.0 ** rng.uniform(-20, 10, 300))
    divisors += [repr(float(v)) for v in steps]
    out = subprocess.run([str(exe)] + divisors, capture_output=True, text=True, check=True).stdout.split()
    n, bad = int(out[0]), int(out[1])
    assert n == 40000 * len(divisors)
    assert bad == 0


def test_a_wrong_reciprocal_is_caught(tmp_path):
    # the same check fails for a reciprocal off by one ulp with no correction
    # step (a planted mistake), so the comparison above has teeth
    src = tmp_path / "divw.c"
    src.write_text(SRC.replace("q = fma(fma(-b, q, a), y, q);\n    q = fma(fma(-b, q, a), y, q);", "")
                   .replace("double y = one / b;", "double y = nextafter(one / b, 0.0);"))
    exe = tmp_path / "divw"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe), repr(360.0 / 40076.0), "0.0123"], capture_output=True, text=True,
                         check=True).stdout.split()
    assert int(out[1]) > 0
