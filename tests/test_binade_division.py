"""The exact constant division behind the binade weight tables.

``div_const`` (csrc/binade.hpp) computes x / d for the Lagrange weights'
divisors d in {3, 6, 12, 20, 24, 30, 120} as q = RN(x y), y = RN(1/d),
r = fma(-q, d, x), RN(q + r y) instead of a division.  It must round exactly
as IEEE division for every numerator the tables produce (alpha^k times 1, 2,
5 or 7 with alpha = 0 or a multiple of 2^-52 in [0, 1)), and does for every
normal numerator with a normal quotient.  This test runs the same three
operations in C (gcc, with the C library's correctly rounded fma) on those
numerators and on random normal doubles across the exponent range, and
compares them with the host's division bit for bit.  It also checks that a
subnormal numerator can differ, the reason the tables use the method only
for departure points with |t| >= 1.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

C_SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double unit(void) { return (double)(xr() >> 11) * 0x1p-53; }
static double qd(double x, double d, double y) { double q = x * y; double r = fma(-q, d, x); return fma(r, y, q); }
static int same(double a, double b) { return memcmp(&a, &b, 8) == 0; }
int main(int argc, char** argv) {
    long n = atol(argv[1]);
    const double ds[] = {3, 6, 12, 20, 24, 30, 120};
    long bad = 0, sub_diff = 0;
    for (int k = 0; k < 7; ++k) {
        const double d = ds[k], y = 1.0 / d;
        for (long i = 0; i < n; ++i) {
            double x;
            if (i % 3 != 2) {
                /* binade numerators: alpha a multiple of 2^-52 (or 2^-52 * 2^-e) */
                int e = (int)(xr() % 53);
                double a = ldexp(floor(unit() * 0x1p52), -52 - (int)(xr() % 8)) ;
                if (i % 5 == 0) a = ldexp(1.0, -52 + e) * (double)(xr() % 7 + 1);
                if (a >= 1.0) a = a - floor(a);
                double p2 = a * a, p3 = p2 * a, p4 = p3 * a, p5 = p4 * a;
                const double m[] = {1.0, 2.0, 5.0, 7.0};
                double pw[] = {a, p2, p3, p4, p5};
                x = pw[xr() % 5] * m[xr() % 4];
            } else {
                /* random normal doubles, exponents 2^-1000 .. 2^1000 */
                x = ldexp(0.5 + 0.5 * unit(), (int)(xr() % 2001) - 1000);
            }
            if (!same(x / d, qd(x, d, y))) ++bad;
        }
        for (long i = 0; i < n / 10; ++i) { /* subnormal numerators */
            uint64_t b = xr() & 0x000FFFFFFFFFFFFFULL;
            double x;
            memcpy(&x, &b, 8);
            if (!same(x / d, qd(x, d, y))) ++sub_diff;
        }
    }
    printf("%ld %ld\n", bad, sub_diff);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_fma_division_rounds_like_ieee_division():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "div.c")
        exe = os.path.join(tmp, "div")
        with open(src, "w") as f:
            f.write(C_SRC.replace("#include <string.h>", "#include <string.h>\n#include <stdlib.h>"))
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe, "3000000"], check=True, capture_output=True, text=True).stdout.split()
    bad, sub_diff = int(out[0]), int(out[1])
    assert bad == 0
    assert sub_diff > 0  # the method is not exact there: the tables never feed it subnormals
