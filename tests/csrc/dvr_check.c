/* The step kernels divide by constant and per-row divisors through their
 * precomputed reciprocal (dvr in swm_amd/csrc/swm_kernels.h): q0 = a*y,
 * two remainder/correction steps with FMAs, q0 itself for zero and infinite
 * quotients.  This program restates that sequence with the C library's fma
 * (correctly rounded, like the GPU's) and compares it bit for bit with IEEE
 * a / b on random operands, on quotients next to rounding midpoints, on
 * divisors whose significands are nearly all ones, on negative divisors and
 * on signed zeros and infinities.  Usage: dvr_check N; exit status 0 when
 * every case agrees. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st[2] = {0x9E3779B97F4A7C15ull, 0xD1B54A32D192ED03ull};
static uint64_t next(void) {   /* xoroshiro128+ */
    const uint64_t s0 = st[0];
    uint64_t s1 = st[1];
    const uint64_t r = s0 + s1;
    s1 ^= s0;
    st[0] = ((s0 << 55) | (s0 >> 9)) ^ s1 ^ (s1 << 14);
    st[1] = (s1 << 36) | (s1 >> 28);
    return r;
}
static double rnd(int emin, int emax) {   /* random sign, significand and exponent in [emin, emax] */
    uint64_t b = ((uint64_t)(emin + (int)(next() % (uint64_t)(emax - emin + 1)) + 1023) << 52) |
                 (next() & ((1ull << 52) - 1));
    if (next() & 1) b |= 1ull << 63;
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static double dvr(double a, double b, double y) {
    const double q0 = a * y;
    double r = fma(-b, q0, a);
    const double q1 = fma(r, y, q0);
    r = fma(-b, q1, a);
    const double q2 = fma(r, y, q1);
    return (q0 == 0 || isinf(q0)) ? q0 : q2;
}
static int same(double x, double y) { return memcmp(&x, &y, 8) == 0 || (isnan(x) && isnan(y)); }

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    long bad = 0;
    for (long k = 0; k < n; ++k) {
        double b = rnd(-30, 30);
        if (k % 4 == 0) {   /* significand 1.111...1 with a few low bits cleared */
            uint64_t bb;
            memcpy(&bb, &b, 8);
            bb |= ((1ull << 52) - 1) ^ (next() & 0xff);
            memcpy(&b, &bb, 8);
        }
        const double y = 1.0 / b;
        double a;
        if (k % 3 == 0) {   /* a / b next to the midpoint m between two doubles */
            const double q = rnd(-20, 20);
            const double m = q + copysign(ldexp(1.0, ilogb(q) - 53), q);
            a = b * m;
            if (next() & 1) a = nextafter(a, (next() & 1) ? INFINITY : -INFINITY);
        } else {
            a = rnd(-60, 60);
        }
        if (!same(a / b, dvr(a, b, y))) {
            if (bad < 5) printf("a=%a b=%a: %a != %a\n", a, b, a / b, dvr(a, b, y));
            ++bad;
        }
    }
    const double sp[4] = {0.0, -0.0, INFINITY, -INFINITY};
    for (int z = 0; z < 8; ++z) {
        const double a = sp[z % 4], b = z < 4 ? 3.7 : -3.7;
        if (!same(a / b, dvr(a, b, 1.0 / b))) {
            printf("a=%a b=%a: %a != %a\n", a, b, a / b, dvr(a, b, 1.0 / b));
            ++bad;
        }
    }
    printf("%ld cases, %ld differ\n", n + 8, bad);
    return bad != 0;
}
