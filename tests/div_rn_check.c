/* Exhaustive check of the reciprocal used by the stencil kernels' shared
 * division (rcp_rn in liquidsimulation_amd/csrc/lps_div.h): for every float d
 * in [2^-78, 2^40) and every estimate y0 within one ulp of 1/d (the hardware
 * reciprocal's documented accuracy), one fused Newton step
 *     e = 1 - d*y0,  y = y0 + e*y0
 * must give the correctly rounded 1.0f/d, except when d's significand is all
 * ones and y0 is the lower candidate (counted apart; the device program
 * div_rn_check.hip shows the hardware estimate is never that one).  Also draws
 * random quotients a/d
 * (a and d in [1, 2)) and checks the Markstein correction built on that
 * reciprocal against the IEEE quotient; the device program div_rn_check.hip
 * covers every such pair.  The uncorrected product a * (1/d) must miss some of
 * the random quotients (the check can fail).  Exit status 0 when no case fails.
 *
 * With arguments DIV... (table steps, e.g. 0x1.a36e2ep-17) it checks instead
 * the pressure passes' table index (table_index_q in lps_kernels.hip): for
 * every float d in [0, 1001 * DIV), the truncated corrected quotient
 * trunc(markstein(d, DIV, 1.0f / DIV)) equals trunc(d / DIV); the uncorrected
 * trunc(d * (1.0f / DIV)) must miss some (the check can fail). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float newton(float d, float y0) {
    const float e = fmaf(-d, y0, 1.f);
    return fmaf(e, y0, y0);
}

static float markstein(float a, float d, float y) {
    const float q = a * y;
    const float r = fmaf(-q, d, a);
    return fmaf(r, y, q);
}

static uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int check_index(float div) {
    const float y = 1.f / div, dmax = 1001.f * div;
    const long u1 = f2u(dmax);
    long bad = 0, plain = 0;
#pragma omp parallel for reduction(+ : bad, plain) schedule(static)
    for (long u = 0; u < u1; u++) {
        const float d = u2f((uint32_t)u);
        const uint32_t k = (uint32_t)(d / div);
        if ((uint32_t)markstein(d, div, y) != k) bad++;
        if ((uint32_t)(d * y) != k) plain++;
    }
    printf("div %a: %ld distances checked, %ld wrong (%ld wrong uncorrected)\n", div, u1, bad, plain);
    return bad != 0 || plain == 0;
}

int main(int argc, char** argv) {
    if (argc > 1) {
        int rc = 0;
        for (int k = 1; k < argc; k++) rc |= check_index(strtof(argv[k], NULL));
        return rc;
    }
    const long u0 = f2u(0x1p-78f), u1 = f2u(0x1p40f);
    long bad = 0, n = 0, ones = 0;
#pragma omp parallel for reduction(+ : bad, n, ones) schedule(static)
    for (long u = u0; u < u1; u++) {
        const float d = u2f((uint32_t)u);
        const float rn = 1.f / d;
        const double t = 1.0 / (double)d;
        for (int k = -1; k <= 1; k++) {
            const float y0 = u2f(f2u(rn) + (uint32_t)k);
            const double up = (double)u2f(f2u(y0) + 1u) - (double)y0, dn = (double)y0 - (double)u2f(f2u(y0) - 1u);
            if (fabs((double)y0 - t) > (up > dn ? up : dn)) continue; /* not within one ulp */
            n++;
            if (newton(d, y0) != rn) {
                if ((u & 0x7fffff) == 0x7fffff)
                    ones++;
                else
                    bad++;
            }
        }
    }
    long qbad = 0, plain = 0;
    const long qn = 1L << 30;
#pragma omp parallel for reduction(+ : qbad, plain) schedule(static)
    for (long k = 0; k < qn; k++) {
        uint64_t s = (uint64_t)k * 0x2545F4914F6CDD1Dull;
        const uint64_t r = splitmix(&s);
        const float a = u2f(0x3f800000u | (uint32_t)(r & 0x7fffffu));
        const float d = u2f(0x3f800000u | (uint32_t)((r >> 23) & 0x7fffffu));
        if (markstein(a, d, 1.f / d) != a / d) qbad++;
        if (a * (1.f / d) != a / d) plain++;
    }
    printf("%ld reciprocal estimates checked, %ld wrong (%ld with an all-ones significand); "
           "%ld random quotients, %ld wrong (%ld wrong uncorrected)\n", n, bad, ones, qn, qbad, plain);
    return bad != 0 || qbad != 0 || plain == 0;
}
