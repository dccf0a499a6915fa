/* Exhaustive check of the square-root correction used by the stencil kernels
 * (sqrt_rn in liquidsimulation_amd/csrc/lps_kernels.hip): for every float s in
 * [2^-96, 2^64) and every estimate r within one ulp of sqrt(s), choosing lo
 * when s - lo*r <= 0, hi when s - hi*r > 0 (fused), else r, must give the
 * correctly rounded sqrtf(s).  Exit status 0 when no case fails. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float correct(float r, float s) {
    const float lo = u2f(f2u(r) - 1u), hi = u2f(f2u(r) + 1u);
    if (fmaf(-lo, r, s) <= 0.f) return lo;
    if (fmaf(-hi, r, s) > 0.f) return hi;
    return r;
}

int main(void) {
    const long u0 = f2u(0x1p-96f), u1 = f2u(0x1p64f);
    long bad = 0, n = 0;
#pragma omp parallel for reduction(+ : bad, n) schedule(static)
    for (long u = u0; u < u1; u++) {
        const float s = u2f((uint32_t)u);
        const float rn = sqrtf(s);
        const double t = sqrt((double)s);
        for (int k = -1; k <= 1; k++) {
            const float r = u2f(f2u(rn) + (uint32_t)k);
            const double up = (double)u2f(f2u(r) + 1u) - (double)r, dn = (double)r - (double)u2f(f2u(r) - 1u);
            if (fabs((double)r - t) > (up > dn ? up : dn)) continue;  /* not within one ulp */
            n++;
            if (correct(r, s) != rn) bad++;
        }
    }
    printf("%ld estimates checked, %ld wrong\n", n, bad);
    return bad != 0;
}
