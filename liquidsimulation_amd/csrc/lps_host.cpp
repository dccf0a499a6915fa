// lps_host.cpp -- host-side derived quantities of the solver parameters.
//
// Kernel radius, lookup tables, particle mass and the PCISPH delta are set by
// the reference on the host whenever a parameter changes:
//   SPHSolver::recalcKernelRadius   LiquidPhysics/src/SPHSolver.cpp:509-523
//   SPHKernel::recalcCachedWeights  LiquidPhysics/src/Kernels/SPHKernel.cpp:113-137
//   poly6 / spiky / viscosity       Kernels/DefaultKernel.cpp:12-42,
//                                   Kernels/PressureKernel.cpp:12-42,
//                                   Kernels/ViscosityKernel.cpp:12-40
//   SPHSolver::recalcParticleMass   SPHSolver.cpp:480-507
//   PCISPHSolver::calcDelta         PCISPHSolver.cpp:286-328
// The tables are a parity contract (every kernel value is table[int(d/step)]),
// so they are produced here with the reference's float arithmetic: the kernel
// coefficients are evaluated in double and rounded to float, powers of the
// radius are accumulated in float.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "lps_common.h"

namespace {

constexpr int kSubdivision = LPS_TABLE_SIZE - 1;

float table_lookup(const float* tab, float h, float step, float d) {
    if (d >= h) return 0.f;
    return tab[(int)(d / step)];
}

// The 7x7x7 lattice of step 1.6 r_p used by both the mass rule and delta.
template <class F>
void for_lattice(float particle_radius, F&& fn) {
    const float step = 1.6f * particle_radius;
    for (int a = -3; a <= 3; a++)
        for (int b = -3; b <= 3; b++)
            for (int c = -3; c <= 3; c++) fn((float)a * step, (float)b * step, (float)c * step);
}

}  // namespace

void lps_compute_derived(const lps_params* p, Derived* d) {
    const float r = p->kernel_radius_factor * p->particle_radius;
    // powers as SPHKernel::setRadius accumulates them
    float pw[10];
    pw[1] = r;
    for (int k = 2; k <= 6; k++) pw[k] = pw[k - 1] * r;
    pw[9] = pw[6] * pw[3];
    const float r2 = pw[2];
    const float step = r / (float)kSubdivision;
    const double pi = M_PI;
    const float c_poly = (float)(315.0 / (64.0 * pi * (double)pw[9]));
    const float c_poly_d = (float)(-945.0 / (32.0 * pi * (double)pw[9]));
    const float c_spiky_d = (float)(-45.0 / (pi * (double)pw[6]));
    const float c_visc_dd = (float)(45.0 / (pi * (double)pw[6]));
    for (int i = 0; i <= kSubdivision; i++) {
        const float x = (float)i * step;
        const float q = r2 - x * x;
        d->tables[0][i] = c_poly * q * q * q;
        d->tables[1][i] = c_poly_d * q * q;
        d->tables[2][i] = c_poly_d * (r2 - x * x) * (3.f * r2 - 7.f * x * x);
        const float u = r - x;
        d->tables[3][i] = c_spiky_d * u * u;
        d->tables[4][i] = c_visc_dd * (r - x);
    }
    d->h = r;
    d->div = step;
    if (p->mass_override > 0.f) {
        d->mass = p->mass_override;
    } else {
        float sum = 0.f;
        for_lattice(p->particle_radius, [&](float x, float y, float z) {
            sum += table_lookup(d->tables[0], r, step, std::sqrt(x * x + y * y + z * z));
        });
        d->mass = p->rest_density / sum;
    }
    // s_max: largest float s with sqrtf(s) < h.  sqrtf is correctly rounded and
    // monotone, so "sqrtf(s) < h" is equivalent to "s <= s_max" for every float s
    // (NaN compares false on both sides).  The stencils test the squared
    // distance first and take the square root only for accepted pairs.
    float s = r * r;
    while (std::sqrt(s) >= r) s = std::nextafter(s, 0.f);
    while (std::sqrt(std::nextafter(s, INFINITY)) < r) s = std::nextafter(s, INFINITY);
    d->s_max = s;
    // a squared distance evaluated with fused multiply-adds differs from the
    // unfused one by a few units in the last place; 2^-18 covers that
    d->s_relax = std::nextafter(s * (1.f + 1.f / 262144.f), INFINITY);
    d->inv_div = 1.f / step;
    lps_index_thresholds(r, step, d->thr);
}

// The lookup index of the OpenCL kernels, min(trunc(d / step), 1000) with
// d = sqrtf(s) (e.g. SPHKernels.cl:237), as a function of the squared distance s.
static uint32_t lookup_index(float s, float step) {
    const float q = std::sqrt(s) / step;
    if (!(q < (float)LPS_TABLE_SIZE)) return LPS_TABLE_SIZE - 1;
    return std::min((uint32_t)std::trunc(q), (uint32_t)(LPS_TABLE_SIZE - 1));
}

void lps_index_thresholds(float h, float div, float* thr) {
    // the index is monotone in s (sqrtf and the division round to nearest and
    // are monotone), and non-negative floats order like their bit patterns:
    // binary search per k over the bit patterns up to just beyond h^2
    uint32_t hi_bits;
    const float s_hi = std::nextafter(h * h * 1.01f, INFINITY);
    std::memcpy(&hi_bits, &s_hi, sizeof(float));
    auto at = [](uint32_t b) {
        float f;
        std::memcpy(&f, &b, sizeof(float));
        return f;
    };
    thr[0] = 0.f;
    for (uint32_t k = 1; k < (uint32_t)LPS_TABLE_SIZE; k++) {
        if (lookup_index(s_hi, div) < k) {
            thr[k] = INFINITY;
            continue;
        }
        uint32_t lo = 0, hi = hi_bits;  // index(at(lo)) < k <= index(at(hi)) once lo is excluded
        if (lookup_index(0.f, div) >= k) {
            thr[k] = 0.f;
            continue;
        }
        while (hi - lo > 1u) {
            const uint32_t mid = lo + (hi - lo) / 2u;
            if (lookup_index(at(mid), div) >= k) hi = mid;
            else lo = mid;
        }
        thr[k] = at(hi);
    }
    thr[LPS_TABLE_SIZE] = INFINITY;
}

float lps_compute_delta(const lps_params* p, const Derived* d, float dt) {
    float gx = 0.f, gy = 0.f, gz = 0.f, gg = 0.f;
    for_lattice(p->particle_radius, [&](float x, float y, float z) {
        const float dist = std::sqrt(x * x + y * y + z * z);
        if (dist > 0.f) {
            const float w = table_lookup(d->tables[3], d->h, d->div, dist);
            const float ux = (x / dist) * w, uy = (y / dist) * w, uz = (z / dist) * w;
            gx += ux;
            gy += uy;
            gz += uz;
            gg += ux * ux + uy * uy + uz * uz;
        }
    });
    const float denom = -(gx * gx + gy * gy + gz * gz) - gg;
    float beta = dt * d->mass / p->rest_density;
    beta = 2.f * beta * beta;
    return (std::fabs(denom) > 0.f) ? -1.f / (beta * denom) : 0.f;
}
