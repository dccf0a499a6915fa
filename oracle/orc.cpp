// oracle/orc.cpp -- CPU oracle for the SPH/PCISPH step (TEST INFRASTRUCTURE).
//
// A restatement of the reference solver's algorithm, written for checking the
// HIP library and for timing a CPU baseline.  It is never linked into the
// product.  Every function cites the reference file:line it follows (paths
// relative to the reference root, LiquidPhysics/...).
//
// Floating-point conventions (chosen so the GPU path can be compared bit for
// bit): no FMA contraction (built with -ffp-contract=off), IEEE division and
// square root, and the evaluation orders written below.  Where the reference
// relies on OpenCL built-ins whose evaluation order is implementation-defined
// (distance, length, dot, normalize) the order used here is stated at the
// helper; that choice is "parity unpinned" against the reference's OpenCL
// compiler and is pinned only against the reference's sequential Vector3D
// operations (Vector3D.cpp:55-71), which use the same orders.
#include "orc.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kSub = 1000;  // SPHKernel subdivision, Kernels/SPHKernel.cpp:4-7

enum { ST_OK = 0, ST_INVALID = 1, ST_NONFINITE = 4, ST_GRID_OVERFLOW = 5 };

// ---------------------------------------------------------------- tables --
// Kernels/SPHKernel.cpp:93-137 (setRadius, recalcCachedWeights, recalcDivisionStep)
// Kernels/DefaultKernel.cpp:12-42, PressureKernel.cpp:12-42, ViscosityKernel.cpp:12-40
// Coefficients are evaluated in double (M_PI) and stored in the kernel's
// member type, as the reference stores them in float members.
template <class T>
struct Tables {
    T h = 0, div = 0;
    T r2 = 0;
    std::vector<T> w;    // poly6 W              (DefaultKernel::calcKernelWeight)
    std::vector<T> w1;   // poly6 first deriv    (DefaultKernel::calcFirstDerivativeWeight)
    std::vector<T> w2;   // poly6 second deriv   (DefaultKernel::calcSecondDerivativeWeight)
    std::vector<T> p1;   // spiky first deriv    (PressureKernel::calcFirstDerivativeWeight)
    std::vector<T> v2;   // viscosity laplacian  (ViscosityKernel::calcSecondDerivativeWeight)
};

template <class T>
Tables<T> make_tables(T radius) {
    Tables<T> t;
    T r = radius;
    T r2 = r * r;
    T r3 = r2 * r;
    T r4 = r3 * r;
    T r5 = r4 * r;
    T r6 = r5 * r;
    T r9 = r6 * r3;
    (void)r4; (void)r5;
    T div = r / (T)kSub;                                         // SPHKernel.cpp:134-137
    T c_w = (T)(315.0 / (64.0 * M_PI * (double)r9));              // DefaultKernel.cpp:29-32
    T c_w1 = (T)(-945.0 / (32.0 * M_PI * (double)r9));            // DefaultKernel.cpp:34-37
    T c_w2 = (T)(-945.0 / (32.0 * M_PI * (double)r9));            // DefaultKernel.cpp:39-42
    T c_p1 = (T)(-45.0 / (M_PI * (double)r6));                    // PressureKernel.cpp:34-37
    T c_v2 = (T)(45.0 / (M_PI * (double)r6));                     // ViscosityKernel.cpp:37-40
    t.h = r;
    t.div = div;
    t.r2 = r2;
    for (int i = 0; i <= kSub; i++) {                              // SPHKernel.cpp:124-131
        T d = (T)i * div;
        T tw = r2 - d * d;
        t.w.push_back(c_w * tw * tw * tw);                          // DefaultKernel.cpp:12-16
        t.w1.push_back(c_w1 * tw * tw);                             // DefaultKernel.cpp:18-22
        t.w2.push_back(c_w2 * (r2 - d * d) * ((T)3 * r2 - (T)7 * d * d));  // DefaultKernel.cpp:24-27
        T tp = r - d;
        t.p1.push_back(c_p1 * tp * tp);                             // PressureKernel.cpp:18-22
        t.v2.push_back(c_v2 * (r - d));                             // ViscosityKernel.cpp:26-29
    }
    return t;
}

template <class T>
Tables<T> cast_tables(const Tables<float>& f) {
    Tables<T> t;
    t.h = (T)f.h;
    t.div = (T)f.div;
    t.r2 = (T)f.r2;
    for (size_t i = 0; i < f.w.size(); i++) {
        t.w.push_back((T)f.w[i]);
        t.w1.push_back((T)f.w1[i]);
        t.w2.push_back((T)f.w2[i]);
        t.p1.push_back((T)f.p1[i]);
        t.v2.push_back((T)f.v2[i]);
    }
    return t;
}

// SPHKernel::get*Weight (SPHKernel.cpp:13-44): zero at or beyond the radius,
// otherwise table[int(d / step)].
template <class T>
inline T lookup(const std::vector<T>& tab, const Tables<T>& t, T d) {
    if (d >= t.h) return (T)0;
    int key = (int)(d / t.div);
    return tab[(size_t)key];
}

template <class T>
T kernel_radius(const orc_params* p) {
    // SPHSolver::recalcKernelRadius (SPHSolver.cpp:509-523): h = factor * r_p in float
    float h = p->kernel_radius_factor * p->particle_radius;
    return (T)h;
}

// SPHSolver::recalcParticleMass (SPHSolver.cpp:480-507): rest density divided
// by the poly6 sum over the 7^3 lattice of step 1.6 r_p.
template <class T>
T lattice_mass(const orc_params* p, const Tables<T>& t) {
    // The FP64 rerun takes the FP32 constants (1.6f, r_p) as exact inputs.
    T step = (T)1.6f * (T)p->particle_radius;
    T sum = 0;
    for (int i = -3; i <= 3; i++)
        for (int j = -3; j <= 3; j++)
            for (int k = -3; k <= 3; k++) {
                T x = (T)i * step, y = (T)j * step, z = (T)k * step;
                T d = std::sqrt(x * x + y * y + z * z);          // Vector3D::magnitude
                sum += lookup(t.w, t, d);
            }
    return (T)p->rest_density / sum;
}

template <class T>
T particle_mass(const orc_params* p, const Tables<T>& t) {
    if (p->mass_override > 0.f) return (T)p->mass_override;
    return lattice_mass<T>(p, t);
}

// PCISPHSolver::calcDelta (PCISPHSolver.cpp:286-328), FP32.
float delta_f32(const orc_params* p, const Tables<float>& t, float mass, float dt) {
    float step = 1.6f * p->particle_radius;
    float d1x = 0.f, d1y = 0.f, d1z = 0.f, d2 = 0.f;
    for (int i = -3; i <= 3; i++)
        for (int j = -3; j <= 3; j++)
            for (int k = -3; k <= 3; k++) {
                float x = (float)i * step, y = (float)j * step, z = (float)k * step;
                float d = std::sqrt(x * x + y * y + z * z);
                if (d > 0.f) {
                    float dx = x / d, dy = y / d, dz = z / d;
                    float w = lookup(t.p1, t, d);
                    float gx = dx * w, gy = dy * w, gz = dz * w;
                    d1x += gx; d1y += gy; d1z += gz;
                    d2 += gx * gx + gy * gy + gz * gz;
                }
            }
    float denom = -(d1x * d1x + d1y * d1y + d1z * d1z) - d2;
    float beta = dt * mass / p->rest_density;
    beta = 2.f * beta * beta;
    float delta = 0.f;
    if (std::fabs(denom) > 0.f) delta = -1.f / (beta * denom);
    return delta;
}

// ------------------------------------------------------------------ grid --
struct Grid {
    float bmin[3], bmax[3], off[3], gs;
    uint32_t gsize[3];
    uint32_t cells;
    std::vector<uint32_t> keys;   // incoming order
    std::vector<uint32_t> perm;   // sorted slot -> incoming index
    std::vector<uint32_t> skeys;  // sorted keys
    std::vector<uint32_t> lb;     // [cells + 1] lower_bound table, lb[cells] = n
    int status = ST_OK;
};

// Upper bound on the cell count accepted by the oracle (also a runaway guard).
constexpr uint64_t kMaxCells = (uint64_t)1 << 31;

// Stable LSD radix sort as the reference runs it: 4 passes of 8 bits, 256
// chunks of floor(n/256) keys (last chunk takes the remainder), per-chunk
// histograms laid out bucket-major and scanned in that order
// (SPHKernels.cl:84-146, SPHSolver.cpp:1049-1119).  Returns the permutation.
std::vector<uint32_t> radix_sort_reference(const std::vector<uint32_t>& keys, int nthreads_omp) {
    const uint32_t n = (uint32_t)keys.size();
    const uint32_t T = 256, B = 256;
    std::vector<uint32_t> k0 = keys, k1(n), v0(n), v1(n);
    std::iota(v0.begin(), v0.end(), 0u);
    std::vector<uint32_t> counts((size_t)T * B);
    (void)nthreads_omp;
    for (uint32_t pass = 0; pass < 4; pass++) {
        std::fill(counts.begin(), counts.end(), 0u);
        uint32_t ppt = n / T;
        auto range = [&](uint32_t t, uint32_t& s, uint32_t& e) {
            s = ppt * t;
            e = (t == T - 1) ? n : s + ppt;
        };
#pragma omp parallel for schedule(static)
        for (int ti = 0; ti < (int)T; ti++) {
            uint32_t s, e;
            range((uint32_t)ti, s, e);
            for (uint32_t j = s; j < e; j++) {
                uint32_t b = (k0[j] >> (8 * pass)) & 0xFFu;
                counts[(size_t)b * T + ti]++;
            }
        }
        uint32_t sum = 0;
        for (size_t i = 0; i < counts.size(); i++) {
            uint32_t c = counts[i];
            counts[i] = sum;
            sum += c;
        }
#pragma omp parallel for schedule(static)
        for (int ti = 0; ti < (int)T; ti++) {
            uint32_t s, e;
            range((uint32_t)ti, s, e);
            for (uint32_t j = s; j < e; j++) {
                uint32_t b = (k0[j] >> (8 * pass)) & 0xFFu;
                uint32_t dst = counts[(size_t)b * T + ti]++;
                k1[dst] = k0[j];
                v1[dst] = v0[j];
            }
        }
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    return v0;
}

std::vector<uint32_t> stable_sort_perm(const std::vector<uint32_t>& keys) {
    std::vector<uint32_t> perm(keys.size());
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    return perm;
}

// SPHSolver::buildParallelGrid host part (SPHSolver.cpp:1011-1040), the key
// kernel calcGridIndices (SPHKernels.cl:66-82), the sort, and the cell list
// of buildCellList (SPHKernels.cl:148-186) as a lower_bound table.
template <class PosT>
Grid build_grid(const orc_params* p, int32_t n, const PosT* x, bool use_radix) {
    Grid g;
    float minX, minY, minZ, maxX, maxY, maxZ;
    minX = minY = minZ = FLT_MAX;
    maxX = maxY = maxZ = FLT_MIN;  // as the reference initialises it (SPHSolver.cpp:1014)
    bool finite = true;
    for (int32_t i = 0; i < n; i++) {
        float px = (float)x[3 * i], py = (float)x[3 * i + 1], pz = (float)x[3 * i + 2];
        if (!std::isfinite(px) || !std::isfinite(py) || !std::isfinite(pz)) finite = false;
        if (px < minX) minX = px;
        if (py < minY) minY = py;
        if (pz < minZ) minZ = pz;
        if (px > maxX) maxX = px;
        if (py > maxY) maxY = py;
        if (pz > maxZ) maxZ = pz;
    }
    g.bmin[0] = minX; g.bmin[1] = minY; g.bmin[2] = minZ;
    g.bmax[0] = maxX; g.bmax[1] = maxY; g.bmax[2] = maxZ;
    if (!finite || n <= 0) {
        g.status = (n <= 0) ? ST_INVALID : ST_NONFINITE;
        return g;
    }
    float gs = kernel_radius<float>(p) * 1.f;  // m_gridSpacingFactor = 1 (SPHSolver.cpp:13)
    g.gs = gs;
    g.off[0] = -minX; g.off[1] = -minY; g.off[2] = -minZ;
    float ex[3] = {(maxX - minX) / gs, (maxY - minY) / gs, (maxZ - minZ) / gs};
    uint64_t cells = 1;
    for (int a = 0; a < 3; a++) {
        if (!(ex[a] < 2.0e9f)) { g.status = ST_GRID_OVERFLOW; return g; }
        g.gsize[a] = (uint32_t)((int)ex[a] + 1);
        cells *= g.gsize[a];
    }
    if (cells >= kMaxCells) { g.status = ST_GRID_OVERFLOW; return g; }
    g.cells = (uint32_t)cells;
    g.keys.resize(n);
    for (int32_t i = 0; i < n; i++) {
        float px = (float)x[3 * i], py = (float)x[3 * i + 1], pz = (float)x[3 * i + 2];
        uint32_t xg = (uint32_t)std::trunc((px + g.off[0]) / gs);
        uint32_t yg = (uint32_t)std::trunc((py + g.off[1]) / gs);
        uint32_t zg = (uint32_t)std::trunc((pz + g.off[2]) / gs);
        g.keys[i] = xg + g.gsize[0] * yg + g.gsize[0] * g.gsize[1] * zg;
    }
    g.perm = use_radix ? radix_sort_reference(g.keys, 0) : stable_sort_perm(g.keys);
    g.skeys.resize(n);
    for (int32_t i = 0; i < n; i++) g.skeys[i] = g.keys[g.perm[i]];
    g.lb.resize((size_t)g.cells + 1);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c <= (int64_t)g.cells; c++)
        g.lb[c] = (uint32_t)(std::lower_bound(g.skeys.begin(), g.skeys.end(), (uint32_t)c) - g.skeys.begin());
    return g;
}

void fill_trace_grid(const Grid& g, int32_t n, orc_trace* tr) {
    if (!tr) return;
    for (int a = 0; a < 3; a++) {
        tr->bbox_min[a] = g.bmin[a];
        tr->bbox_max[a] = g.bmax[a];
        tr->grid_offset[a] = g.off[a];
        tr->grid_size[a] = g.gsize[a];
    }
    tr->grid_spacing = g.gs;
    tr->cell_count = g.cells;
    if (g.status != ST_OK) return;
    if (tr->keys) std::memcpy(tr->keys, g.skeys.data(), sizeof(uint32_t) * n);
    if (tr->perm) std::memcpy(tr->perm, g.perm.data(), sizeof(uint32_t) * n);
    if (tr->cell_start) {
        uint32_t m = std::min(tr->cell_cap, g.cells);
        std::memcpy(tr->cell_start, g.lb.data(), sizeof(uint32_t) * m);
    }
}

// -------------------------------------------------------------- helpers --
template <class T>
struct V3 {
    T x, y, z;
};

// OpenCL distance(a, b) = length(a - b); evaluated as sqrt((dx*dx + dy*dy) + dz*dz)
// with d = a - b, the order of Vector3D::magnitude (Vector3D.cpp:55-58).
template <class T>
inline T dist(const V3<T>& a, const V3<T>& b) {
    T dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
template <class T>
inline T len3(T x, T y, T z) { return std::sqrt(x * x + y * y + z * z); }

// Pressure-force direction (x_i - x_j) / d of the OpenCL kernels
// (SPHKernels.cl:406,865).  For two distinct particles at exactly the same
// point (d == 0) the reference yields 0/0 = NaN (hazard H9; the box collider
// snaps every particle that crosses three walls onto the same corner point).
// Here such a pair gets the unit direction +y for the higher sorted index and
// -y for the lower one: antisymmetric, so momentum is still conserved, and it
// separates the pair instead of leaving it stacked.
template <class T>
inline V3<T> pair_dir(const V3<T>& xi, const V3<T>& xj, T d, int32_t i, int32_t j) {
    if (d > (T)0) return V3<T>{(xi.x - xj.x) / d, (xi.y - xj.y) / d, (xi.z - xj.z) / d};
    return V3<T>{(T)0, i > j ? (T)1 : (T)-1, (T)0};
}

struct Colliders {
    const orc_box* boxes; int32_t nbox;
    const orc_sphere* spheres; int32_t nsphere;
    const orc_plane* planes; int32_t nplane;
};

// resolveCollision (SPHKernels.cl:640-659) / StaticCollisionObject.cpp:21-38
template <class T>
inline void resolve(V3<T>& pos, V3<T>& vel, T e, T mu, V3<T> n, V3<T> point) {
    T sv = vel.x * n.x + vel.y * n.y + vel.z * n.z;
    if (sv < (T)0) {
        V3<T> vn{n.x * sv, n.y * sv, n.z * sv};
        V3<T> vt{vel.x - vn.x, vel.y - vn.y, vel.z - vn.z};
        vn.x *= -e; vn.y *= -e; vn.z *= -e;
        vt.x *= mu; vt.y *= mu; vt.z *= mu;
        vel = V3<T>{vn.x + vt.x, vn.y + vt.y, vn.z + vt.z};
    }
    pos = point;
}

// handleCollisionWithBox (SPHKernels.cl:488-598) == StaticCollisionBox.cpp:20-118
template <class T>
inline void collide_box(V3<T>& pos, V3<T>& vel, T r, T e, T mu, const orc_box& b) {
    T cx = (T)b.center[0], cy = (T)b.center[1], cz = (T)b.center[2];
    T hx = (T)b.half[0], hy = (T)b.half[1], hz = (T)b.half[2];
    T dx = std::fabs(pos.x - cx), dy = std::fabs(pos.y - cy), dz = std::fabs(pos.z - cz);
    bool hit = false;
    V3<T> n{0, 0, 0};
    V3<T> pt{0, 0, 0};
    if (!b.boundary) {
        T px = (hx - dx) + r, py = (hy - dy) + r, pz = (hz - dz) + r;
        if (px > 0 && py > 0 && pz > 0) {
            hit = true;
            pt = pos;
            if (px <= py && px <= pz) {
                n.x = (pos.x - cx >= 0) ? (T)1 : (T)-1;
                pt.x = cx + n.x * (hx + r);
            } else if (py <= px && py <= pz) {
                n.y = (pos.y - cy >= 0) ? (T)1 : (T)-1;
                pt.y = cy + n.y * (hy + r);
            } else if (pz <= py && pz <= px) {
                n.z = (pos.z - cz >= 0) ? (T)1 : (T)-1;
                pt.z = cz + n.z * (hz + r);
            }
        }
    } else {
        T px = (dx - hx) + r, py = (dy - hy) + r, pz = (dz - hz) + r;
        if (px > 0 || py > 0 || pz > 0) {
            hit = true;
            pt = pos;
            if (px > 0) {
                n.x = (pos.x - cx <= 0) ? (T)1 : (T)-1;
                pt.x = cx - n.x * (hx - r);
            }
            if (py > 0) {
                n.y = (pos.y - cy <= 0) ? (T)1 : (T)-1;
                pt.y = cy - n.y * (hy - r);
            }
            if (pz > 0) {
                n.z = (pos.z - cz <= 0) ? (T)1 : (T)-1;
                pt.z = cz - n.z * (hz - r);
            }
            // normalize as Vector3D::normalize (Vector3D.cpp:61-66): v * (1/|v|)
            T m = len3(n.x, n.y, n.z);
            if (m > 0) {
                T inv = (T)1 / m;
                n.x *= inv; n.y *= inv; n.z *= inv;
            }
        }
    }
    if (hit) resolve(pos, vel, e, mu, n, pt);
}

// handleCollisionWithSphere, OpenCL form (SPHKernels.cl:600-638): the normal is
// normalised before the contact point is formed (hazard H8: the NONE path does not).
template <class T>
inline void collide_sphere_ocl(V3<T>& pos, V3<T>& vel, T r, T e, T mu, const orc_sphere& s) {
    V3<T> c{(T)s.center[0], (T)s.center[1], (T)s.center[2]};
    T R = (T)s.radius;
    T d = dist(c, pos);
    if (!s.boundary) {
        T pen = (R + r) - d;
        if (pen > 0) {
            // a particle exactly at the centre has no normal (0/0 in the
            // reference); it is pushed along +y
            V3<T> n = d > 0 ? V3<T>{(pos.x - c.x) / d, (pos.y - c.y) / d, (pos.z - c.z) / d} : V3<T>{0, 1, 0};
            T k = R + r;
            V3<T> pt{c.x + n.x * k, c.y + n.y * k, c.z + n.z * k};
            resolve(pos, vel, e, mu, n, pt);
        }
    } else {
        T pen = d - (R - r);
        if (pen > 0) {
            V3<T> n{(c.x - pos.x) / d, (c.y - pos.y) / d, (c.z - pos.z) / d};
            T k = R - r;
            V3<T> pt{c.x - n.x * k, c.y - n.y * k, c.z - n.z * k};
            resolve(pos, vel, e, mu, n, pt);
        }
    }
}

// StaticCollisionSphere (StaticCollisionSphere.cpp:20-56), NONE form: contact
// point from the unnormalised (x - c), then the normal is normalised (H8).
template <class T>
inline void collide_sphere_seq(V3<T>& pos, V3<T>& vel, T r, T e, T mu, const orc_sphere& s) {
    V3<T> c{(T)s.center[0], (T)s.center[1], (T)s.center[2]};
    T R = (T)s.radius;
    T d = dist(c, pos);
    T pen = s.boundary ? d - (R - r) : (R + r) - d;
    if (pen > 0) {
        V3<T> n, pt;
        if (!s.boundary) {
            n = V3<T>{pos.x - c.x, pos.y - c.y, pos.z - c.z};
            T k = R + r;
            pt = V3<T>{c.x + n.x * k, c.y + n.y * k, c.z + n.z * k};
        } else {
            n = V3<T>{c.x - pos.x, c.y - pos.y, c.z - pos.z};
            T k = R - r;
            pt = V3<T>{c.x - n.x * k, c.y - n.y * k, c.z - n.z * k};
        }
        T m = len3(n.x, n.y, n.z);
        if (m > 0) {
            T inv = (T)1 / m;
            n.x *= inv; n.y *= inv; n.z *= inv;
        }
        resolve(pos, vel, e, mu, n, pt);
    }
}

// Plane collider (no reference counterpart): the fluid side is n.x >= offset.
// Penetration r - s with s = n.x - offset; the contact point moves the particle
// along n by the penetration; the response reuses resolveCollision.
template <class T>
inline void collide_plane(V3<T>& pos, V3<T>& vel, T r, T e, T mu, const orc_plane& pl) {
    V3<T> n{(T)pl.normal[0], (T)pl.normal[1], (T)pl.normal[2]};
    T s = (n.x * pos.x + n.y * pos.y + n.z * pos.z) - (T)pl.offset;
    T pen = r - s;
    if (pen > 0) {
        V3<T> pt{pos.x + n.x * pen, pos.y + n.y * pen, pos.z + n.z * pen};
        resolve(pos, vel, e, mu, n, pt);
    }
}

// OpenCL collider order: all boxes, all spheres (SPHKernels.cl:472-480), then planes.
template <class T>
inline void collide_all_ocl(V3<T>& pos, V3<T>& vel, const orc_params* p, const Colliders& cl) {
    T r = (T)p->particle_radius, e = (T)p->restitution, mu = (T)p->friction;
    for (int32_t j = 0; j < cl.nbox; j++) collide_box(pos, vel, r, e, mu, cl.boxes[j]);
    for (int32_t j = 0; j < cl.nsphere; j++) collide_sphere_ocl(pos, vel, r, e, mu, cl.spheres[j]);
    for (int32_t j = 0; j < cl.nplane; j++) collide_plane(pos, vel, r, e, mu, cl.planes[j]);
}

// ------------------------------------------------------------ ocl step --
template <class T>
struct OclState {
    int32_t n;
    std::vector<V3<T>> x, v, vh, f, fp, xp;
    std::vector<T> rho, p, rp;
    std::vector<uint8_t> first;
};

// 27-cell stencil of the OpenCL kernels (e.g. SPHKernels.cl:219-244): cells
// are visited z-outer, y, x-inner around the cell of `cellpos`, each cell's
// sorted range [cellList[c], cellList[c+1] or n) in ascending order.
template <class T, class F>
inline void for_neighbors(const Grid& g, int32_t n, const V3<T>& cellpos, F&& fn) {
    const float gs = g.gs;
    int xg = (int)std::trunc(((float)cellpos.x + g.off[0]) / gs);
    int yg = (int)std::trunc(((float)cellpos.y + g.off[1]) / gs);
    int zg = (int)std::trunc(((float)cellpos.z + g.off[2]) / gs);
    for (int z = zg - 1; z <= zg + 1; z++)
        for (int y = yg - 1; y <= yg + 1; y++)
            for (int x = xg - 1; x <= xg + 1; x++) {
                if (x >= 0 && (uint32_t)x < g.gsize[0] && y >= 0 && (uint32_t)y < g.gsize[1] &&
                    z >= 0 && (uint32_t)z < g.gsize[2]) {
                    uint32_t gi = (uint32_t)x + g.gsize[0] * (uint32_t)y +
                                  g.gsize[0] * g.gsize[1] * (uint32_t)z;
                    uint32_t end = (gi + 1 < g.cells) ? g.lb[gi + 1] : (uint32_t)n;
                    for (uint32_t j = g.lb[gi]; j < end; j++) fn((int32_t)j);
                }
            }
}

// For the FP32 cell coordinates the grid is built from FP32 positions; in the
// FP64 rerun the stencil uses the FP32 image of the (FP32-exact) input positions.
template <class T>
inline V3<T> cell_anchor(const std::vector<V3<float>>& xf, int32_t i) {
    return V3<T>{(T)xf[i].x, (T)xf[i].y, (T)xf[i].z};
}

// Test hook: 0 makes the stall exit apply the last correction (the rule's
// earlier form, kept so a test can show what the best-correction exit fixes).
static int g_best_exit = 1;
inline bool best_exit_enabled() { return g_best_exit != 0; }

// Learned warm start of a step (no reference counterpart).  `seed` is the
// network's pressure per particle in this step's sorted order (NULL: no
// network, an unbounded value).  With `pprev` (the previous step's applied
// pressure, incoming particle order) and negative pressures disabled, the
// first correction's pressure becomes max(p1, min(seed, kappa * pprev)),
// p1 the zero start's own first-correction pressure; otherwise it becomes
// the seed.
struct WarmStart {
    const float* seed = nullptr;
    const float* pprev = nullptr;
    float kappa = 0.f;
};

template <class T>
int step_ocl(const orc_params* p, const Colliders& cl, float dtf, int32_t n, T* x, T* v, T* vh,
             uint8_t* first, orc_trace* tr, int nthreads, const WarmStart* warm = nullptr) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (tr) { tr->status = ST_OK; tr->iterations = 0; }
    if (n <= 0) return ST_INVALID;
    Tables<float> tf = make_tables<float>(kernel_radius<float>(p));
    // The FP64 rerun takes the FP32 model (tables, mass) as exact inputs and
    // evaluates the step in double: deriving them in double would move the
    // lattice distances across lookup-index ties (d / step = 500 exactly for
    // the nearest lattice neighbours), a 0.2 % change of the mass that is a
    // property of the piecewise-constant tables, not of the step's arithmetic.
    Tables<T> tb = cast_tables<T>(tf);
    const T m = (T)particle_mass<float>(p, tf);
    const T dt = (T)dtf;
    const T h = tb.h, ds = tb.div;
    const T rho0 = (T)p->rest_density;

    // ---- grid (from FP32 positions) and reorder (SPHSolver.cpp:1008-1132);
    // the multi-threaded baseline sorts with the reference's radix scheme, the
    // checker with std::stable_sort (same permutation, checked in tests)
    Grid g = build_grid(p, n, x, nthreads != 1);
    fill_trace_grid(g, n, tr);
    if (g.status != ST_OK) {
        if (tr) tr->status = g.status;
        return g.status;
    }
    OclState<T> s;
    s.n = n;
    s.x.resize(n); s.v.resize(n); s.vh.resize(n); s.first.resize(n);
    std::vector<V3<float>> xf(n);
    for (int32_t i = 0; i < n; i++) {
        uint32_t k = g.perm[i];
        s.x[i] = V3<T>{x[3 * k], x[3 * k + 1], x[3 * k + 2]};
        s.v[i] = V3<T>{v[3 * k], v[3 * k + 1], v[3 * k + 2]};
        s.vh[i] = V3<T>{vh[3 * k], vh[3 * k + 1], vh[3 * k + 2]};
        s.first[i] = first[k];
        xf[i] = V3<float>{(float)s.x[i].x, (float)s.x[i].y, (float)s.x[i].z};
    }
    s.rho.assign(n, 0); s.p.assign(n, 0); s.f.assign(n, V3<T>{0, 0, 0});

    // ---- calcDensityPressure (SPHKernels.cl:190-254)
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; i++) {
        V3<T> xi = s.x[i];
        T sum = 0;
        for_neighbors(g, n, cell_anchor<T>(xf, i), [&](int32_t j) {
            T d = dist(s.x[j], xi);
            if (d < h) {
                uint32_t idx = (uint32_t)std::trunc(d / ds);
                sum += tb.w[idx];
            }
        });
        T rho = m * sum;
        T pr = (T)p->stiffness * (rho - rho0);
        if (pr < (T)0) pr *= (T)p->negative_pressure_factor;
        s.rho[i] = rho;
        s.p[i] = pr;
    }
    if (tr && tr->rho) for (int32_t i = 0; i < n; i++) ((T*)tr->rho)[i] = s.rho[i];

    // ---- accumulateNonPressureForces (SPHKernels.cl:256-349)
    const T gx = (T)p->gravity[0], gy = (T)p->gravity[1], gz = (T)p->gravity[2];
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; i++) {
        V3<T> xi = s.x[i], vi = s.v[i];
        T ri = s.rho[i];
        V3<T> F{gx * ri, gy * ri, gz * ri};
        V3<T> sn{0, 0, 0}, vf{0, 0, 0};
        T lc = 0;
        for_neighbors(g, n, cell_anchor<T>(xf, i), [&](int32_t j) {
            V3<T> xj = s.x[j], vj = s.v[j];
            T rj = s.rho[j];
            T d = dist(xj, xi);
            if (d < h) {
                V3<T> dir{xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
                uint32_t idx = (uint32_t)std::trunc(d / ds);
                T w1 = tb.w1[idx];
                sn.x += dir.x * w1 / rj;
                sn.y += dir.y * w1 / rj;
                sn.z += dir.z * w1 / rj;
                lc += tb.w2[idx] / rj;
                if (i != j) {
                    T wv = tb.v2[idx];
                    vf.x += (vj.x - vi.x) * wv / rj;
                    vf.y += (vj.y - vi.y) * wv / rj;
                    vf.z += (vj.z - vi.z) * wv / rj;
                }
            }
        });
        sn.x *= m; sn.y *= m; sn.z *= m;
        T snl = len3(sn.x, sn.y, sn.z);
        if (snl > (T)p->surface_tension_threshold) {
            T k = -(T)p->surface_tension * lc * m;
            F.x += (sn.x / snl) * k;
            F.y += (sn.y / snl) * k;
            F.z += (sn.z / snl) * k;
        }
        T mu = (T)p->viscosity;
        F.x += vf.x * mu * m;
        F.y += vf.y * mu * m;
        F.z += vf.z * mu * m;
        s.f[i] = F;
    }

    if (p->method == 0) {
        // ---- accumulatePressureForces (SPHKernels.cl:351-422)
        std::vector<V3<T>> fnew(n);
#pragma omp parallel for schedule(dynamic, 256)
        for (int32_t i = 0; i < n; i++) {
            V3<T> xi = s.x[i];
            T ri = s.rho[i], pi = s.p[i];
            T tfac = pi / (ri * ri);
            V3<T> pf{0, 0, 0};
            for_neighbors(g, n, cell_anchor<T>(xf, i), [&](int32_t j) {
                V3<T> xj = s.x[j];
                T d = dist(xj, xi);
                if (d < h) {
                    V3<T> dir = pair_dir(xi, xj, d, i, j);
                    uint32_t idx = (uint32_t)std::trunc(d / ds);
                    if (i != j) {
                        T rj = s.rho[j];
                        T sc = tfac + s.p[j] / (rj * rj);
                        T w = tb.p1[idx];
                        pf.x += dir.x * sc * w;
                        pf.y += dir.y * sc * w;
                        pf.z += dir.z * sc * w;
                    }
                }
            });
            V3<T> F = s.f[i];
            F.x += (-pf.x) * m * ri;
            F.y += (-pf.y) * m * ri;
            F.z += (-pf.z) * m * ri;
            fnew[i] = F;
        }
        s.f = fnew;
        if (tr && tr->pressure) for (int32_t i = 0; i < n; i++) ((T*)tr->pressure)[i] = s.p[i];
    } else {
        // ---- PCISPH OpenCL branch (PCISPHSolver.cpp:171-239) with the device
        // convergence rule: stop after iteration k (1-based) when k >= min and
        // max_i |e_i| / rho0 < ratio, e_i = rho*_i - rho0 scaled by npf when
        // the pressure update is negative (PCISPHSolver.cpp:113-121,152-162); at most max.
        float deltaf = delta_f32(p, tf, (float)particle_mass<float>(p, tf), dtf);
        T delta = (T)deltaf;
        if (tr) tr->delta = deltaf;
        s.xp.assign(n, V3<T>{0, 0, 0});
        s.rp.assign(n, 0);
        s.fp.assign(n, V3<T>{0, 0, 0});
        std::vector<T> pp(n, (T)0);  // pciInit (SPHKernels.cl:661-672)
        std::vector<T> err(n, (T)0);
        const T npf = (T)p->negative_pressure_factor;
        int32_t maxit = p->max_iterations > 0 ? p->max_iterations : 1;
        int32_t minit = p->min_iterations;
        int32_t k = 0;
        bool nonfinite = false;
        // stall rule and its exit (no reference counterpart): from correction
        // 2 on, the correction kb with the smallest maximum error is
        // remembered together with what its predictions came from (the
        // pressure force, pressure and predicted density of correction
        // kb - 1).  A loop stopped by the stall rule applies those; every
        // other stop applies the last correction's, as the reference does.
        T best = 0;
        int32_t kb = 0;
        std::vector<V3<T>> fp_best;
        std::vector<V3<T>> fp_prev;
        std::vector<T> p_best, rp_best, p_prev, rp_prev;
        const bool keep_best = p->stall_iterations > 0 && best_exit_enabled();
        int32_t stall = 0, converged = 0, stalled = 0;
        for (k = 1; k <= maxit; k++) {
            // pciIntegrate + pciHandleCollisions (SPHKernels.cl:676-735)
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; i++) {
                V3<T> hv = s.vh[i];
                V3<T> pos = s.x[i];
                T ri = s.rho[i];
                V3<T> a{(s.f[i].x + s.fp[i].x) / ri, (s.f[i].y + s.fp[i].y) / ri,
                        (s.f[i].z + s.fp[i].z) / ri};
                if (s.first[i]) {
                    hv.x = s.v[i].x - a.x * dt / (T)2;
                    hv.y = s.v[i].y - a.y * dt / (T)2;
                    hv.z = s.v[i].z - a.z * dt / (T)2;
                }
                hv.x += a.x * dt; hv.y += a.y * dt; hv.z += a.z * dt;
                pos.x += hv.x * dt; pos.y += hv.y * dt; pos.z += hv.z * dt;
                collide_all_ocl(pos, hv, p, cl);
                s.xp[i] = pos;
            }
            if (keep_best && k >= 2) {
                p_prev = pp;
                rp_prev = s.rp;
                fp_prev = s.fp;
            }
            // pciCalcDensityPressure (SPHKernels.cl:737-804)
#pragma omp parallel for schedule(dynamic, 256)
            for (int32_t i = 0; i < n; i++) {
                V3<T> xpi = s.xp[i];
                T sum = 0;
                for_neighbors(g, n, cell_anchor<T>(xf, i), [&](int32_t j) {
                    T d = dist(s.xp[j], xpi);
                    if (d < h) {
                        uint32_t idx = (uint32_t)std::trunc(d / ds);
                        sum += tb.w[idx];
                    }
                });
                T rp = m * sum;
                T e = rp - rho0;
                T dp = delta * e;
                if (dp < (T)0) {
                    dp *= npf;
                    e *= npf;
                }
                s.rp[i] = rp;
                pp[i] += dp;
                err[i] = std::fabs(e);
            }
            if (warm && k == 1) {
                // learned warm start (no reference counterpart): unless the
                // loop stops at the first correction, its pressure is replaced
                // by the given seed (sorted order) before the pressure force
                T mx1 = 0;
                bool nf1 = false;
                for (int32_t i = 0; i < n; i++) {
                    if (std::isnan(err[i])) nf1 = true;
                    else mx1 = std::max(mx1, err[i]);
                }
                float r1 = (float)mx1 / p->rest_density;
                if (sizeof(T) == sizeof(double)) r1 = (float)(mx1 / rho0);
                const bool stop1 = nf1 || (k >= minit && r1 < p->max_density_error_ratio) || k >= maxit;
                if (!stop1) {
                    const bool envelope = warm->pprev && npf == (T)0;
                    for (int32_t i = 0; i < n; i++) {
                        if (!envelope) {
                            pp[i] = (T)warm->seed[i];
                            continue;
                        }
                        const T nn = warm->seed ? (T)warm->seed[i] : (T)INFINITY;
                        const T cap = (T)warm->kappa * (T)warm->pprev[g.perm[i]];
                        pp[i] = std::max(pp[i], std::min(nn, cap));
                    }
                }
            }
            // pciCalcPressureForce (SPHKernels.cl:806-878)
#pragma omp parallel for schedule(dynamic, 256)
            for (int32_t i = 0; i < n; i++) {
                V3<T> xpi = s.xp[i];
                T ri = s.rp[i];
                T tfac = pp[i] / (ri * ri);
                V3<T> pf{0, 0, 0};
                for_neighbors(g, n, cell_anchor<T>(xf, i), [&](int32_t j) {
                    V3<T> xpj = s.xp[j];
                    T d = dist(xpj, xpi);
                    if (d < h) {
                        if (i != j) {
                            V3<T> dir = pair_dir(xpi, xpj, d, i, j);
                            uint32_t idx = (uint32_t)std::trunc(d / ds);
                            T rj = s.rp[j];
                            T sc = tfac + pp[j] / (rj * rj);
                            T w = tb.p1[idx];
                            pf.x += dir.x * sc * w;
                            pf.y += dir.y * sc * w;
                            pf.z += dir.z * sc * w;
                        }
                    }
                });
                s.fp[i] = V3<T>{(-pf.x) * m * ri, (-pf.y) * m * ri, (-pf.z) * m * ri};
            }
            T mx = 0;
            for (int32_t i = 0; i < n; i++) {
                if (std::isnan(err[i])) nonfinite = true;
                else mx = std::max(mx, err[i]);
            }
            float ratio = (float)mx / p->rest_density;
            if (sizeof(T) == sizeof(double)) ratio = (float)(mx / rho0);
            if (tr && k <= 64) tr->err_ratio[k - 1] = ratio;
            if (tr) tr->final_err_ratio = ratio;
            if (nonfinite) break;
            // stall rule: count corrections that do not improve on the best
            // maximum error so far (from correction 2 on)
            if (k >= 2) {
                if (kb == 0 || mx < best) {
                    best = mx;
                    kb = k;
                    stall = 0;
                    if (keep_best) {
                        fp_best = fp_prev;
                        p_best = p_prev;
                        rp_best = rp_prev;
                    }
                } else {
                    stall++;
                }
            }
            if (k >= minit && ratio < p->max_density_error_ratio) {
                converged = 1;
                break;
            }
            if (k >= minit && p->stall_iterations > 0 && stall >= p->stall_iterations) {
                stalled = 1;
                break;
            }
        }
        if (k > maxit) k = maxit;
        if (stalled && keep_best) {
            s.fp = fp_best;
            pp = p_best;
            s.rp = rp_best;
            if (tr) tr->final_err_ratio = (float)(best / rho0);
            if (tr && sizeof(T) == sizeof(float)) tr->final_err_ratio = (float)best / p->rest_density;
        }
        if (tr) {
            tr->iterations = k;
            tr->converged = converged;
        }
        // pciAddPressureForce (SPHKernels.cl:880-890)
        for (int32_t i = 0; i < n; i++) {
            s.f[i].x += s.fp[i].x; s.f[i].y += s.fp[i].y; s.f[i].z += s.fp[i].z;
        }
        if (tr && tr->pressure) for (int32_t i = 0; i < n; i++) ((T*)tr->pressure)[i] = pp[i];
        if (tr && tr->rho_pred) for (int32_t i = 0; i < n; i++) ((T*)tr->rho_pred)[i] = s.rp[i];
        if (tr && tr->xpred)
            for (int32_t i = 0; i < n; i++) {
                ((T*)tr->xpred)[3 * i] = s.xp[i].x;
                ((T*)tr->xpred)[3 * i + 1] = s.xp[i].y;
                ((T*)tr->xpred)[3 * i + 2] = s.xp[i].z;
            }
        if (tr && tr->fp)
            for (int32_t i = 0; i < n; i++) {
                ((T*)tr->fp)[3 * i] = s.fp[i].x;
                ((T*)tr->fp)[3 * i + 1] = s.fp[i].y;
                ((T*)tr->fp)[3 * i + 2] = s.fp[i].z;
            }
        if (nonfinite) {
            if (tr) tr->status = ST_NONFINITE;
        }
    }
    if (tr && tr->force)
        for (int32_t i = 0; i < n; i++) {
            ((T*)tr->force)[3 * i] = s.f[i].x;
            ((T*)tr->force)[3 * i + 1] = s.f[i].y;
            ((T*)tr->force)[3 * i + 2] = s.f[i].z;
        }

    // ---- integrate (SPHKernels.cl:424-455) + handleCollisions (SPHKernels.cl:457-486)
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) {
        V3<T> hv = s.vh[i], pos = s.x[i];
        T ri = s.rho[i];
        V3<T> a{s.f[i].x / ri, s.f[i].y / ri, s.f[i].z / ri};
        if (s.first[i]) {
            hv.x = s.v[i].x - a.x * dt / (T)2;
            hv.y = s.v[i].y - a.y * dt / (T)2;
            hv.z = s.v[i].z - a.z * dt / (T)2;
            s.first[i] = 0;
        }
        V3<T> old = hv;
        hv.x += a.x * dt; hv.y += a.y * dt; hv.z += a.z * dt;
        pos.x += hv.x * dt; pos.y += hv.y * dt; pos.z += hv.z * dt;
        collide_all_ocl(pos, hv, p, cl);
        s.vh[i] = hv;
        s.x[i] = pos;
        s.v[i] = V3<T>{(old.x + hv.x) / (T)2, (old.y + hv.y) / (T)2, (old.z + hv.z) / (T)2};
    }
    for (int32_t i = 0; i < n; i++) {
        x[3 * i] = s.x[i].x; x[3 * i + 1] = s.x[i].y; x[3 * i + 2] = s.x[i].z;
        v[3 * i] = s.v[i].x; v[3 * i + 1] = s.v[i].y; v[3 * i + 2] = s.v[i].z;
        vh[3 * i] = s.vh[i].x; vh[3 * i + 1] = s.vh[i].y; vh[3 * i + 2] = s.vh[i].z;
        first[i] = s.first[i];
    }
    return tr ? tr->status : ST_OK;
}

// ------------------------------------------------------------ seq step --
// SPHSpatialGrid (SPHSpatialGrid.cpp:19-90): particles are bucketed by the
// cell (int)(x/h) (truncation toward zero); a query visits cells i-outer,
// j, k-inner over [(int)((x-h)/h), (int)((x+h)/h)] and keeps candidates with
// |x_i - x_j|^2 <= h^2, each bucket in insertion order.  The reference keys
// buckets by the hash 541i+79j+31k; within the queried 3x3x3 block no two cells
// share a hash and foreign cells that do are farther than h, so bucketing by
// the cell triple yields the same lists (checked against oracle/_ref).
struct SeqGrid {
    float gs;
    std::vector<int32_t> ci, cj, ck;
    // map cell triple -> list via sorted (cell, index) pairs
    std::vector<std::pair<uint64_t, int32_t>> entries;
    static uint64_t pack(int32_t i, int32_t j, int32_t k) {
        return ((uint64_t)(uint32_t)(i + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(j + (1 << 20)) << 21) |
               (uint64_t)(uint32_t)(k + (1 << 20));
    }
};

template <class T>
std::vector<std::vector<int32_t>> seq_neighbor_lists(float h, int32_t n, const T* x) {
    SeqGrid sg;
    sg.gs = h;
    sg.entries.resize(n);
    for (int32_t i = 0; i < n; i++) {
        int32_t a = (int32_t)((float)x[3 * i] / h), b = (int32_t)((float)x[3 * i + 1] / h),
                c = (int32_t)((float)x[3 * i + 2] / h);
        sg.entries[i] = {SeqGrid::pack(a, b, c), i};
    }
    std::stable_sort(sg.entries.begin(), sg.entries.end(),
                     [](const std::pair<uint64_t, int32_t>& u, const std::pair<uint64_t, int32_t>& w) {
                         return u.first < w.first;
                     });
    std::vector<std::vector<int32_t>> lists(n);
    const float h2 = h * h;
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; i++) {
        float px = (float)x[3 * i], py = (float)x[3 * i + 1], pz = (float)x[3 * i + 2];
        int32_t iMin = (int32_t)((px - h) / h), iMax = (int32_t)((px + h) / h);
        int32_t jMin = (int32_t)((py - h) / h), jMax = (int32_t)((py + h) / h);
        int32_t kMin = (int32_t)((pz - h) / h), kMax = (int32_t)((pz + h) / h);
        std::vector<int32_t>& out = lists[i];
        for (int32_t a = iMin; a <= iMax; a++)
            for (int32_t b = jMin; b <= jMax; b++)
                for (int32_t c = kMin; c <= kMax; c++) {
                    uint64_t key = SeqGrid::pack(a, b, c);
                    auto it = std::lower_bound(sg.entries.begin(), sg.entries.end(),
                                               std::make_pair(key, (int32_t)-1));
                    for (; it != sg.entries.end() && it->first == key; ++it) {
                        int32_t j = it->second;
                        float dx = px - (float)x[3 * j], dy = py - (float)x[3 * j + 1],
                              dz = pz - (float)x[3 * j + 2];
                        if (dx * dx + dy * dy + dz * dz <= h2) out.push_back(j);
                    }
                }
    }
    return lists;
}

template <class T>
inline void collide_all_seq(V3<T>& pos, V3<T>& vel, const orc_params* p, const orc_box* boxes,
                            const orc_sphere* spheres, const orc_plane* planes, const int32_t* kinds,
                            const int32_t* index, int32_t ncoll) {
    T r = (T)p->particle_radius, e = (T)p->restitution, mu = (T)p->friction;
    for (int32_t c = 0; c < ncoll; c++) {
        if (kinds[c] == 0) collide_box(pos, vel, r, e, mu, boxes[index[c]]);
        else if (kinds[c] == 1) collide_sphere_seq(pos, vel, r, e, mu, spheres[index[c]]);
        else collide_plane(pos, vel, r, e, mu, planes[index[c]]);
    }
}

// NONE path of PhysicSolver::update (PhysicSolver.cpp:12-21) in FP32.
int step_seq_f32(const orc_params* p, const orc_box* boxes, const orc_sphere* spheres,
                 const orc_plane* planes, const int32_t* kinds, const int32_t* index, int32_t ncoll,
                 float dt, int32_t n, float* xa, float* va, float* vha, uint8_t* firsta, orc_trace* tr) {
    typedef float T;
    if (tr) { tr->status = ST_OK; tr->iterations = 0; }
    if (n <= 0) return ST_INVALID;
    Tables<T> tb = make_tables<T>(kernel_radius<T>(p));
    const T m = particle_mass<T>(p, tb);
    const T h = tb.h;
    const T rho0 = p->rest_density;
    std::vector<V3<T>> x(n), v(n), vh(n), f(n, V3<T>{0, 0, 0}), oldvh(n);
    std::vector<T> rho(n), pr(n);
    std::vector<uint8_t> first(firsta, firsta + n);
    for (int32_t i = 0; i < n; i++) {
        x[i] = V3<T>{xa[3 * i], xa[3 * i + 1], xa[3 * i + 2]};
        v[i] = V3<T>{va[3 * i], va[3 * i + 1], va[3 * i + 2]};
        vh[i] = V3<T>{vha[3 * i], vha[3 * i + 1], vha[3 * i + 2]};
        // m_oldHalfVelocity is zero for a particle that has not stepped yet
        // (SPHParticle.cpp:4-14).  For a particle that has, the NONE path's
        // aproximateVelocity side effect (PCISPHSolver.cpp:96) recomputes v from
        // the same two halves that set it at the end of the previous step, so it
        // is the identity and is applied only to first-step particles below.
        oldvh[i] = V3<T>{0, 0, 0};
    }
    // buildCachedNeighborLists (SPHSolver.cpp:431-440)
    auto lists = seq_neighbor_lists<T>(h, n, xa);
    // calcParticleDensityPressure NONE (SPHSolver.cpp:444-464)
    for (int32_t i = 0; i < n; i++) {
        T sum = 0;
        for (int32_t j : lists[i]) {
            V3<T> d{x[j].x - x[i].x, x[j].y - x[i].y, x[j].z - x[i].z};
            sum += lookup(tb.w, tb, len3(d.x, d.y, d.z));
        }
        rho[i] = m * sum;
        T pp = (T)p->stiffness * (rho[i] - rho0);
        if (pp < 0.f) pp *= p->negative_pressure_factor;
        pr[i] = pp;
    }
    if (tr && tr->rho) std::memcpy(tr->rho, rho.data(), sizeof(T) * n);
    // accumulateNonPressureForces NONE (SPHSolver.cpp:203-248)
    for (int32_t i = 0; i < n; i++) {
        V3<T> F{0.f + p->gravity[0] * rho[i], 0.f + p->gravity[1] * rho[i], 0.f + p->gravity[2] * rho[i]};
        V3<T> sn{0, 0, 0};
        for (int32_t j : lists[i]) {
            V3<T> dir{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
            V3<T> dd{x[j].x - x[i].x, x[j].y - x[i].y, x[j].z - x[i].z};
            T d = len3(dd.x, dd.y, dd.z);
            T w = lookup(tb.w1, tb, d);
            sn.x += dir.x * w / rho[j];
            sn.y += dir.y * w / rho[j];
            sn.z += dir.z * w / rho[j];
        }
        sn.x *= m; sn.y *= m; sn.z *= m;
        T snl = len3(sn.x, sn.y, sn.z);
        if (snl > p->surface_tension_threshold) {
            T lc = 0;
            for (int32_t j : lists[i]) {
                V3<T> dd{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
                lc += lookup(tb.w2, tb, len3(dd.x, dd.y, dd.z)) / rho[j];
            }
            T k = -p->surface_tension * lc * m;
            F.x += (sn.x / snl) * k;
            F.y += (sn.y / snl) * k;
            F.z += (sn.z / snl) * k;
        }
        V3<T> vf{0, 0, 0};
        for (int32_t j : lists[i]) {
            if (j == i) continue;
            V3<T> dd{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
            T w = lookup(tb.v2, tb, len3(dd.x, dd.y, dd.z));
            vf.x += ((v[j].x - v[i].x) / rho[j]) * w;
            vf.y += ((v[j].y - v[i].y) / rho[j]) * w;
            vf.z += ((v[j].z - v[i].z) / rho[j]) * w;
        }
        T k2 = p->viscosity * m;
        vf.x *= k2; vf.y *= k2; vf.z *= k2;
        F.x += vf.x; F.y += vf.y; F.z += vf.z;
        f[i] = F;
    }
    if (p->method == 0) {
        // accumulatePressureForces NONE (SPHSolver.cpp:270-291)
        std::vector<V3<T>> fn = f;
        for (int32_t i = 0; i < n; i++) {
            V3<T> pf{0, 0, 0};
            T tfac = pr[i] / (rho[i] * rho[i]);
            for (int32_t j : lists[i]) {
                if (j == i) continue;
                V3<T> dd{x[i].x - x[j].x, x[i].y - x[j].y, x[i].z - x[j].z};
                T d = len3(dd.x, dd.y, dd.z);
                V3<T> dir{dd.x / d, dd.y / d, dd.z / d};
                T w = lookup(tb.p1, tb, d);
                T sc = tfac + pr[j] / (rho[j] * rho[j]);
                pf.x += dir.x * w * sc;
                pf.y += dir.y * w * sc;
                pf.z += dir.z * w * sc;
            }
            T k = -(m * rho[i]);
            pf.x *= k; pf.y *= k; pf.z *= k;
            fn[i].x += pf.x; fn[i].y += pf.y; fn[i].z += pf.z;
        }
        f = fn;
        if (tr && tr->pressure) std::memcpy(tr->pressure, pr.data(), sizeof(T) * n);
    } else {
        // PCISPHSolver::accumulatePressureForces NONE (PCISPHSolver.cpp:55-170)
        Tables<float> tf = make_tables<float>(kernel_radius<float>(p));
        float delta = delta_f32(p, tf, m, dt);
        if (tr) tr->delta = delta;
        std::vector<V3<T>> xp(n), vhp(n), fp(n, V3<T>{0, 0, 0});
        std::vector<T> rp(n, 0.f), de(n, 0.f);
        for (int32_t i = 0; i < n; i++) pr[i] = 0.f;
        int32_t maxit = p->max_iterations > 0 ? p->max_iterations : 1;
        int32_t k;
        for (k = 1; k <= maxit; k++) {
            for (int32_t i = 0; i < n; i++) {
                V3<T> hv = vh[i];
                V3<T> a{(f[i].x + fp[i].x) / rho[i], (f[i].y + fp[i].y) / rho[i], (f[i].z + fp[i].z) / rho[i]};
                if (first[i]) {
                    hv.x = v[i].x - a.x * dt / 2.f;
                    hv.y = v[i].y - a.y * dt / 2.f;
                    hv.z = v[i].z - a.z * dt / 2.f;
                }
                vhp[i] = V3<T>{hv.x + a.x * dt, hv.y + a.y * dt, hv.z + a.z * dt};
                xp[i] = V3<T>{x[i].x + vhp[i].x * dt, x[i].y + vhp[i].y * dt, x[i].z + vhp[i].z * dt};
            }
            for (int32_t i = 0; i < n; i++) {
                collide_all_seq(xp[i], vhp[i], p, boxes, spheres, planes, kinds, index, ncoll);
                // aproximateVelocity on the particle itself (hazard H4, PCISPHSolver.cpp:96)
                if (first[i])
                    v[i] = V3<T>{(oldvh[i].x + vh[i].x) / 2.f, (oldvh[i].y + vh[i].y) / 2.f,
                                 (oldvh[i].z + vh[i].z) / 2.f};
            }
            for (int32_t i = 0; i < n; i++) {
                T sum = 0;
                for (int32_t j : lists[i]) {
                    V3<T> dd{xp[j].x - xp[i].x, xp[j].y - xp[i].y, xp[j].z - xp[i].z};
                    sum += lookup(tb.w, tb, len3(dd.x, dd.y, dd.z));
                }
                T rpp = m * sum;
                T e = rpp - rho0;
                T dp = delta * e;
                if (dp < 0.f) {
                    e *= p->negative_pressure_factor;
                    dp *= p->negative_pressure_factor;
                }
                rp[i] = rpp;
                de[i] = e;
                pr[i] = pr[i] + dp;
            }
            for (int32_t i = 0; i < n; i++) {
                V3<T> pf{0, 0, 0};
                T tfac = pr[i] / (rp[i] * rp[i]);
                for (int32_t j : lists[i]) {
                    if (j == i) continue;
                    V3<T> dd{xp[j].x - xp[i].x, xp[j].y - xp[i].y, xp[j].z - xp[i].z};
                    T d = len3(dd.x, dd.y, dd.z);
                    V3<T> dir{(xp[i].x - xp[j].x) / d, (xp[i].y - xp[j].y) / d, (xp[i].z - xp[j].z) / d};
                    T w = lookup(tb.p1, tb, d);
                    T sc = tfac + pr[j] / (rp[j] * rp[j]);
                    pf.x += dir.x * w * sc;
                    pf.y += dir.y * w * sc;
                    pf.z += dir.z * w * sc;
                }
                T kk = -(m * rho[i]);  // hazard H2: the NONE path multiplies by rho_i
                pf.x *= kk; pf.y *= kk; pf.z *= kk;
                if (std::isnan(pf.x)) pf = V3<T>{0, 0, 0};
                fp[i] = pf;
            }
            T mx = 0.f;
            for (int32_t i = 0; i < n; i++) mx = std::max(mx, std::fabs(de[i]));
            float ratio = mx / rho0;
            if (tr && k <= 64) tr->err_ratio[k - 1] = ratio;
            if (tr) tr->final_err_ratio = ratio;
            if (k >= p->min_iterations && std::fabs(ratio) < p->max_density_error_ratio) break;
        }
        if (k > maxit) k = maxit;
        if (tr) tr->iterations = k;
        for (int32_t i = 0; i < n; i++) {
            f[i].x += fp[i].x; f[i].y += fp[i].y; f[i].z += fp[i].z;
        }
        if (tr && tr->pressure) std::memcpy(tr->pressure, pr.data(), sizeof(T) * n);
        if (tr && tr->rho_pred) std::memcpy(tr->rho_pred, rp.data(), sizeof(T) * n);
        if (tr && tr->xpred)
            for (int32_t i = 0; i < n; i++) {
                ((T*)tr->xpred)[3 * i] = xp[i].x;
                ((T*)tr->xpred)[3 * i + 1] = xp[i].y;
                ((T*)tr->xpred)[3 * i + 2] = xp[i].z;
            }
        if (tr && tr->fp)
            for (int32_t i = 0; i < n; i++) {
                ((T*)tr->fp)[3 * i] = fp[i].x;
                ((T*)tr->fp)[3 * i + 1] = fp[i].y;
                ((T*)tr->fp)[3 * i + 2] = fp[i].z;
            }
    }
    if (tr && tr->force)
        for (int32_t i = 0; i < n; i++) {
            ((T*)tr->force)[3 * i] = f[i].x;
            ((T*)tr->force)[3 * i + 1] = f[i].y;
            ((T*)tr->force)[3 * i + 2] = f[i].z;
        }
    // integrate (SPHParticle.cpp:30-44) + handleCollisions NONE (SPHSolver.cpp:339-355)
    for (int32_t i = 0; i < n; i++) {
        V3<T> a{f[i].x / rho[i], f[i].y / rho[i], f[i].z / rho[i]};
        if (first[i]) {
            vh[i].x = v[i].x - a.x * dt / 2.f;
            vh[i].y = v[i].y - a.y * dt / 2.f;
            vh[i].z = v[i].z - a.z * dt / 2.f;
            first[i] = 0;
        }
        oldvh[i] = vh[i];
        vh[i].x += a.x * dt; vh[i].y += a.y * dt; vh[i].z += a.z * dt;
        x[i].x += vh[i].x * dt; x[i].y += vh[i].y * dt; x[i].z += vh[i].z * dt;
    }
    for (int32_t i = 0; i < n; i++) {
        collide_all_seq(x[i], vh[i], p, boxes, spheres, planes, kinds, index, ncoll);
        v[i] = V3<T>{(oldvh[i].x + vh[i].x) / 2.f, (oldvh[i].y + vh[i].y) / 2.f, (oldvh[i].z + vh[i].z) / 2.f};
    }
    for (int32_t i = 0; i < n; i++) {
        xa[3 * i] = x[i].x; xa[3 * i + 1] = x[i].y; xa[3 * i + 2] = x[i].z;
        va[3 * i] = v[i].x; va[3 * i + 1] = v[i].y; va[3 * i + 2] = v[i].z;
        vha[3 * i] = vh[i].x; vha[3 * i + 1] = vh[i].y; vha[3 * i + 2] = vh[i].z;
        firsta[i] = first[i];
    }
    return ST_OK;
}

}  // namespace

// ================================================================= C API ==
extern "C" {

void orc_set_best_exit(int32_t on) { g_best_exit = on; }

int orc_derived_f32(const orc_params* p, float* h, float* div_step, float* mass, float* tables) {
    Tables<float> t = make_tables<float>(kernel_radius<float>(p));
    if (h) *h = t.h;
    if (div_step) *div_step = t.div;
    if (mass) *mass = particle_mass<float>(p, t);
    if (tables) {
        std::memcpy(tables + 0 * ORC_TABLE_SIZE, t.w.data(), sizeof(float) * ORC_TABLE_SIZE);
        std::memcpy(tables + 1 * ORC_TABLE_SIZE, t.w1.data(), sizeof(float) * ORC_TABLE_SIZE);
        std::memcpy(tables + 2 * ORC_TABLE_SIZE, t.w2.data(), sizeof(float) * ORC_TABLE_SIZE);
        std::memcpy(tables + 3 * ORC_TABLE_SIZE, t.p1.data(), sizeof(float) * ORC_TABLE_SIZE);
        std::memcpy(tables + 4 * ORC_TABLE_SIZE, t.v2.data(), sizeof(float) * ORC_TABLE_SIZE);
    }
    return ST_OK;
}

int orc_derived_f64(const orc_params* p, double* h, double* div_step, double* mass, double* tables) {
    Tables<double> t = make_tables<double>(kernel_radius<double>(p));
    if (h) *h = t.h;
    if (div_step) *div_step = t.div;
    if (mass) *mass = particle_mass<double>(p, t);
    if (tables) {
        std::memcpy(tables + 0 * ORC_TABLE_SIZE, t.w.data(), sizeof(double) * ORC_TABLE_SIZE);
        std::memcpy(tables + 1 * ORC_TABLE_SIZE, t.w1.data(), sizeof(double) * ORC_TABLE_SIZE);
        std::memcpy(tables + 2 * ORC_TABLE_SIZE, t.w2.data(), sizeof(double) * ORC_TABLE_SIZE);
        std::memcpy(tables + 3 * ORC_TABLE_SIZE, t.p1.data(), sizeof(double) * ORC_TABLE_SIZE);
        std::memcpy(tables + 4 * ORC_TABLE_SIZE, t.v2.data(), sizeof(double) * ORC_TABLE_SIZE);
    }
    return ST_OK;
}

float orc_delta_f32(const orc_params* p, float dt) {
    Tables<float> t = make_tables<float>(kernel_radius<float>(p));
    return delta_f32(p, t, particle_mass<float>(p, t), dt);
}

int orc_grid(const orc_params* p, int32_t n, const float* x, orc_trace* tr, int32_t use_radix) {
    Grid g = build_grid(p, n, x, use_radix != 0);
    fill_trace_grid(g, n, tr);
    if (tr) tr->status = g.status;
    return g.status;
}

int orc_step_ocl_f32(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                     int32_t nsphere, const orc_plane* planes, int32_t nplane, float dt, int32_t n, float* x,
                     float* v, float* vh, uint8_t* first, orc_trace* tr, int32_t nthreads) {
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    return step_ocl<float>(p, cl, dt, n, x, v, vh, first, tr, nthreads);
}

int orc_step_ocl_seeded_f32(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                            int32_t nsphere, const orc_plane* planes, int32_t nplane, float dt, int32_t n, float* x,
                            float* v, float* vh, uint8_t* first, const float* seed, orc_trace* tr,
                            int32_t nthreads) {
    if (!seed) return ST_INVALID;
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    WarmStart w;
    w.seed = seed;
    return step_ocl<float>(p, cl, dt, n, x, v, vh, first, tr, nthreads, &w);
}

int orc_step_ocl_warm_f32(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                          int32_t nsphere, const orc_plane* planes, int32_t nplane, float dt, int32_t n, float* x,
                          float* v, float* vh, uint8_t* first, const float* seed, const float* pprev, float kappa,
                          orc_trace* tr, int32_t nthreads) {
    if (!pprev) return ST_INVALID;
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    WarmStart w;
    w.seed = seed;
    w.pprev = pprev;
    w.kappa = kappa;
    return step_ocl<float>(p, cl, dt, n, x, v, vh, first, tr, nthreads, &w);
}

int orc_step_ocl_f64(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                     int32_t nsphere, const orc_plane* planes, int32_t nplane, float dt, int32_t n, double* x,
                     double* v, double* vh, uint8_t* first, orc_trace* tr, int32_t nthreads) {
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    return step_ocl<double>(p, cl, dt, n, x, v, vh, first, tr, nthreads);
}

int orc_step_seq_f32(const orc_params* p, const orc_box* boxes, const orc_sphere* spheres,
                     const orc_plane* planes, const int32_t* kinds, const int32_t* index, int32_t ncoll,
                     float dt, int32_t n, float* x, float* v, float* vh, uint8_t* first, orc_trace* tr) {
    return step_seq_f32(p, boxes, spheres, planes, kinds, index, ncoll, dt, n, x, v, vh, first, tr);
}

int orc_collide_ocl(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                    int32_t nsphere, const orc_plane* planes, int32_t nplane, int32_t n, float* x,
                    float* vel) {
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    for (int32_t i = 0; i < n; i++) {
        V3<float> pos{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        V3<float> v{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        collide_all_ocl(pos, v, p, cl);
        x[3 * i] = pos.x; x[3 * i + 1] = pos.y; x[3 * i + 2] = pos.z;
        vel[3 * i] = v.x; vel[3 * i + 1] = v.y; vel[3 * i + 2] = v.z;
    }
    return ST_OK;
}

int orc_neighbors_seq(float h, int32_t n, const float* x, int32_t i, int32_t* out, int32_t cap) {
    auto lists = seq_neighbor_lists<float>(h, n, x);
    int32_t c = (int32_t)lists[i].size();
    for (int32_t k = 0; k < c && k < cap; k++) out[k] = lists[i][k];
    return c;
}

// O(N^2) density with the same lookup rule, accumulated in double (test pin).
int orc_brute_density(const orc_params* p, int32_t n, const float* x, double* rho) {
    Tables<float> t = make_tables<float>(kernel_radius<float>(p));
    float m = particle_mass<float>(p, t);
    for (int32_t i = 0; i < n; i++) {
        double s = 0;
        V3<float> xi{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        for (int32_t j = 0; j < n; j++) {
            V3<float> xj{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
            float d = dist(xj, xi);
            if (d < t.h) s += t.w[(uint32_t)std::trunc(d / t.div)];
        }
        rho[i] = (double)m * s;
    }
    return ST_OK;
}

int orc_run_ocl_f32(const orc_params* p, const orc_box* boxes, int32_t nbox, const orc_sphere* spheres,
                    int32_t nsphere, const orc_plane* planes, int32_t nplane, float dt, int32_t nsteps,
                    int32_t n, float* x, float* v, float* vh, uint8_t* first, int32_t* iterations_out,
                    float* err_out, int32_t nthreads) {
    Colliders cl{boxes, nbox, spheres, nsphere, planes, nplane};
    for (int32_t s = 0; s < nsteps; s++) {
        orc_trace tr;
        std::memset(&tr, 0, sizeof(tr));
        int st = step_ocl<float>(p, cl, dt, n, x, v, vh, first, &tr, nthreads);
        if (iterations_out) iterations_out[s] = tr.iterations;
        if (err_out) err_out[s] = tr.iterations > 0 ? tr.final_err_ratio : 0.f;
        if (st != ST_OK) return st;
    }
    return ST_OK;
}

}  // extern "C"
