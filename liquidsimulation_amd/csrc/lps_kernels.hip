// lps_kernels.hip -- hand-written gfx950 kernels of the SPH/PCISPH step.
//
// One launch per phase of PhysicSolver::update (LiquidPhysics/src/PhysicSolver.cpp:12-21):
//   grid      k_grid_one (or k_scene_grid + k_grid_finish), k_keys,
//             k_radix_{hist,scan,scatter} x P, k_reorder (+ run-end marks), k_cell_scan
//   density   k_density
//   forces    k_forces (+ first PCISPH prediction)
//   pressure  SPH: k_pressure_sph;  PCISPH: {k_pci_density, k_pci_pressure (+ the
//             loop rule in block 0, + next prediction)} per correction
//   integrate k_integrate (+ collisions, + bounding box of the next step)
// Particle kernels run 256 threads (4 wave64) per block, one particle per thread
// (the *_q variants of the stencil kernels: a quad of threads per particle),
// with a block index remapped so each of the 8 XCDs works on a contiguous range
// of the cell-sorted particles and keeps its neighbours in its own L2.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <type_traits>

#include "lps_common.h"
#include "lps_div.h"
#include "lps_kernels.h"

#define LPS_XCDS 8u

namespace {

// ------------------------------------------------------------ utilities --
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Logical block of an XCD-contiguous mapping: physical block b runs on XCD b % 8;
// XCD k gets logical blocks [k*per, (k+1)*per).  Returns UINT_MAX for padding blocks.
__device__ __forceinline__ uint32_t xcd_block(uint32_t nlogical) {
    const uint32_t per = (nlogical + LPS_XCDS - 1) / LPS_XCDS;
    const uint32_t b = blockIdx.x;
    const uint32_t lb = (b % LPS_XCDS) * per + b / LPS_XCDS;
    return lb < nlogical ? lb : UINT_MAX;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_sum_f32(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// 64-lane mask of the lanes holding the same 8-bit digit (8 ballots).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
    uint64_t peers = active;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// LDS copies of the kernel tables, staged by a workgroup of NT threads.  A
// thread loads all of its entries of every table it stages before it
// stores any, so the workgroup waits out one memory round trip; a strided
// copy loop (the thread count unknown to the compiler) waits once per pass
// and table.
template <uint32_t NT, uint32_t N>
struct TableSlice {
    static constexpr uint32_t IT = (N + NT - 1) / NT;
    float v[IT];
    __device__ __forceinline__ void load(const float* __restrict__ src) {
#pragma unroll
        for (uint32_t r = 0; r < IT; r++) {
            const uint32_t k = threadIdx.x + r * NT;
            v[r] = k < N ? src[k] : 0.f;
        }
    }
    __device__ __forceinline__ void store(float* dst) const {
#pragma unroll
        for (uint32_t r = 0; r < IT; r++) {
            const uint32_t k = threadIdx.x + r * NT;
            if (k < N) dst[k] = v[r];
        }
    }
};

// One table (at src) and, when with_thr, the index thresholds (stored after
// the five tables).
template <uint32_t NT>
__device__ __forceinline__ void stage_table(float* dst, const float* src, float* thr, const float* tables,
                                            bool with_thr = true) {
    TableSlice<NT, LPS_TABLE_SIZE> a;
    TableSlice<NT, LPS_THR_SIZE> t;
    a.load(src);
    if (with_thr) t.load(tables + 5 * LPS_TABLE_SIZE);
    a.store(dst);
    if (with_thr) t.store(thr);
}

// The forces' three tables (W', W'' and the viscosity Laplacian) and the
// thresholds.
template <uint32_t NT>
__device__ __forceinline__ void stage_force_tables(float* w1, float* w2, float* v2, float* thr, const float* tables) {
    TableSlice<NT, LPS_TABLE_SIZE> a, b, c;
    TableSlice<NT, LPS_THR_SIZE> t;
    a.load(tables + 1 * LPS_TABLE_SIZE);
    b.load(tables + 2 * LPS_TABLE_SIZE);
    c.load(tables + 4 * LPS_TABLE_SIZE);
    t.load(tables + 5 * LPS_TABLE_SIZE);
    a.store(w1);
    b.store(w2);
    c.store(v2);
    t.store(thr);
}

// ---------------------------------------------------- grid traversal --
struct GridView {
    float off0, off1, off2, gs;
    int gx, gy, gz;
    uint32_t gxy;
    uint32_t base;  // global index of the scene's cell 0
};

__device__ __forceinline__ GridView grid_view(const DevGrid* g, const SceneGrid* sg) {
    GridView v;
    v.off0 = sg->offset[0];
    v.off1 = sg->offset[1];
    v.off2 = sg->offset[2];
    v.gs = g->spacing;
    v.gx = (int)sg->gsize[0];
    v.gy = (int)sg->gsize[1];
    v.gz = (int)sg->gsize[2];
    v.gxy = sg->gsize[0] * sg->gsize[1];
    v.base = sg->cell_base;
    return v;
}

// The scene of a particle block and its constants; the block's real
// particles are the slots below info.slot0 + info.count.
struct BlockScene {
    uint32_t s;
    SceneInfo info;
};
__device__ __forceinline__ BlockScene block_scene(const lps_dev::Scenes& S, uint32_t lb) {
    BlockScene b;
    b.s = S.block_scene[lb];
    b.info = S.info[b.s];
    return b;
}

// The correction kernels may run a particle block of LPS_BLOCK slots as
// LPS_BLOCK / WG workgroups of WG threads (the autotuner's choice): fewer
// threads per workgroup spread the late corrections of an ensemble, when few
// scenes still run, over more compute units.  Pieces of the active blocks of
// list L are split XCD-contiguously over the 8 XCDs, so the pieces of one
// block share an L2.  Returns the particle block and, in first_slot, the
// workgroup's first slot; UINT_MAX when the workgroup has no piece.
//
// A single scene's lists hold all of its blocks in order while its loop runs
// (a list is rebuilt only when a scene stops, and the blocks of the
// corrections queued past its stop return at the stop check instead), so
// there the piece follows from the block index alone and the workgroup
// starts its particles' loads without two dependent reads.
template <int WG>
__device__ __forceinline__ uint32_t active_piece(const DevGrid* g, const lps_dev::Scenes& S, uint32_t L,
                                                 uint32_t& first_slot) {
    constexpr uint32_t P = LPS_BLOCK / WG;
    const bool one = S.count == 1u;
    const uint32_t act = (one ? S.nblocks : g->act_count[L]) * P;
    const uint32_t per = (act + LPS_XCDS - 1) / LPS_XCDS;
    const uint32_t b = blockIdx.x;
    if (b / LPS_XCDS >= per) return UINT_MAX;
    const uint32_t lp = (b % LPS_XCDS) * per + b / LPS_XCDS;
    if (lp >= act) return UINT_MAX;
    const uint32_t lb = one ? lp / P : S.act_list[L * S.nblocks + lp / P];
    first_slot = lb * LPS_BLOCK + (lp % P) * (uint32_t)WG;
    return lb;
}

// The solver constants as seen by one scene: its gravity and collider counts.
__device__ __forceinline__ SimConst scene_const(const SimConst& c, const SceneInfo& si) {
    SimConst k = c;
    k.gravity[0] = si.gravity[0];
    k.gravity[1] = si.gravity[1];
    k.gravity[2] = si.gravity[2];
    k.nbox = si.nbox;
    k.nsphere = si.nsphere;
    k.nplane = si.nplane;
    return k;
}

// The 27-cell stencil of the OpenCL kernels (e.g. SPHKernels.cl:219-244):
// cells z-outer, y, x-inner around the cell of `anchor`, each cell's sorted
// range ascending.  Cells x-1, x, x+1 of one row are consecutive keys, so a
// row is the single range [LB[k(x-1)], LB[k(x+1)+1]) with the same visiting order.
template <class F>
__device__ __forceinline__ void for_neighbors(const GridView& G, const uint32_t* __restrict__ lb, float ax,
                                              float ay, float az, F&& fn) {
    const int xg = (int)truncf((ax + G.off0) / G.gs);
    const int yg = (int)truncf((ay + G.off1) / G.gs);
    const int zg = (int)truncf((az + G.off2) / G.gs);
    const int x0 = max(xg - 1, 0), x1 = min(xg + 1, G.gx - 1);
    for (int z = zg - 1; z <= zg + 1; z++) {
        if (z < 0 || z >= G.gz) continue;
        for (int y = yg - 1; y <= yg + 1; y++) {
            if (y < 0 || y >= G.gy) continue;
            const uint32_t base = G.base + (uint32_t)G.gx * (uint32_t)y + G.gxy * (uint32_t)z;
            const uint32_t b = lb[base + (uint32_t)x0];
            const uint32_t e = lb[base + (uint32_t)x1 + 1u];
            for (uint32_t j = b; j < e; j++) fn(j);
        }
    }
}

// The same stencil in two passes per window of up to 32 candidates of a row:
// pass 1 runs a cheap filter over every candidate (eight independent loads in
// flight; squared distance with fused multiply-adds against s_relax, which
// bounds its rounding difference to the exact form) and records a per-lane
// accept mask; pass 2 (visit_window) visits the kept candidates in ascending
// order, four loads at a time (position and pre(j)), forms the exact squared
// distance s2 (the kernels' unfused expression) and calls fn(j, x_j, s2,
// pre(j), x_j - x_i) when s2 <= s_max, i.e. when d = sqrtf(s2) < h.  Accepted set and
// visiting order are those of for_neighbors, so sums are bit-identical.
// Candidates are addressed by 32-bit byte offsets from xs (particle counts
// stay below 2^28); pass 1 reads up to 7 entries past a row's end, so xs
// carries LPS_STENCIL_PAD entries of padding (masked out, never used).
//
// Recorded masks (NbMasks): MC_RECORD also stores the accept mask of the
// first window of each in-grid row (0 for an empty row); MC_REPLAY, in a
// later pass over the same positions and anchor cells, runs pass 2 on the
// recorded masks (those pass 1 would compute again) and filters live only
// the windows past the first of a row longer than 32 candidates, or the whole
// stencil when the slot was not recorded.  Every mode accepts the same pairs
// in the same order.
enum : int { MC_LIVE = 0, MC_RECORD = 1, MC_REPLAY = 2 };

// Compact positions for the recording passes' filter.  The sorted positions
// (by the reorder) and each prediction are also stored as three 16-bit
// fixed-point coordinates, u = floor((x + off) * Q) mod 2^16 with Q = 4096 /
// cell size (8 bytes: x | y << 16, z), so that pass 1 of the density and
// predicted-density passes reads 8 bytes per candidate instead of 12 and tests
// it with packed 16-bit differences and two dot2 instructions.  The
// filter only has to keep every neighbour (pass 2 decides exactly):
// * differences wrap mod 2^16, so they are exact while |dx| < 8 cells; a
//   neighbour pair (d <= h <= cell size) differs by at most 4096 units plus
//   rounding on each axis, and a wrapped difference can only add candidates;
// * for |(x + off) * Q| < 2^22 the stored value is within (-1.5, 0.5) of the
//   exact one (floor, two roundings of at most 0.25), so a difference is off
//   by less than 2 units per axis and a distance by less than 3.5;
// * the test keeps s_q <= (ceil(sqrt(s_max) * Q) + 8)^2, so every pair with
//   s2 <= s_max is kept;
// * a particle whose own |(x + off) * Q| reaches 2^21 on some axis filters
//   with floats instead; its neighbours then all lie within 2^22.
// Far predictions (beyond 2^22) store a saturated value that can only add
// candidates.  The dot products start from the bias -(T + 1), T the squared
// radius in units, so a candidate is kept exactly when the saturated result
// is negative: s_q <= T gives s_q - (T + 1) in [-(T + 1), -1], and s_q > T a
// positive result (sums up to 3 * 2^30 saturate at 2^31 - 1).  alignbit
// shifts each sign bit into the mask, one instruction per candidate.
struct CompactFilter {
    const uint2* cq;  // nullptr: float filter
    uint2 qi;         // this particle's entry
    int bias;         // -(T + 1)
};

typedef short lps_short2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float compact_scale(const GridView& G) { return 4096.f / G.gs; }

__device__ __forceinline__ uint32_t quantize1(float x, float off, float qs) {
    const float v = fminf(fmaxf((x + off) * qs, -0x1p30f), 0x1p30f);
    return (uint32_t)(int)floorf(v) & 0xffffu;
}

__device__ __forceinline__ uint2 quantize(float x, float y, float z, const GridView& G, float qs) {
    return make_uint2(quantize1(x, G.off0, qs) | (quantize1(y, G.off1, qs) << 16), quantize1(z, G.off2, qs));
}

// Whether particle (x, y, z) may use the compact filter.
__device__ __forceinline__ bool compact_near(float x, float y, float z, const GridView& G, float qs) {
    return fabsf((x + G.off0) * qs) < 0x1p21f && fabsf((y + G.off1) * qs) < 0x1p21f &&
           fabsf((z + G.off2) * qs) < 0x1p21f;
}

__device__ __forceinline__ int compact_bias(float s_max, float qs) {
    const float r = ceilf(sqrtf(s_max) * qs) + 8.f;
    return -(int)(r * r) - 1;
}

// s_q + bias, saturated; negative exactly when s_q <= T.
__device__ __forceinline__ uint32_t compact_diff(const uint2& e, const uint2& qi, int bias) {
    const lps_short2 dxy = __builtin_bit_cast(lps_short2, e.x) - __builtin_bit_cast(lps_short2, qi.x);
    const lps_short2 dz = __builtin_bit_cast(lps_short2, e.y) - __builtin_bit_cast(lps_short2, qi.y);
    return (uint32_t)__builtin_amdgcn_sdot2(dxy, dxy, __builtin_amdgcn_sdot2(dz, dz, bias, true), true);
}

// A visitor called for every slot of a trip with whether the slot is an
// accepted pair, for branch-free bodies.  The density sums add +0 for the
// other slots, which leaves a sum of non-negative terms (never -0) unchanged,
// so the result is that of the branching visit.
template <class F>
struct Predicated {
    F f;
};
template <class T>
struct is_predicated : std::false_type {};
template <class F>
struct is_predicated<Predicated<F>> : std::true_type {};
template <class F>
__device__ __forceinline__ Predicated<F> predicated(F f) {
    return Predicated<F>{f};
}

template <class F, class P>
__device__ __forceinline__ void visit_slot(F& fn, bool ok, uint32_t j, const float4& q, float s2, const P& pv,
                                           const float3& dl) {
    if constexpr (is_predicated<std::decay_t<F>>::value)
        fn.f(ok, j, q, s2, pv, dl);
    else if (ok)
        fn(j, q, s2, pv, dl);
}

// Pass 2 visits VB kept candidates per trip (their loads issue together);
// the float filter tests four candidates per trip.  With VB = 2 the recording
// passes stay at or below 64 VGPRs, i.e. eight waves per SIMD: the rare
// paths (the float filter of a far particle or of a replayed row's later
// windows, and the checked division) otherwise set the register count of the
// whole kernel.  A row holds about four kept candidates, so VB = 3 or 4
// mostly fills slots with repeats, and VB = 1 leaves one load in flight; in
// the recording passes all measure within 1 % at 524k particles.
template <int VB, class Acc, class Pre, class F>
__device__ __forceinline__ void visit_window(const char* __restrict__ xb, uint32_t w0, uint32_t mask, float px,
                                             float py, float pz, float s_max, Acc& accept_extra, Pre& pre, F& fn) {
    while (mask) {
        uint32_t kk[VB];
        uint32_t cnt = 1u;
        kk[0] = (uint32_t)__ffs((int)mask) - 1u;
        mask &= mask - 1u;
#pragma unroll
        for (int t = 1; t < VB; t++) {
            kk[t] = kk[0];
            if (mask) {
                kk[t] = (uint32_t)__ffs((int)mask) - 1u;
                mask &= mask - 1u;
                cnt++;
            }
        }
        float4 q[VB];
        decltype(pre(0u)) pv[VB];
#pragma unroll
        for (int t = 0; t < VB; t++) {
            q[t] = *reinterpret_cast<const float4*>(xb + ((w0 + kk[t]) << 4));
            pv[t] = pre(w0 + kk[t]);
        }
        // every load of the trip issues before the first use of any
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < VB; t++) {
            // slots past cnt repeat candidate kk[0]; one predicate keeps the
            // exec-mask bookkeeping to a single region per candidate
            const uint32_t j = w0 + kk[t];
            const float dx = q[t].x - px, dy = q[t].y - py, dz = q[t].z - pz;
            const float s2 = dx * dx + dy * dy + dz * dz;
            visit_slot(fn, (uint32_t)t < cnt && s2 <= s_max && accept_extra(j), j, q[t], s2, pv[t],
                       make_float3(dx, dy, dz));
        }
    }
}

// One stencil row [b, e) of for_neighbors_masked: the recorded first window
// (replay), then pass 1 and pass 2 over the remaining windows of 32.
template <int VB, class Acc, class Pre, class F>
__device__ __forceinline__ void visit_row(const char* __restrict__ xb, const CompactFilter& CF, uint32_t b, uint32_t e,
                                          uint32_t* __restrict__ row, uint32_t i, bool replay, bool record, float px,
                                          float py, float pz, float s_max, float s_relax, Acc& accept_extra, Pre& pre,
                                          F& fn) {
    const char* __restrict__ cb = reinterpret_cast<const char*>(CF.cq);
    auto keep = [&](const float4& q) {
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) <= s_relax;
    };
    uint32_t w0 = b;
    if (replay) {
        visit_window<VB>(xb, b, row[i], px, py, pz, s_max, accept_extra, pre, fn);
        w0 += 32u;
    }
    uint32_t first = 0u;
    for (; w0 < e; w0 += 32u) {
        const uint32_t n = min(32u, e - w0);
        uint32_t mask = 0u;
        if (cb) {
            for (uint32_t u = 0; u < n; u += 8u) {
                const char* p8 = cb + ((w0 + u) << 3);
                uint2 e[8];
#pragma unroll
                for (int t = 0; t < 8; t++) e[t] = *reinterpret_cast<const uint2*>(p8 + 8 * t);
                uint32_t m8 = 0u;
#pragma unroll
                for (int t = 7; t >= 0; t--) m8 = __builtin_amdgcn_alignbit(m8, compact_diff(e[t], CF.qi, CF.bias), 31u);
                mask |= m8 << u;
            }
        } else {
            for (uint32_t u = 0; u < n; u += 4u) {
                const char* p8 = xb + ((w0 + u) << 4);
                float4 q[4];
#pragma unroll
                for (int t = 0; t < 4; t++) q[t] = *reinterpret_cast<const float4*>(p8 + 16 * t);
                uint32_t m8 = 0u;
#pragma unroll
                for (int t = 0; t < 4; t++) m8 |= (keep(q[t]) ? 1u : 0u) << t;
                mask |= m8 << u;
            }
        }
        mask &= n == 32u ? 0xFFFFFFFFu : (1u << n) - 1u;
        if (w0 == b) first = mask;
        visit_window<VB>(xb, w0, mask, px, py, pz, s_max, accept_extra, pre, fn);
    }
    if (record) row[i] = first;
}

template <int VB, class Acc, class Pre, class F>
__device__ __forceinline__ void for_neighbors_masked(const GridView& G, const uint32_t* __restrict__ lb, float ax,
                                                     float ay, float az, const float4* __restrict__ xs, float px,
                                                     float py, float pz, float s_max, float s_relax,
                                                     const lps_dev::NbMasks& M, uint32_t i, int mode,
                                                     const CompactFilter& CF, Acc&& accept_extra, Pre&& pre, F&& fn) {
    const char* __restrict__ xb = reinterpret_cast<const char*>(xs);
    const int xg = (int)truncf((ax + G.off0) / G.gs);
    const int yg = (int)truncf((ay + G.off1) / G.gs);
    const int zg = (int)truncf((az + G.off2) / G.gs);
    const int x0 = max(xg - 1, 0), x1 = min(xg + 1, G.gx - 1);
    const bool replay = mode == MC_REPLAY && M.ok[i] != 0u;
    const bool record = mode == MC_RECORD;
    for (uint32_t rz = 0; rz < 3u; rz++) {
        const int z = zg - 1 + (int)rz;
        if (z < 0 || z >= G.gz) continue;
        for (uint32_t ry = 0; ry < 3u; ry++) {
            const int y = yg - 1 + (int)ry;
            if (y < 0 || y >= G.gy) continue;
            const uint32_t base = G.base + (uint32_t)G.gx * (uint32_t)y + G.gxy * (uint32_t)z;
            visit_row<VB>(xb, CF, lb[base + (uint32_t)x0], lb[base + (uint32_t)x1 + 1u],
                          M.mask + (3u * rz + ry) * M.stride, i, replay, record, px, py, pz, s_max, s_relax,
                          accept_extra, pre, fn);
        }
    }
    if (record) M.ok[i] = 1u;
}

// The replaying walk with the three rows of a z-plane merged: each lane visits
// its recorded kept candidates of rows y-1, y, y+1 as one sequence, VB per
// trip, so a plane costs the wave the trips of its fullest lane instead of the
// sum over its rows of each row's fullest lane.  A trip's gathers then come
// from up to three row ranges of the plane.  Accepted set and visiting order
// are those of for_neighbors_masked in MC_REPLAY mode; a plane where some lane
// of the wave has no recorded masks or a row longer than one window of 32
// candidates is walked row by row as there.
template <int VB, class Acc, class Pre, class F>
__device__ __forceinline__ void for_neighbors_planes(const GridView& G, const uint32_t* __restrict__ lb, float ax,
                                                     float ay, float az, const float4* __restrict__ xs, float px,
                                                     float py, float pz, float s_max, float s_relax,
                                                     const lps_dev::NbMasks& M, uint32_t i, Acc&& accept_extra,
                                                     Pre&& pre, F&& fn) {
    const char* __restrict__ xb = reinterpret_cast<const char*>(xs);
    const int xg = (int)truncf((ax + G.off0) / G.gs);
    const int yg = (int)truncf((ay + G.off1) / G.gs);
    const int zg = (int)truncf((az + G.off2) / G.gs);
    const int x0 = max(xg - 1, 0), x1 = min(xg + 1, G.gx - 1);
    const bool replay = M.ok[i] != 0u;
    for (uint32_t rz = 0; rz < 3u; rz++) {
        const int z = zg - 1 + (int)rz;
        if (z < 0 || z >= G.gz) continue;
        uint32_t b[3], m[3];
        bool merge = replay;
#pragma unroll
        for (uint32_t ry = 0; ry < 3u; ry++) {
            const int y = yg - 1 + (int)ry;
            b[ry] = m[ry] = 0u;
            if (y >= 0 && y < G.gy) {
                const uint32_t base = G.base + (uint32_t)G.gx * (uint32_t)y + G.gxy * (uint32_t)z;
                b[ry] = lb[base + (uint32_t)x0];
                const uint32_t n = lb[base + (uint32_t)x1 + 1u] - b[ry];
                if (n > 32u) merge = false;
                else if (n != 0u && replay) m[ry] = M.mask[(3u * rz + ry) * M.stride + i];
            }
        }
        // one choice for the lanes in this plane
        if (__all(merge)) {
            uint32_t m0 = m[0], m1 = m[1], m2 = m[2];
            while (m0 | m1 | m2) {
                uint32_t jj[VB];
                uint32_t cnt = 0u;
#pragma unroll
                for (int t = 0; t < VB; t++) {
                    // the first remaining candidate in row order (none: repeat the trip's first)
                    const bool r0 = m0 != 0u, r1 = m1 != 0u;
                    const uint32_t mm = r0 ? m0 : (r1 ? m1 : m2);
                    const uint32_t bb = r0 ? b[0] : (r1 ? b[1] : b[2]);
                    const uint32_t rest = mm & (mm - 1u);
                    if (r0) m0 = rest;
                    else if (r1) m1 = rest;
                    else m2 = rest;
                    jj[t] = (t > 0 && mm == 0u) ? jj[0] : bb + (uint32_t)__ffs((int)mm) - 1u;
                    cnt += mm != 0u ? 1u : 0u;
                }
                float4 q[VB];
                decltype(pre(0u)) pv[VB];
#pragma unroll
                for (int t = 0; t < VB; t++) {
                    q[t] = *reinterpret_cast<const float4*>(xb + (jj[t] << 4));
                    pv[t] = pre(jj[t]);
                }
#pragma unroll
                for (int t = 0; t < VB; t++) {
                    const uint32_t j = jj[t];
                    const float dx = q[t].x - px, dy = q[t].y - py, dz = q[t].z - pz;
                    const float s2 = dx * dx + dy * dy + dz * dz;
                    visit_slot(fn, (uint32_t)t < cnt && s2 <= s_max && accept_extra(j), j, q[t], s2, pv[t],
                               make_float3(dx, dy, dz));
                }
            }
        } else {
            for (uint32_t ry = 0; ry < 3u; ry++) {
                const int y = yg - 1 + (int)ry;
                if (y < 0 || y >= G.gy) continue;
                const uint32_t base = G.base + (uint32_t)G.gx * (uint32_t)y + G.gxy * (uint32_t)z;
                visit_row<VB>(xb, CompactFilter{}, lb[base + (uint32_t)x0], lb[base + (uint32_t)x1 + 1u],
                              M.mask + (3u * rz + ry) * M.stride, i, replay, false, px, py, pz, s_max, s_relax,
                              accept_extra, pre, fn);
            }
        }
    }
}

// Lookup index min(trunc(sqrtf(s2) / div), 1000) of a squared distance
// s2 <= s_max from the threshold table (see Derived::thr): the hardware
// square root and 1/div give the quotient to within 5e-4; one below that,
// truncated, is the exact index or one less, and the threshold decides.  The
// guess is held below 1000, so the index stays in the table for every s2
// (NaN and +inf included), which the branch-free density visits rely on.
__device__ __forceinline__ uint32_t table_index(float s2, const float* __restrict__ thr, float inv_div) {
    const float qa = __builtin_amdgcn_sqrtf(s2) * inv_div - 1e-3f;
    const uint32_t g = min((uint32_t)fmaxf(qa, 0.f), (uint32_t)(LPS_TABLE_SIZE - 2));
    return g + (s2 >= thr[g + 1u] ? 1u : 0u);
}

// The same index when d = sqrtf(s2) is already at hand (a closer first guess).
__device__ __forceinline__ uint32_t table_index_d(float s2, float d, const float* __restrict__ thr, float inv_div) {
    const float qa = d * inv_div - 1e-3f;
    const uint32_t g = min((uint32_t)fmaxf(qa, 0.f), (uint32_t)(LPS_TABLE_SIZE - 2));
    return g + (s2 >= thr[g + 1u] ? 1u : 0u);
}

// The index trunc(d / div) itself from d = sqrtf(s2), by the shared-reciprocal
// quotient of lps_div.h with inv_div = RN(1 / div): d is 0 or at least 2^-75
// (the square root of a positive float), so the quotient is the IEEE one when
// div lies in [2^-78, 2^40) (index_q_ok, uniform), and d < h keeps the index
// at or below 1000.  Four instructions and no threshold read.
__device__ __forceinline__ bool index_q_ok(float div) { return div >= 0x1p-78f && div < 0x1p40f; }

__device__ __forceinline__ uint32_t table_index_q(float d, float div, float inv_div) {
    return (uint32_t)lps_div::div_rn(d, div, inv_div);
}

// Correctly rounded square root (sqrtf).  For s >= 2^-96 the hardware
// estimate r is within one ulp of sqrt(s) and needs no range scaling; of r
// and its float neighbours lo < r < hi the nearest to sqrt(s) is lo when
// s <= lo * r, hi when s > hi * r, else r (signs of fused residuals).  The
// products are the squares of the geometric means of neighbouring
// candidates; each lies ulp^2 / 4 below the square of the arithmetic
// midpoint it stands for, and no float s falls in that gap, so the choice is
// the rounding to nearest (checked exhaustively over [2^-96, 2^64) for every
// estimate within one ulp by tests/test_sqrt_rn.py).
__device__ __forceinline__ float sqrt_rn(float s) {
    if (s >= 0x1p-96f) {
        const float r = __builtin_amdgcn_sqrtf(s);
        const float lo = __uint_as_float(__float_as_uint(r) - 1u);
        const float hi = __uint_as_float(__float_as_uint(r) + 1u);
        if (__builtin_fmaf(-lo, r, s) <= 0.f) return lo;
        if (__builtin_fmaf(-hi, r, s) > 0.f) return hi;
        return r;
    }
    return sqrtf(s);
}

// ---------------------------------------------------------- colliders --
// resolveCollision (SPHKernels.cl:640-659)
__device__ __forceinline__ void resolve(float3& pos, float3& vel, float e, float mu, float3 n, float3 pt) {
    const float sv = vel.x * n.x + vel.y * n.y + vel.z * n.z;
    if (sv < 0.f) {
        float3 vn = make_float3(n.x * sv, n.y * sv, n.z * sv);
        float3 vt = make_float3(vel.x - vn.x, vel.y - vn.y, vel.z - vn.z);
        vn.x *= -e; vn.y *= -e; vn.z *= -e;
        vt.x *= mu; vt.y *= mu; vt.z *= mu;
        vel = make_float3(vn.x + vt.x, vn.y + vt.y, vn.z + vt.z);
    }
    pos = pt;
}

// handleCollisionWithBox (SPHKernels.cl:488-598); the boundary normal is
// normalised as v * (1/|v|).
__device__ __forceinline__ void collide_box(float3& pos, float3& vel, float r, float e, float mu,
                                            const lps_box& b) {
    const float cx = b.center[0], cy = b.center[1], cz = b.center[2];
    const float hx = b.half[0], hy = b.half[1], hz = b.half[2];
    const float dx = fabsf(pos.x - cx), dy = fabsf(pos.y - cy), dz = fabsf(pos.z - cz);
    bool hit = false;
    float3 n = make_float3(0.f, 0.f, 0.f);
    float3 pt = make_float3(0.f, 0.f, 0.f);
    if (!b.boundary) {
        const float px = (hx - dx) + r, py = (hy - dy) + r, pz = (hz - dz) + r;
        if (px > 0.f && py > 0.f && pz > 0.f) {
            hit = true;
            pt = pos;
            if (px <= py && px <= pz) {
                n.x = (pos.x - cx >= 0.f) ? 1.f : -1.f;
                pt.x = cx + n.x * (hx + r);
            } else if (py <= px && py <= pz) {
                n.y = (pos.y - cy >= 0.f) ? 1.f : -1.f;
                pt.y = cy + n.y * (hy + r);
            } else if (pz <= py && pz <= px) {
                n.z = (pos.z - cz >= 0.f) ? 1.f : -1.f;
                pt.z = cz + n.z * (hz + r);
            }
        }
    } else {
        const float px = (dx - hx) + r, py = (dy - hy) + r, pz = (dz - hz) + r;
        if (px > 0.f || py > 0.f || pz > 0.f) {
            hit = true;
            pt = pos;
            if (px > 0.f) {
                n.x = (pos.x - cx <= 0.f) ? 1.f : -1.f;
                pt.x = cx - n.x * (hx - r);
            }
            if (py > 0.f) {
                n.y = (pos.y - cy <= 0.f) ? 1.f : -1.f;
                pt.y = cy - n.y * (hy - r);
            }
            if (pz > 0.f) {
                n.z = (pos.z - cz <= 0.f) ? 1.f : -1.f;
                pt.z = cz - n.z * (hz - r);
            }
            const float m = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
            if (m > 0.f) {
                const float inv = 1.f / m;
                n.x *= inv; n.y *= inv; n.z *= inv;
            }
        }
    }
    if (hit) resolve(pos, vel, e, mu, n, pt);
}

// handleCollisionWithSphere (SPHKernels.cl:600-638): normal normalised first.
__device__ __forceinline__ void collide_sphere(float3& pos, float3& vel, float r, float e, float mu,
                                               const lps_sphere& s) {
    const float3 c = make_float3(s.center[0], s.center[1], s.center[2]);
    const float R = s.radius;
    const float ddx = c.x - pos.x, ddy = c.y - pos.y, ddz = c.z - pos.z;
    const float d = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
    if (!s.boundary) {
        if ((R + r) - d > 0.f) {
            // a particle exactly at the centre (0/0 in the reference) is pushed along +y
            const float3 n = d > 0.f ? make_float3((pos.x - c.x) / d, (pos.y - c.y) / d, (pos.z - c.z) / d)
                                     : make_float3(0.f, 1.f, 0.f);
            const float k = R + r;
            resolve(pos, vel, e, mu, n, make_float3(c.x + n.x * k, c.y + n.y * k, c.z + n.z * k));
        }
    } else {
        if (d - (R - r) > 0.f) {
            const float3 n = make_float3((c.x - pos.x) / d, (c.y - pos.y) / d, (c.z - pos.z) / d);
            const float k = R - r;
            resolve(pos, vel, e, mu, n, make_float3(c.x - n.x * k, c.y - n.y * k, c.z - n.z * k));
        }
    }
}

// Plane (no reference counterpart): fluid side n.x >= offset; penetration
// r - (n.x - offset); contact point moved along n; resolveCollision response.
__device__ __forceinline__ void collide_plane(float3& pos, float3& vel, float r, float e, float mu,
                                              const lps_plane& pl) {
    const float3 n = make_float3(pl.normal[0], pl.normal[1], pl.normal[2]);
    const float s = (n.x * pos.x + n.y * pos.y + n.z * pos.z) - pl.offset;
    const float pen = r - s;
    if (pen > 0.f)
        resolve(pos, vel, e, mu, n, make_float3(pos.x + n.x * pen, pos.y + n.y * pen, pos.z + n.z * pen));
}

// handleCollisions order: boxes, spheres (SPHKernels.cl:472-480), then planes.
__device__ __forceinline__ void collide_all(float3& pos, float3& vel, const SimConst& c, const lps_box* boxes,
                                            const lps_sphere* spheres, const lps_plane* planes) {
    for (uint32_t j = 0; j < c.nbox; j++) collide_box(pos, vel, c.particle_radius, c.restitution, c.friction, boxes[j]);
    for (uint32_t j = 0; j < c.nsphere; j++)
        collide_sphere(pos, vel, c.particle_radius, c.restitution, c.friction, spheres[j]);
    for (uint32_t j = 0; j < c.nplane; j++)
        collide_plane(pos, vel, c.particle_radius, c.restitution, c.friction, planes[j]);
}

// pciIntegrate + pciHandleCollisions (SPHKernels.cl:676-735): predicted position
// from acceleration (f + fp) / rho; the first-step back-step reads v.
__device__ __forceinline__ float4 predict(float4 x, float4 v, float4 vh, bool first, float fx, float fy, float fz,
                                          float rho, const SimConst& c, const lps_box* boxes,
                                          const lps_sphere* spheres, const lps_plane* planes) {
    const float ax = fx / rho, ay = fy / rho, az = fz / rho;
    float3 hv = make_float3(vh.x, vh.y, vh.z);
    if (first) {
        hv.x = v.x - ax * c.dt / 2.f;
        hv.y = v.y - ay * c.dt / 2.f;
        hv.z = v.z - az * c.dt / 2.f;
    }
    hv.x += ax * c.dt; hv.y += ay * c.dt; hv.z += az * c.dt;
    float3 p = make_float3(x.x + hv.x * c.dt, x.y + hv.y * c.dt, x.z + hv.z * c.dt);
    collide_all(p, hv, c, boxes, spheres, planes);
    return make_float4(p.x, p.y, p.z, 0.f);
}

__device__ __forceinline__ bool step_dead(const DevGrid* g) { return g->status != 0u; }

// Pressure-force direction (x_i - x_j) / d (SPHKernels.cl:406,865), from the
// stencil's differences dl = x_j - x_i: -dl equals x_i - x_j except that an
// exact zero difference comes out as -0.  The three quotients share one
// reciprocal of d (lps_div.h) where that is exact, which also turns a zero
// numerator into +0 as in the reference; the IEEE division elsewhere gives
// -0 for it, and a -0 quotient adds nothing different to the force sums,
// which start at +0 and so never hold -0.  `fastdiv` is the caller's uniform
// bound d < 2^40; NOCHECK says every coordinate is known to be 0 or >= 2^-56
// in magnitude (the numerators, their differences, then need no check, and
// the caller also guarantees the bound).  Two distinct particles at exactly the same point (d == 0;
// the box collider snaps every particle crossing three walls onto one corner
// point) would give 0/0 in the reference; the pair gets the antisymmetric
// unit direction +y for the higher sorted index and -y for the lower one,
// which separates it.
template <bool NOCHECK>
__device__ __forceinline__ void pair_dir(const float3& dl, float d, uint32_t i, uint32_t j, bool fastdiv, float& ux,
                                         float& uy, float& uz) {
    // one branch on the common path (the shared reciprocal implies d > 0)
    const float ax = -dl.x, ay = -dl.y, az = -dl.z;
    bool fast;
    if constexpr (NOCHECK) {
        fast = lps_div::div_rn_ok_d(d);
    } else {
        const uint32_t key = min(min(lps_div::num_key(ax), lps_div::num_key(ay)), lps_div::num_key(az));
        fast = fastdiv && lps_div::div_rn_ok(d, key);
    }
    if (fast) {
        const float y = lps_div::rcp_rn(d);
        ux = lps_div::div_rn(ax, d, y);
        uy = lps_div::div_rn(ay, d, y);
        uz = lps_div::div_rn(az, d, y);
    } else if (d > 0.f) {
        ux = ax / d;
        uy = ay / d;
        uz = az / d;
    } else {
        ux = 0.f;
        uy = i > j ? 1.f : -1.f;
        uz = 0.f;
    }
}

}  // namespace

// ================================================================ grid ====
// Bounding box of the positions (SPHSolver.cpp:1011-1029).  Each block writes
// its partial min/max (orderable-uint encoding) and a non-finite flag; the
// grid set-up kernel reduces the partials, so there is no global atomic.
// The same record carries the block's kinetic and potential energy (no
// counterpart in the reference), summed per scene by the grid set-up.
__device__ __forceinline__ void bbox_partial(uint32_t* __restrict__ part, uint32_t lb, bool valid, float x,
                                             float y, float z, float ke, float pe) {
    __shared__ uint32_t s_red[9][LPS_BLOCK / 64];
    uint32_t mn[3], mx[3], nf = 0;
    const float v[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
        const bool fin = valid && isfinite(v[a]);
        if (valid && !isfinite(v[a])) nf = 1;
        mn[a] = wave_min_u32(fin ? f2o(v[a]) : 0xFFFFFFFFu);
        mx[a] = wave_max_u32(fin ? f2o(v[a]) : 0u);
    }
    nf = wave_max_u32(nf);
    const float wke = wave_sum_f32(valid ? ke : 0.f), wpe = wave_sum_f32(valid ? pe : 0.f);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        for (int a = 0; a < 3; a++) {
            s_red[a][w] = mn[a];
            s_red[3 + a][w] = mx[a];
        }
        s_red[6][w] = nf;
        s_red[7][w] = __float_as_uint(wke);
        s_red[8][w] = __float_as_uint(wpe);
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const uint32_t a = threadIdx.x;
        uint32_t r = (a < 3) ? 0xFFFFFFFFu : 0u;
        for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) r = (a < 3) ? min(r, s_red[a][k]) : max(r, s_red[a][k]);
        part[lb * LPS_BBOX_PART + a] = r;
    } else if (threadIdx.x < 9) {
        float e = 0.f;
        for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) e += __uint_as_float(s_red[threadIdx.x][k]);
        part[lb * LPS_BBOX_PART + threadIdx.x] = __float_as_uint(e);
    }
}

// Kinetic and potential energy of one particle: m |v|^2 / 2 and -m g.x.
__device__ __forceinline__ void particle_energy(float m, const float* g, float x, float y, float z, float vx,
                                                float vy, float vz, float& ke, float& pe) {
    ke = 0.5f * m * (vx * vx + vy * vy + vz * vz);
    pe = -m * (g[0] * x + g[1] * y + g[2] * z);
}

__global__ __launch_bounds__(LPS_BLOCK) void k_bbox(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                    float mass, lps_dev::Scenes S, uint32_t* __restrict__ part,
                                                    uint32_t nlogical) {
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const BlockScene B = block_scene(S, lb);
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    const bool valid = i < B.info.slot0 + B.info.count;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    float ke = 0.f, pe = 0.f;
    if (valid) {
        p = pos[i];
        const float4 v = vel[i];
        particle_energy(mass, B.info.gravity, p.x, p.y, p.z, v.x, v.y, v.z, ke, pe);
    }
    bbox_partial(part, lb, valid, p.x, p.y, p.z, ke, pe);
}

// Kinetic and potential energy of scene sidx from its blocks' partial records,
// summed in double by one block, to out[2 * sidx] and out[2 * sidx + 1]
// (coherent host memory).
__device__ __forceinline__ void scene_energy_block(const lps_dev::Scenes& S, const uint32_t* __restrict__ part,
                                                   uint32_t sidx, double* __restrict__ out) {
    __shared__ double s_e[2][LPS_BLOCK / 64];
    const SceneInfo si = S.info[sidx];
    const uint32_t nparts = si.padded / LPS_BLOCK;
    double ke = 0.0, pe = 0.0;
    // unrolled so that several records' loads are in flight at once (the
    // sums keep their order)
#pragma unroll 8
    for (uint32_t k = threadIdx.x; k < nparts; k += LPS_BLOCK) {
        const uint32_t* q = part + (si.block0 + k) * LPS_BBOX_PART;
        ke += (double)__uint_as_float(q[7]);
        pe += (double)__uint_as_float(q[8]);
    }
    ke = wave_sum_f64(ke);
    pe = wave_sum_f64(pe);
    if (lane_id() == 0) {
        s_e[0][threadIdx.x >> 6] = ke;
        s_e[1][threadIdx.x >> 6] = pe;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < LPS_BLOCK / 64; w++) {
            ke += s_e[0][w];
            pe += s_e[1][w];
        }
        out[2 * sidx] = ke;
        out[2 * sidx + 1] = pe;
    }
    __syncthreads();
}

__global__ __launch_bounds__(LPS_BLOCK) void k_scene_energy(lps_dev::Scenes S, const uint32_t* __restrict__ part,
                                                            double* __restrict__ out) {
    scene_energy_block(S, part, blockIdx.x, out);
}

// Grid parameters of each scene (SPHSolver.cpp:1031-1040), one block per
// scene over the partial boxes of its particle blocks.  min starts at FLT_MAX
// and max at FLT_MIN as in the reference (SPHSolver.cpp:1013-1014).
__device__ __forceinline__ void scene_grid_block(const lps_dev::Scenes& S, const uint32_t* __restrict__ part,
                                                 float h, uint32_t sidx, double* __restrict__ energy) {
    // the energy of the state entering this step, for the host
    if (energy) scene_energy_block(S, part, sidx, energy);
    __shared__ uint32_t s_r[7][LPS_BLOCK / 64];
    const SceneInfo si = S.info[sidx];
    const uint32_t nparts = si.padded / LPS_BLOCK;
    uint32_t r[7];
    for (int a = 0; a < 3; a++) {
        r[a] = f2o(FLT_MAX);
        r[3 + a] = f2o(FLT_MIN);
    }
    r[6] = 0u;
#pragma unroll 8
    for (uint32_t k = threadIdx.x; k < nparts; k += LPS_BLOCK) {
        const uint32_t* q = part + (si.block0 + k) * LPS_BBOX_PART;
        for (int a = 0; a < 3; a++) {
            r[a] = min(r[a], q[a]);
            r[3 + a] = max(r[3 + a], q[3 + a]);
        }
        r[6] |= q[6];
    }
    for (int a = 0; a < 7; a++) r[a] = (a < 3) ? wave_min_u32(r[a]) : wave_max_u32(r[a]);
    if (lane_id() == 0)
        for (int a = 0; a < 7; a++) s_r[a][threadIdx.x >> 6] = r[a];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int a = 0; a < 7; a++)
        for (uint32_t k = 1; k < LPS_BLOCK / 64; k++) r[a] = (a < 3) ? min(r[a], s_r[a][k]) : max(r[a], s_r[a][k]);
    SceneGrid* sg = S.grid + sidx;  // written by thread 0 only
    uint32_t bad = r[6] ? 1u : 0u;
    float mn[3], mx[3];
    const float gs = h * 1.f;
    uint64_t cells = 1;
    for (int a = 0; a < 3; a++) {
        mn[a] = o2f(r[a]);
        mx[a] = o2f(r[3 + a]);
        sg->bbox_min[a] = mn[a];
        sg->bbox_max[a] = mx[a];
        sg->offset[a] = -mn[a];
        const float ex = (mx[a] - mn[a]) / gs;
        if (!(ex < 2.0e9f)) {
            bad = bad ? bad : 2u;
            sg->gsize[a] = 1u;
        } else {
            sg->gsize[a] = (uint32_t)((int)ex + 1);
        }
        cells *= sg->gsize[a];
    }
    if (cells >= 0x7FFFFFFFull) bad = bad ? bad : 2u;
    sg->cells = bad ? 1u : (uint32_t)cells;
    sg->bad = bad;
}

__global__ __launch_bounds__(LPS_BLOCK) void k_scene_grid(lps_dev::Scenes S, const uint32_t* __restrict__ part,
                                                          float h, double* __restrict__ energy) {
    scene_grid_block(S, part, h, blockIdx.x, energy);
}

// Cell numbering across scenes (scene s owns cells [cell_base, cell_base +
// cells] including its empty-slot cell), capacity check, and the per-step
// reset of the loop state.  Thread 0 reads every scene's grid.
__device__ __forceinline__ void grid_finish_block(DevGrid* g, const lps_dev::Scenes& S, float h,
                                                  uint32_t cell_cap) {
    const uint32_t t = threadIdx.x;
    for (uint32_t b = t; b < S.nblocks; b += LPS_BLOCK) S.act_list[b] = b;
    for (uint32_t sidx = t; sidx < S.count; sidx += LPS_BLOCK) {
        SceneGrid* sg = S.grid + sidx;
        sg->stop_iter = 0u;
        sg->best_bits = 0u;
        sg->best_k = 0u;
        sg->best_buf = 0u;
        sg->stall = 0u;
        sg->converged = 0u;
        sg->final_bits = 0u;
        sg->final_buf = 0u;
        sg->buf_of[1] = 0u;
        for (int k = 0; k < LPS_MAX_ITER_SLOTS; k++) sg->err_bits[k] = 0u;
    }
    if (t != 0) return;
    uint32_t status = 0u;
    uint64_t total = 0;
    for (uint32_t sidx = 0; sidx < S.count; sidx++) {
        SceneGrid* sg = S.grid + sidx;
        if (sg->bad == 1u) status = LPS_ERR_NONFINITE;
        if (sg->bad == 2u && !status) status = LPS_ERR_GRID_OVERFLOW;
        sg->cell_base = (uint32_t)min(total, (uint64_t)0xFFFFFFFFull);
        total += (uint64_t)sg->cells + 1u;
    }
    g->spacing = h * 1.f;
    g->cells_needed = total > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total;
    if (!status && total > (uint64_t)cell_cap) status = LPS_ERR_GRID_OVERFLOW;
    g->total_cells = status ? 1u : (uint32_t)total;
    const uint32_t cm1 = g->total_cells - 1u;
    g->key_bits = cm1 ? 32u - (uint32_t)__clz(cm1) : 0u;
    g->passes_done = 0u;
    g->status = status;
    g->running = S.count;
    g->iterations = 0u;
    g->final_err = 0u;
    g->converged = 0u;
    g->tiny_pred = 0u;
    g->act_count[0] = S.nblocks;
    g->act_count[1] = 0u;
    g->list_of[1] = 0u;
    if (!status) g->steps_ok += 1u;
}

__global__ __launch_bounds__(LPS_BLOCK) void k_grid_finish(DevGrid* g, lps_dev::Scenes S, float h,
                                                           uint32_t cell_cap) {
    grid_finish_block(g, S, h, cell_cap);
}

// Both of the above for a single scene, in one block: thread 0 writes the
// scene's grid and then reads it (one launch less per step).
__global__ __launch_bounds__(LPS_BLOCK) void k_grid_one(DevGrid* g, lps_dev::Scenes S,
                                                        const uint32_t* __restrict__ part, float h,
                                                        uint32_t cell_cap, double* __restrict__ energy) {
    scene_grid_block(S, part, h, 0u, energy);
    __syncthreads();
    grid_finish_block(g, S, h, cell_cap);
}

// calcGridIndices (SPHKernels.cl:66-82) within the scene's grid, offset by the
// scene's first cell; empty slots get the scene's empty-slot cell so they sort
// after its particles.
// Cell keys (calcGridIndices), one radix tile of LPS_RADIX_TILE slots per
// block in the radix kernels' item order, and the tile's histogram of the
// first digit (the first radix pass then starts at its scan).
__global__ __launch_bounds__(LPS_BLOCK) void k_keys(const float4* __restrict__ pos, uint32_t n, const DevGrid* g,
                                                    lps_dev::Scenes S, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals, uint32_t* __restrict__ tile_max,
                                                    uint32_t* __restrict__ hist, uint32_t ntiles) {
    if (step_dead(g)) return;
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0u;
    // the cell table's per-tile maxima of the run ends start at zero; the
    // reorder writes those of the non-empty tiles
    const uint32_t ncell_tiles = (g->total_cells + LPS_SCAN_TILE) / LPS_SCAN_TILE;
    for (uint32_t b = blockIdx.x * LPS_BLOCK + threadIdx.x; b < ncell_tiles; b += gridDim.x * LPS_BLOCK)
        tile_max[b] = 0u;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    // the tile's four positions are loaded before any key is formed
    float4 p[LPS_RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t i = tile * LPS_RADIX_TILE + r * LPS_BLOCK + threadIdx.x;
        p[r] = i < n ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // the scene and grid of each of the tile's particle blocks (slot i lies
    // in block i / LPS_BLOCK, uniform over a round), all loaded before the
    // first key is formed
    // (blocks past the last one read the last one's, unused)
    uint32_t sidx[LPS_RADIX_ITEMS], kend[LPS_RADIX_ITEMS], kempty[LPS_RADIX_ITEMS];
    GridView Gr[LPS_RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++)
        sidx[r] = S.block_scene[min((tile * LPS_RADIX_TILE + r * LPS_BLOCK) / LPS_BLOCK, S.nblocks - 1u)];
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const SceneGrid* sg = S.grid + sidx[r];
        kend[r] = S.info[sidx[r]].slot0 + S.info[sidx[r]].count;
        kempty[r] = sg->cell_base + sg->cells;
        Gr[r] = grid_view(g, sg);
    }
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t i = tile * LPS_RADIX_TILE + r * LPS_BLOCK + threadIdx.x;
        const bool valid = i < n;
        uint32_t key = 0u;
        if (valid) {
            key = kempty[r];
            if (i < kend[r]) {
                const GridView& G = Gr[r];
                const uint32_t xg = (uint32_t)truncf((p[r].x + G.off0) / G.gs);
                const uint32_t yg = (uint32_t)truncf((p[r].y + G.off1) / G.gs);
                const uint32_t zg = (uint32_t)truncf((p[r].z + G.off2) / G.gs);
                key = G.base + xg + (uint32_t)G.gx * yg + G.gxy * zg;
            }
            keys[i] = key;
            vals[i] = i;
        }
        const uint32_t d = key & 0xFFu;
        const uint64_t act = __ballot(valid);
        const uint64_t peers = match_digit(d, act);
        if (valid && lane_id() == (uint32_t)__ffsll((long long)peers) - 1u) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
}

// ---- stable LSD radix sort by key, 8-bit digits, reduce-then-scan ----------
// A pass is skipped when its digit is above the key width of this step's
// cell count; skipped passes leave the result in the buffer given by the
// parity of passes_done.
__global__ __launch_bounds__(LPS_BLOCK) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                          uint32_t shift, const DevGrid* g,
                                                          uint32_t* __restrict__ hist, uint32_t ntiles) {
    if (step_dead(g) || shift >= g->key_bits) return;
    __shared__ uint32_t cnt[256];
    const uint32_t tile = blockIdx.x;
    // the tile's keys load before the counters are cleared
    uint32_t kr[LPS_RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t idx = tile * LPS_RADIX_TILE + r * LPS_BLOCK + threadIdx.x;
        kr[r] = idx < n ? keys[idx] : 0u;
    }
    cnt[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t idx = tile * LPS_RADIX_TILE + r * LPS_BLOCK + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t d = valid ? (kr[r] >> shift) & 0xFFu : 0u;
        const uint64_t act = __ballot(valid);
        const uint64_t peers = match_digit(d, act);
        if (valid && lane_id() == (uint32_t)__ffsll((long long)peers) - 1u) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
}

// Exclusive scan of each digit's per-tile counts (one block per digit); the
// digit totals are scanned by every scatter block.  Together this is the
// bucket-major scan of the reference (SPHSolver.cpp:1069-1078).
__global__ __launch_bounds__(LPS_BLOCK) void k_radix_scan(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                          uint32_t* __restrict__ digit_total, DevGrid* g,
                                                          uint32_t shift) {
    if (step_dead(g) || shift >= g->key_bits) return;
    __shared__ uint32_t wsum[LPS_BLOCK / 64];
    const uint32_t d = blockIdx.x, t = threadIdx.x, lane = lane_id(), w = t >> 6;
    uint32_t* row = hist + d * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += LPS_BLOCK) {
        const uint32_t k = base + t;
        const uint32_t v = k < ntiles ? row[k] : 0u;
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
            if (lane >= (uint32_t)off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (uint32_t q = 0; q < LPS_BLOCK / 64; q++) {
            if (q < w) wpre += wsum[q];
            tot += wsum[q];
        }
        if (k < ntiles) row[k] = carry + wpre + incl - v;
        carry += tot;
        __syncthreads();
    }
    if (t == 0) {
        digit_total[d] = carry;
        if (d == 0) g->passes_done += 1u;
    }
}

// Stable scatter of one tile: items are taken in index order (round-major,
// then wave, then lane); ranks within a wave come from the digit match mask,
// across waves and rounds from per-digit running offsets in LDS.
__global__ __launch_bounds__(LPS_BLOCK) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin,
                                                             uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                             uint32_t n, uint32_t shift, const DevGrid* g,
                                                             const uint32_t* __restrict__ hist, uint32_t ntiles,
                                                             const uint32_t* __restrict__ digit_total) {
    if (step_dead(g) || shift >= g->key_bits) return;
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_w[LPS_BLOCK / 64][256];
    __shared__ uint32_t s_ws[LPS_BLOCK / 64];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = lane_id();
    const uint32_t tile = blockIdx.x;
    // every load of the block issues before its first barrier: the digit
    // totals, this tile's offsets and its items
    const uint32_t v = digit_total[t];
    const uint32_t tile_off = hist[t * ntiles + tile];
    uint32_t kr[LPS_RADIX_ITEMS], vr[LPS_RADIX_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t idx = tile * LPS_RADIX_TILE + r * LPS_BLOCK + t;
        kr[r] = idx < n ? kin[idx] : 0u;
        vr[r] = idx < n ? vin[idx] : 0u;
    }
    {
        // exclusive scan of the 256 digit totals
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
            if (lane >= (uint32_t)off) incl += y;
        }
        if (lane == 63) s_ws[w] = incl;
        __syncthreads();
        uint32_t wpre = 0;
        for (uint32_t q = 0; q < w; q++) wpre += s_ws[q];
        s_off[t] = tile_off + wpre + incl - v;
    }
    for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) s_w[k][t] = 0u;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
    for (uint32_t r = 0; r < LPS_RADIX_ITEMS; r++) {
        const uint32_t idx = tile * LPS_RADIX_TILE + r * LPS_BLOCK + t;
        const bool valid = idx < n;
        const uint32_t key = kr[r], val = vr[r];
        const uint32_t d = (key >> shift) & 0xFFu;
        const uint64_t act = __ballot(valid);
        const uint64_t peers = match_digit(d, act);
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0u) s_w[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t base = s_off[t];
            for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) {
                const uint32_t c = s_w[k][t];
                s_w[k][t] = base;
                base += c;
            }
            s_off[t] = base;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dst = s_w[w][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) s_w[k][t] = 0u;
        __syncthreads();
    }
}

// Gather the state into sorted order (the payload moves of permuteParticles,
// SPHKernels.cl:108-146, done once) and mark the end of every run of equal
// keys for the cell table.  On a failed step the copy is the identity so the
// state survives the buffer flip.
__global__ __launch_bounds__(LPS_BLOCK) void k_reorder(
    const uint32_t* __restrict__ keysA, const uint32_t* __restrict__ valsA, const uint32_t* __restrict__ keysB,
    const uint32_t* __restrict__ valsB, uint32_t n, const DevGrid* g, const float4* __restrict__ pos_in,
    const float4* __restrict__ vel_in, const float4* __restrict__ vh_in, const uint8_t* __restrict__ first_in,
    const uint32_t* __restrict__ id_in, const float* __restrict__ pprev_in, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, float4* __restrict__ vh_out, uint8_t* __restrict__ first_out,
    uint32_t* __restrict__ id_out, float* __restrict__ pprev_out, uint32_t* __restrict__ run_end,
    uint32_t* __restrict__ skeys_out, uint32_t* __restrict__ perm_out, lps_dev::Scenes S, uint2* __restrict__ cq,
    uint32_t* __restrict__ tile_max, uint32_t nlogical) {
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    if (i >= n) return;
    const bool dead = step_dead(g);
    uint32_t j = i;
    if (!dead) {
        const bool b = (g->passes_done & 1u) != 0u;
        const uint32_t* keys = b ? keysB : keysA;
        const uint32_t* vals = b ? valsB : valsA;
        j = vals[i];
        const uint32_t k = keys[i];
        skeys_out[i] = k;
        perm_out[i] = j;
        if (i == n - 1u || keys[i + 1u] != k) {
            run_end[k] = i + 1u;
            // the last run of a cell-table tile holds the tile's largest run
            // end (keys are sorted): one plain store per non-empty tile
            if (i == n - 1u || keys[i + 1u] / LPS_SCAN_TILE != k / LPS_SCAN_TILE) tile_max[k / LPS_SCAN_TILE] = i + 1u;
        }
    }
    const float4 p = pos_in[j];
    pos_out[i] = p;
    if (!dead) {
        // the density pass filters candidates on this copy (slots of one
        // block belong to one scene)
        const BlockScene B = block_scene(S, lb);
        const GridView G = grid_view(g, S.grid + B.s);
        cq[i] = quantize(p.x, p.y, p.z, G, compact_scale(G));
    }
    vel_out[i] = vel_in[j];
    vh_out[i] = vh_in[j];
    first_out[i] = first_in[j];
    id_out[i] = id_in[j];
    if (pprev_in) pprev_out[i] = pprev_in[j];
}

// Cell table = exclusive max-scan of the run ends: LB[c] = #keys < c, the
// cellList of buildCellList (SPHKernels.cl:148-186), with LB[cells] = n.  The
// run-end marks are cleared as they are consumed.  block_max holds each
// tile's largest run end (0 for an empty tile), written by the reorder after
// k_keys cleared it.
__global__ __launch_bounds__(LPS_BLOCK) void k_cell_scan(uint32_t* __restrict__ run_end, const DevGrid* g,
                                                         const uint32_t* __restrict__ block_max,
                                                         uint32_t* __restrict__ cell_lb) {
    if (step_dead(g)) return;
    const uint32_t total = g->total_cells + 1u;
    const uint32_t base = blockIdx.x * LPS_SCAN_TILE;
    if (base >= total) return;
    __shared__ uint32_t s[LPS_BLOCK / 64];
    __shared__ uint32_t s_carry;
    const uint32_t t = threadIdx.x, lane = lane_id(), w = t >> 6;
    // carry = max of the previous blocks
    uint32_t c0 = 0;
    for (uint32_t k = t; k < blockIdx.x; k += LPS_BLOCK) c0 = max(c0, block_max[k]);
    c0 = wave_max_u32(c0);
    if (lane == 0) s[w] = c0;
    __syncthreads();
    if (t == 0) {
        uint32_t r = 0;
        for (uint32_t k = 0; k < LPS_BLOCK / 64; k++) r = max(r, s[k]);
        s_carry = r;
    }
    __syncthreads();
    uint32_t carry = s_carry;
    // thread t owns cells base + t*ITEMS .. +ITEMS (contiguous)
    const uint32_t c_begin = base + t * LPS_SCAN_ITEMS;
    uint32_t v[LPS_SCAN_ITEMS];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t k = 0; k < LPS_SCAN_ITEMS; k++) {
        const uint32_t c = c_begin + k;
        v[k] = (c < total) ? run_end[c] : 0u;
        local = max(local, v[k]);
    }
    // exclusive max-scan of `local` across the block
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
        if (lane >= (uint32_t)off) incl = max(incl, y);
    }
    uint32_t excl = (uint32_t)__shfl_up((int)incl, 1, 64);
    if (lane == 0) excl = 0u;
    __syncthreads();
    if (lane == 63) s[w] = incl;
    __syncthreads();
    uint32_t wprefix = 0;
    for (uint32_t k = 0; k < w; k++) wprefix = max(wprefix, s[k]);
    uint32_t run = max(carry, max(wprefix, excl));
#pragma unroll
    for (uint32_t k = 0; k < LPS_SCAN_ITEMS; k++) {
        const uint32_t c = c_begin + k;
        if (c < total) {
            cell_lb[c] = run;
            if (v[k]) run_end[c] = 0u;
        }
        run = max(run, v[k]);
    }
}

// ============================================================= physics ====
// calcDensityPressure (SPHKernels.cl:190-254).  Writes (x, rho) packed for the
// force pass, the pressure, and q = p / rho^2 (the per-particle term of the
// SPH pressure force, SPHKernels.cl:383,412).
__global__ __launch_bounds__(LPS_BLOCK) void k_density(const float4* __restrict__ pos, const DevGrid* g,
                                                       lps_dev::Scenes S, const uint32_t* __restrict__ cell_lb,
                                                       SimConst c, const float* __restrict__ tables,
                                                       float4* __restrict__ posrho, float* __restrict__ press,
                                                       float* __restrict__ q, lps_dev::NbMasks mk,
                                                       uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    // the slot's own state loads while the tables are staged (every slot of
    // a block lies within the allocated slots); the reorder wrote the compact
    // copy of these positions
    const float4 xi = pos[i];
    const uint2 qi = mk.cq[i];
    const BlockScene B = block_scene(S, lb);
    __shared__ float sW[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_table<LPS_BLOCK>(sW, tables, sT, tables);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    float sum = 0.f;
    const float qs = compact_scale(G);
    CompactFilter cf{};
    if (compact_near(xi.x, xi.y, xi.z, G, qs)) cf = CompactFilter{mk.cq, qi, compact_bias(c.s_max, qs)};
    for_neighbors_masked<2>(G, cell_lb, xi.x, xi.y, xi.z, pos, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i, MC_RECORD,
                         cf, [](uint32_t) { return true; }, [](uint32_t) { return 0; },
                         predicated([&](bool ok, uint32_t, const float4&, float s2, int, const float3&) {
                             const float w = sW[table_index(s2, sT, c.inv_div)];
                             sum += ok ? w : 0.f;
                         }));
    const float rho = c.mass * sum;
    float p = c.stiffness * (rho - c.rest_density);
    if (p < 0.f) p *= c.npf;
    posrho[i] = make_float4(xi.x, xi.y, xi.z, rho);
    press[i] = p;
    q[i] = p / (rho * rho);
}

// accumulateNonPressureForces (SPHKernels.cl:256-349); for PCISPH also the
// first prediction of the correction loop (pciInit + pciIntegrate +
// pciHandleCollisions with fp = 0, SPHKernels.cl:661-735).  Held one wave per
// SIMD above what the compiler picks (7 and 6 waves: 102.3 and 110.1 µs
// against 103.3 and 113.4 at 524k); at eight it spills inside the plane loop.
template <bool FEAT>
__global__ __launch_bounds__(LPS_BLOCK) __attribute__((amdgpu_waves_per_eu(FEAT ? 6 : 7))) void k_forces(
    const float4* __restrict__ posrho, const float4* __restrict__ vel, const float4* __restrict__ vh,
    const uint8_t* __restrict__ first, DevGrid* g, const uint32_t* __restrict__ cell_lb, SimConst c0,
    const float* __restrict__ tables, const lps_box* __restrict__ boxes0, const lps_sphere* __restrict__ spheres0,
    const lps_plane* __restrict__ planes0, float4* __restrict__ force, float4* __restrict__ xpred,
    const float* __restrict__ pprev, float* __restrict__ feat, lps_dev::Scenes S, lps_dev::NbMasks mk,
    uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    // the slot's own state loads while the tables are staged
    const float4 xi = posrho[i];
    const float4 vi = vel[i];
    const BlockScene B = block_scene(S, lb);
    __shared__ float sW1[LPS_TABLE_SIZE], sW2[LPS_TABLE_SIZE], sV2[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_force_tables<LPS_BLOCK>(sW1, sW2, sV2, sT, tables);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const SimConst c = scene_const(c0, B.info);
    const lps_box* boxes = boxes0 + B.info.box0;
    const lps_sphere* spheres = spheres0 + B.info.sphere0;
    const lps_plane* planes = planes0 + B.info.plane0;
    const float ri = xi.w;
    float fx = c.gravity[0] * ri, fy = c.gravity[1] * ri, fz = c.gravity[2] * ri;
    float snx = 0.f, sny = 0.f, snz = 0.f, lc = 0.f, vfx = 0.f, vfy = 0.f, vfz = 0.f;
    float nb = 0.f, dv = 0.f;  // warm-start features: neighbour count, velocity divergence
    // the density pass recorded this stencil's masks at the same positions
    for_neighbors_planes<2>(G, cell_lb, xi.x, xi.y, xi.z, posrho, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                         [](uint32_t) { return true; }, [&](uint32_t j) { return vel[j]; },
                         [&](uint32_t j, const float4& xj, float s2, const float4& vj, const float3&) {
        const uint32_t idx = table_index(s2, sT, c.inv_div);
        const float rj = xj.w;
        const float w1 = sW1[idx], wv = sV2[idx];
        // the numerators of every quotient by rho_j (the self pair's velocity
        // terms are zero and unused)
        float n[8] = {(xi.x - xj.x) * w1, (xi.y - xj.y) * w1, (xi.z - xj.z) * w1, sW2[idx],
                      (vj.x - vi.x) * wv, (vj.y - vi.y) * wv, (vj.z - vi.z) * wv, 0.f};
        constexpr int NQ = FEAT ? 8 : 7;
        if constexpr (FEAT)
            n[7] = ((vj.x - vi.x) * (xi.x - xj.x) + (vj.y - vi.y) * (xi.y - xj.y) + (vj.z - vi.z) * (xi.z - xj.z)) * w1;
        uint32_t key = lps_div::num_key(n[0]);
        for (int t = 1; t < NQ; t++) key = min(key, lps_div::num_key(n[t]));
        if (rj < 0x1p40f && lps_div::div_rn_ok(rj, key)) {
            const float y = lps_div::rcp_rn(rj);
            for (int t = 0; t < NQ; t++) n[t] = lps_div::div_rn(n[t], rj, y);
        } else {
            for (int t = 0; t < NQ; t++) n[t] = n[t] / rj;
        }
        snx += n[0];
        sny += n[1];
        snz += n[2];
        lc += n[3];
        if (i != j) {
            vfx += n[4];
            vfy += n[5];
            vfz += n[6];
            if constexpr (FEAT) {
                nb += 1.f;
                dv += n[7];
            }
        }
    });
    snx *= c.mass; sny *= c.mass; snz *= c.mass;
    const float snl = sqrtf(snx * snx + sny * sny + snz * snz);
    if constexpr (FEAT) {
        // feature-major: feature f of slot i at f * stride + i (coalesced stores)
        float* fr = feat + i;
        const size_t fs = mk.stride;
        fr[0 * fs] = ri / c.rest_density - 1.f;
        fr[1 * fs] = nb * (1.f / 32.f);
        fr[2 * fs] = dv * c.mass * c.dt;
        fr[3 * fs] = snx * c.h;
        fr[4 * fs] = sny * c.h;
        fr[5 * fs] = snz * c.h;
        fr[6 * fs] = snl * c.h;
        fr[7 * fs] = pprev[i] / (c.delta * c.rest_density);
        fr[8 * fs] = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z) * c.dt / c.h;
        fr[9 * fs] = first[i] ? 1.f : 0.f;
    }
    if (snl > c.st_threshold) {
        const float k = -c.surface_tension * lc * c.mass;
        fx += (snx / snl) * k;
        fy += (sny / snl) * k;
        fz += (snz / snl) * k;
    }
    fx += vfx * c.viscosity * c.mass;
    fy += vfy * c.viscosity * c.mass;
    fz += vfz * c.viscosity * c.mass;
    force[i] = make_float4(fx, fy, fz, 0.f);
    if (xpred) {
        const float4 xo = make_float4(xi.x, xi.y, xi.z, 0.f);
        const float4 xp = predict(xo, vi, vh[i], first[i] != 0, fx + 0.f, fy + 0.f, fz + 0.f, ri, c, boxes, spheres,
                                  planes);
        xpred[i] = xp;
        mk.cq[i] = quantize(xp.x, xp.y, xp.z, G, compact_scale(G));
        if (lps_div::tiny_coord(xp.x, xp.y, xp.z)) g->tiny_pred = 1u;
        if constexpr (FEAT) {
            const float ex = xp.x - xo.x, ey = xp.y - xo.y, ez = xp.z - xo.z;
            feat[(size_t)10 * mk.stride + i] = sqrtf(ex * ex + ey * ey + ez * ez) / c.h;
        }
    }
}

// accumulatePressureForces (SPHKernels.cl:351-422), SPH only.
__global__ __launch_bounds__(LPS_BLOCK) void k_pressure_sph(const float4* __restrict__ posrho,
                                                            const float* __restrict__ q, const DevGrid* g,
                                                            const uint32_t* __restrict__ cell_lb, SimConst c,
                                                            const float* __restrict__ tables,
                                                            float4* __restrict__ force, lps_dev::Scenes S,
                                                            lps_dev::NbMasks mk, uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    __shared__ float sP1[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_table<LPS_BLOCK>(sP1, tables + 3 * LPS_TABLE_SIZE, sT, tables);
    __syncthreads();
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    const BlockScene B = block_scene(S, lb);
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const float4 xi = posrho[i];
    const float tf = q[i];
    const bool fastdiv = c.s_max < 0x1p78f;  // accepted pairs have d < 2^40
    float pfx = 0.f, pfy = 0.f, pfz = 0.f;
    auto pressure = [&](auto qidx) {
        for_neighbors_planes<2>(G, cell_lb, xi.x, xi.y, xi.z, posrho, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                             [&](uint32_t j) { return i != j; }, [&](uint32_t j) { return q[j]; },
                             [&](uint32_t j, const float4&, float s2, float qj, const float3& dl) {
            const float d = sqrt_rn(s2);
            float ux, uy, uz;
            pair_dir<false>(dl, d, i, j, fastdiv, ux, uy, uz);
            const float w = sP1[decltype(qidx)::value ? table_index_q(d, c.div, c.inv_div)
                                                      : table_index_d(s2, d, sT, c.inv_div)];
            const float sc = tf + qj;
            pfx += ux * sc * w;
            pfy += uy * sc * w;
            pfz += uz * sc * w;
        });
    };
    if (index_q_ok(c.div))
        pressure(std::true_type{});
    else
        pressure(std::false_type{});
    float4 f = force[i];
    f.x += (-pfx) * c.mass * xi.w;
    f.y += (-pfy) * c.mass * xi.w;
    f.z += (-pfz) * c.mass * xi.w;
    force[i] = f;
}

// ===================================================== quad stencils ====
// Small scenes leave most of the chip idle while every lane walks a long
// neighbour list alone: the 16k SPH dam break's compressed fluid gives a
// particle hundreds of candidates, and its 63 particle blocks fill a quarter
// of the SIMDs with one wave each.  The quad kernels give a particle the four
// lanes of a quad.  Pass 1 tests every fourth candidate of a window on each
// lane, and the quad ORs the four masks.  Pass 2 hands the quad's next four
// kept candidates to its four lanes in order; each lane computes the terms
// of its candidate, and every lane of the quad then adds the four terms in
// candidate order (DPP quad broadcasts), so each accumulator takes exactly the
// additions of the one-lane walk.  A rejected candidate contributes +0: the
// accumulators start at +0, a float sum is -0 only when both operands are -0,
// so they never hold -0 and adding +0 leaves them unchanged.  The four lanes
// of a quad share one particle, so they are active together throughout.
constexpr uint32_t LPS_QUAD = 4u;

__device__ __forceinline__ uint32_t quad_sub() { return threadIdx.x & (LPS_QUAD - 1u); }

template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, false));
}

__device__ __forceinline__ uint32_t quad_or(uint32_t m) {
    m |= (uint32_t)__builtin_amdgcn_mov_dpp((int)m, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    m |= (uint32_t)__builtin_amdgcn_mov_dpp((int)m, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    return m;
}

// Pass 1 over one window of n <= 32 candidates from w0, candidates sub,
// sub + 4, ... on lane sub; the quad's accept mask on every lane.
__device__ __forceinline__ uint32_t quad_filter(const char* __restrict__ xb, const CompactFilter CF, uint32_t w0,
                                                uint32_t n, float px, float py, float pz, float s_relax) {
    const uint32_t sub = quad_sub();
    uint32_t bits = 0u;
    if (CF.cq) {
        const char* __restrict__ cb = reinterpret_cast<const char*>(CF.cq);
        uint2 e[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t u = sub + LPS_QUAD * (uint32_t)t;
            e[t] = CF.qi;
            if (u < n) e[t] = *reinterpret_cast<const uint2*>(cb + ((w0 + u) << 3));
        }
#pragma unroll
        for (int t = 0; t < 8; t++) bits |= (compact_diff(e[t], CF.qi, CF.bias) >> 31) << (sub + LPS_QUAD * t);
    } else {
        float4 q[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t u = sub + LPS_QUAD * (uint32_t)t;
            q[t] = u < n ? *reinterpret_cast<const float4*>(xb + ((w0 + u) << 4)) : make_float4(px, py, pz, 0.f);
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float dx = q[t].x - px, dy = q[t].y - py, dz = q[t].z - pz;
            bits |= (__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) <= s_relax ? 1u : 0u)
                    << (sub + LPS_QUAD * t);
        }
    }
    return quad_or(bits) & (n == 32u ? 0xFFFFFFFFu : (1u << n) - 1u);
}

// The quad's next trip from `mask` (advanced past it): this lane's candidate
// j (w0 when it has none: a valid index to load), whether it has one, and
// how many of the quad's lanes have one (0: the window is done).
struct QuadTrip {
    uint32_t j, cnt;
    bool have;
};
__device__ __forceinline__ QuadTrip quad_trip(uint32_t w0, uint32_t& mask) {
    const uint32_t sub = quad_sub();
    uint32_t mm = mask;
    if (sub > 0u) mm &= mm - 1u;
    if (sub > 1u) mm &= mm - 1u;
    if (sub > 2u) mm &= mm - 1u;
    QuadTrip r;
    r.cnt = min((uint32_t)__popc(mask), LPS_QUAD);
#pragma unroll
    for (uint32_t t = 0; t < LPS_QUAD; t++) mask &= mask - 1u;
    r.have = mm != 0u;
    r.j = r.have ? w0 + (uint32_t)__ffs((int)mm) - 1u : w0;
    return r;
}

// Pass 2 of one window for a quad: `term(ok, j, x_j, s2, pre(j), x_j - x_i, t)`
// writes the NT terms of candidate j (zeros when not ok), and acc[a] takes
// the terms of the quad's candidates in order.  The loads of the next trip
// issue before this trip's terms are computed: a small scene has about one
// wave per SIMD, which would otherwise wait out every gather.
template <int NT, class Acc, class Pre, class Term>
__device__ __forceinline__ void quad_visit(const char* __restrict__ xb, uint32_t w0, uint32_t mask, float px, float py,
                                           float pz, float s_max, Acc& accept_extra, Pre& pre, Term& term,
                                           float (&acc)[NT]) {
    QuadTrip tr = quad_trip(w0, mask);
    float4 q = *reinterpret_cast<const float4*>(xb + (tr.j << 4));
    auto pv = pre(tr.j);
    while (tr.cnt) {
        const QuadTrip nx = quad_trip(w0, mask);
        const float4 qn = *reinterpret_cast<const float4*>(xb + (nx.j << 4));
        const auto pvn = pre(nx.j);
        const uint32_t j = tr.j, cnt = tr.cnt;
        const bool have = tr.have;
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float s2 = dx * dx + dy * dy + dz * dz;
        float t[NT];
        term(have && s2 <= s_max && accept_extra(j), j, q, s2, pv, make_float3(dx, dy, dz), t);
#pragma unroll
        for (int a = 0; a < NT; a++) acc[a] += quad_bcast<0>(t[a]);
        if (cnt > 1u) {
#pragma unroll
            for (int a = 0; a < NT; a++) acc[a] += quad_bcast<1>(t[a]);
        }
        if (cnt > 2u) {
#pragma unroll
            for (int a = 0; a < NT; a++) acc[a] += quad_bcast<2>(t[a]);
        }
        if (cnt > 3u) {
#pragma unroll
            for (int a = 0; a < NT; a++) acc[a] += quad_bcast<3>(t[a]);
        }
        tr = nx;
        q = qn;
        pv = pvn;
    }
}

// for_neighbors_masked for a quad (same rows, windows, recorded masks and
// visiting order; the recording lane is the quad's first).
template <int NT, class Acc, class Pre, class Term>
__device__ __forceinline__ void for_neighbors_quad(const GridView& G, const uint32_t* __restrict__ lb, float ax,
                                                   float ay, float az, const float4* __restrict__ xs, float px,
                                                   float py, float pz, float s_max, float s_relax,
                                                   const lps_dev::NbMasks& M, uint32_t i, int mode,
                                                   const CompactFilter& CF, Acc&& accept_extra, Pre&& pre,
                                                   Term&& term, float (&acc)[NT]) {
    const char* __restrict__ xb = reinterpret_cast<const char*>(xs);
    const int xg = (int)truncf((ax + G.off0) / G.gs);
    const int yg = (int)truncf((ay + G.off1) / G.gs);
    const int zg = (int)truncf((az + G.off2) / G.gs);
    const int x0 = max(xg - 1, 0), x1 = min(xg + 1, G.gx - 1);
    const bool replay = mode == MC_REPLAY && M.ok[i] != 0u;
    const bool record = mode == MC_RECORD && quad_sub() == 0u;
    for (uint32_t rz = 0; rz < 3u; rz++) {
        const int z = zg - 1 + (int)rz;
        if (z < 0 || z >= G.gz) continue;
        for (uint32_t ry = 0; ry < 3u; ry++) {
            const int y = yg - 1 + (int)ry;
            if (y < 0 || y >= G.gy) continue;
            uint32_t* __restrict__ row = M.mask + (3u * rz + ry) * M.stride;
            const uint32_t base = G.base + (uint32_t)G.gx * (uint32_t)y + G.gxy * (uint32_t)z;
            const uint32_t b = lb[base + (uint32_t)x0];
            const uint32_t e = lb[base + (uint32_t)x1 + 1u];
            uint32_t w0 = b;
            if (replay) {
                quad_visit<NT>(xb, b, row[i], px, py, pz, s_max, accept_extra, pre, term, acc);
                w0 += 32u;
            }
            uint32_t first = 0u;
            for (; w0 < e; w0 += 32u) {
                const uint32_t mask = quad_filter(xb, CF, w0, min(32u, e - w0), px, py, pz, s_relax);
                if (w0 == b) first = mask;
                quad_visit<NT>(xb, w0, mask, px, py, pz, s_max, accept_extra, pre, term, acc);
            }
            if (record) row[i] = first;
        }
    }
    if (record) M.ok[i] = 1u;
}

// The particle of a quad kernel's thread: launch block lb covers LPS_BLOCK /
// 4 particles, four lanes each.
__device__ __forceinline__ uint32_t quad_particle(uint32_t lb) {
    return lb * (LPS_BLOCK / LPS_QUAD) + threadIdx.x / LPS_QUAD;
}

// k_density with a quad per particle.
__global__ __launch_bounds__(LPS_BLOCK) void k_density_q(const float4* __restrict__ pos, const DevGrid* g,
                                                         lps_dev::Scenes S, const uint32_t* __restrict__ cell_lb,
                                                         SimConst c, const float* __restrict__ tables,
                                                         float4* __restrict__ posrho, float* __restrict__ press,
                                                         float* __restrict__ q, lps_dev::NbMasks mk,
                                                         uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    // the slot's own state loads while the tables are staged (every slot of
    // a block lies within the allocated slots)
    const uint32_t i = quad_particle(lb);
    const float4 xi = pos[i];
    const uint2 qi = mk.cq[i];
    const BlockScene B = block_scene(S, i / LPS_BLOCK);
    __shared__ float sW[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_table<LPS_BLOCK>(sW, tables, sT, tables);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const float qs = compact_scale(G);
    CompactFilter cf{};
    if (compact_near(xi.x, xi.y, xi.z, G, qs)) cf = CompactFilter{mk.cq, qi, compact_bias(c.s_max, qs)};
    float acc[1] = {0.f};
    for_neighbors_quad<1>(G, cell_lb, xi.x, xi.y, xi.z, pos, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i, MC_RECORD,
                          cf, [](uint32_t) { return true; }, [](uint32_t) { return 0; },
                          [&](bool ok, uint32_t, const float4&, float s2, int, const float3&, float (&t)[1]) {
                              t[0] = ok ? sW[table_index(s2, sT, c.inv_div)] : 0.f;
                          },
                          acc);
    if (quad_sub() != 0u) return;
    const float rho = c.mass * acc[0];
    float p = c.stiffness * (rho - c.rest_density);
    if (p < 0.f) p *= c.npf;
    posrho[i] = make_float4(xi.x, xi.y, xi.z, rho);
    press[i] = p;
    q[i] = p / (rho * rho);
}

// k_forces with a quad per particle.
template <bool FEAT>
__global__ __launch_bounds__(LPS_BLOCK) void k_forces_q(
    const float4* __restrict__ posrho, const float4* __restrict__ vel, const float4* __restrict__ vh,
    const uint8_t* __restrict__ first, DevGrid* g, const uint32_t* __restrict__ cell_lb, SimConst c0,
    const float* __restrict__ tables, const lps_box* __restrict__ boxes0, const lps_sphere* __restrict__ spheres0,
    const lps_plane* __restrict__ planes0, float4* __restrict__ force, float4* __restrict__ xpred,
    const float* __restrict__ pprev, float* __restrict__ feat, lps_dev::Scenes S, lps_dev::NbMasks mk,
    uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    // the slot's own state loads while the tables are staged
    const uint32_t i = quad_particle(lb);
    const float4 xi = posrho[i];
    const float4 vi = vel[i];
    const BlockScene B = block_scene(S, i / LPS_BLOCK);
    __shared__ float sW1[LPS_TABLE_SIZE], sW2[LPS_TABLE_SIZE], sV2[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_force_tables<LPS_BLOCK>(sW1, sW2, sV2, sT, tables);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const SimConst c = scene_const(c0, B.info);
    const float ri = xi.w;
    // snx, sny, snz, lc, vfx, vfy, vfz, then (features) nb, dv
    constexpr int NT = FEAT ? 9 : 7;
    float acc[NT];
#pragma unroll
    for (int a = 0; a < NT; a++) acc[a] = 0.f;
    for_neighbors_quad<NT>(G, cell_lb, xi.x, xi.y, xi.z, posrho, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                           MC_REPLAY, CompactFilter{}, [](uint32_t) { return true; },
                           [&](uint32_t j) { return vel[j]; },
                           [&](bool ok, uint32_t j, const float4& xj, float s2, const float4& vj, const float3&,
                               float (&t)[NT]) {
        if (!ok) {
#pragma unroll
            for (int a = 0; a < NT; a++) t[a] = 0.f;
            return;
        }
        const uint32_t idx = table_index(s2, sT, c.inv_div);
        const float rj = xj.w;
        const float w1 = sW1[idx], wv = sV2[idx];
        float n[8] = {(xi.x - xj.x) * w1, (xi.y - xj.y) * w1, (xi.z - xj.z) * w1, sW2[idx],
                      (vj.x - vi.x) * wv, (vj.y - vi.y) * wv, (vj.z - vi.z) * wv, 0.f};
        constexpr int NQ = FEAT ? 8 : 7;
        if constexpr (FEAT)
            n[7] = ((vj.x - vi.x) * (xi.x - xj.x) + (vj.y - vi.y) * (xi.y - xj.y) + (vj.z - vi.z) * (xi.z - xj.z)) * w1;
        uint32_t key = lps_div::num_key(n[0]);
        for (int a = 1; a < NQ; a++) key = min(key, lps_div::num_key(n[a]));
        if (rj < 0x1p40f && lps_div::div_rn_ok(rj, key)) {
            const float y = lps_div::rcp_rn(rj);
            for (int a = 0; a < NQ; a++) n[a] = lps_div::div_rn(n[a], rj, y);
        } else {
            for (int a = 0; a < NQ; a++) n[a] = n[a] / rj;
        }
        const bool other = i != j;
        t[0] = n[0];
        t[1] = n[1];
        t[2] = n[2];
        t[3] = n[3];
        t[4] = other ? n[4] : 0.f;
        t[5] = other ? n[5] : 0.f;
        t[6] = other ? n[6] : 0.f;
        if constexpr (FEAT) {
            t[7] = other ? 1.f : 0.f;
            t[8] = other ? n[7] : 0.f;
        }
    }, acc);
    if (quad_sub() != 0u) return;
    float snx = acc[0], sny = acc[1], snz = acc[2];
    const float lc = acc[3], vfx = acc[4], vfy = acc[5], vfz = acc[6];
    float fx = c.gravity[0] * ri, fy = c.gravity[1] * ri, fz = c.gravity[2] * ri;
    snx *= c.mass; sny *= c.mass; snz *= c.mass;
    const float snl = sqrtf(snx * snx + sny * sny + snz * snz);
    if constexpr (FEAT) {
        float* fr = feat + i;
        const size_t fs = mk.stride;
        fr[0 * fs] = ri / c.rest_density - 1.f;
        fr[1 * fs] = acc[7] * (1.f / 32.f);
        fr[2 * fs] = acc[8] * c.mass * c.dt;
        fr[3 * fs] = snx * c.h;
        fr[4 * fs] = sny * c.h;
        fr[5 * fs] = snz * c.h;
        fr[6 * fs] = snl * c.h;
        fr[7 * fs] = pprev[i] / (c.delta * c.rest_density);
        fr[8 * fs] = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z) * c.dt / c.h;
        fr[9 * fs] = first[i] ? 1.f : 0.f;
    }
    if (snl > c.st_threshold) {
        const float k = -c.surface_tension * lc * c.mass;
        fx += (snx / snl) * k;
        fy += (sny / snl) * k;
        fz += (snz / snl) * k;
    }
    fx += vfx * c.viscosity * c.mass;
    fy += vfy * c.viscosity * c.mass;
    fz += vfz * c.viscosity * c.mass;
    force[i] = make_float4(fx, fy, fz, 0.f);
    if (xpred) {
        const float4 xo = make_float4(xi.x, xi.y, xi.z, 0.f);
        const float4 xp = predict(xo, vi, vh[i], first[i] != 0, fx + 0.f, fy + 0.f, fz + 0.f, ri, c,
                                  boxes0 + B.info.box0, spheres0 + B.info.sphere0, planes0 + B.info.plane0);
        xpred[i] = xp;
        mk.cq[i] = quantize(xp.x, xp.y, xp.z, G, compact_scale(G));
        if (lps_div::tiny_coord(xp.x, xp.y, xp.z)) g->tiny_pred = 1u;
        if constexpr (FEAT) {
            const float ex = xp.x - xo.x, ey = xp.y - xo.y, ez = xp.z - xo.z;
            feat[(size_t)10 * mk.stride + i] = sqrtf(ex * ex + ey * ey + ez * ez) / c.h;
        }
    }
}

// k_pressure_sph with a quad per particle.
__global__ __launch_bounds__(LPS_BLOCK) void k_pressure_sph_q(const float4* __restrict__ posrho,
                                                              const float* __restrict__ q, const DevGrid* g,
                                                              const uint32_t* __restrict__ cell_lb, SimConst c,
                                                              const float* __restrict__ tables,
                                                              float4* __restrict__ force, lps_dev::Scenes S,
                                                              lps_dev::NbMasks mk, uint32_t nlogical) {
    if (step_dead(g)) return;
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    // the slot's own state loads while the tables are staged
    const uint32_t i = quad_particle(lb);
    const float4 xi = posrho[i];
    const float tf = q[i];
    const BlockScene B = block_scene(S, i / LPS_BLOCK);
    __shared__ float sP1[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_table<LPS_BLOCK>(sP1, tables + 3 * LPS_TABLE_SIZE, sT, tables);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const bool fastdiv = c.s_max < 0x1p78f;  // accepted pairs have d < 2^40
    const bool qidx = index_q_ok(c.div);
    float acc[3] = {0.f, 0.f, 0.f};
    for_neighbors_quad<3>(G, cell_lb, xi.x, xi.y, xi.z, posrho, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                          MC_REPLAY, CompactFilter{}, [&](uint32_t j) { return i != j; },
                          [&](uint32_t j) { return q[j]; },
                          [&](bool ok, uint32_t j, const float4&, float s2, float qj, const float3& dl, float (&t)[3]) {
        if (!ok) {
            t[0] = t[1] = t[2] = 0.f;
            return;
        }
        const float d = sqrt_rn(s2);
        float ux, uy, uz;
        pair_dir<false>(dl, d, i, j, fastdiv, ux, uy, uz);
        const float w = sP1[qidx ? table_index_q(d, c.div, c.inv_div) : table_index_d(s2, d, sT, c.inv_div)];
        const float sc = tf + qj;
        t[0] = ux * sc * w;
        t[1] = uy * sc * w;
        t[2] = uz * sc * w;
    }, acc);
    if (quad_sub() != 0u) return;
    float4 f = force[i];
    f.x += (-acc[0]) * c.mass * xi.w;
    f.y += (-acc[1]) * c.mass * xi.w;
    f.z += (-acc[2]) * c.mass * xi.w;
    force[i] = f;
}

// pciCalcDensityPressure (SPHKernels.cl:737-804) + the density-error maximum
// of correction k (max_i |e_i|, e_i = rho*_i - rho0 scaled by npf when the
// pressure update is negative; PCISPHSolver.cpp:113-121,152-156).  Cells come
// from the current position, distances from the predicted ones.  Each block
// adds its maximum to the scene's error of correction k; block 0 of the
// pressure pass that follows applies the loop rule (loop_rule_block).
template <bool FEAT, int WG>
__global__ __launch_bounds__(WG) void k_pci_density(
    const float4* __restrict__ pos, const float4* __restrict__ xpred, DevGrid* g, const uint32_t* __restrict__ cell_lb,
    SimConst c, const float* __restrict__ tables, float* __restrict__ rho_pred, float* __restrict__ press,
    float4* __restrict__ xq, uint32_t k, float* __restrict__ feat, const float4* __restrict__ posrho,
    lps_dev::Scenes S, lps_dev::NbMasks mk, uint32_t nlogical) {
    const uint32_t slot = (k - 1u) % LPS_MAX_ITER_SLOTS;
    // a correction queued past the stop, or of a dead step, does nothing
    if (step_dead(g) || g->running == 0u) return;
    // only blocks of scenes whose loop is running are scheduled
    const uint32_t L = g->list_of[k % LPS_MAX_ITER_SLOTS];
    uint32_t s0 = 0u;
    const uint32_t lb = active_piece<WG>(g, S, L, s0);
    if (lb == UINT_MAX) return;
    __shared__ float sW[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    __shared__ uint32_t s_max[WG / 64];
    const bool active = true;
    const uint32_t i = active ? s0 + threadIdx.x : UINT_MAX;
    // the slot's own state loads while the tables are staged (every slot of
    // an active piece lies within the allocated slots)
    const float4 xa = pos[i];
    const float4 xi = xpred[i];
    const uint2 qi = mk.cq[i];
    const BlockScene B = block_scene(S, lb);
    stage_table<WG>(sW, tables, sT, tables);
    __syncthreads();
    uint32_t ebits = 0;
    if (active && i < B.info.slot0 + B.info.count) {
        const GridView G = grid_view(g, S.grid + B.s);
        float sum = 0.f;
        const float qs = compact_scale(G);
        CompactFilter cf{};
        if (compact_near(xi.x, xi.y, xi.z, G, qs)) cf = CompactFilter{mk.cq, qi, compact_bias(c.s_max, qs)};
        // recorded for the pressure pass of this correction (same positions and cells)
        for_neighbors_masked<2>(G, cell_lb, xa.x, xa.y, xa.z, xpred, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                             MC_RECORD, cf, [](uint32_t) { return true; }, [](uint32_t) { return 0; },
                             predicated([&](bool ok, uint32_t, const float4&, float s2, int, const float3&) {
                                 const float w = sW[table_index(s2, sT, c.inv_div)];
                                 sum += ok ? w : 0.f;
                             }));
        const float rp = c.mass * sum;
        float e = rp - c.rest_density;
        float dp = c.delta * e;
        if (dp < 0.f) {
            dp *= c.npf;
            e *= c.npf;
        }
        // this correction's buffer (chosen by the previous loop rule) and
        // the previous correction's pressure
        const SceneGrid* sg = S.grid + B.s;
        const size_t ob = (size_t)sg->buf_of[k % LPS_MAX_ITER_SLOTS] * mk.stride;
        const float pold = k == 1u ? 0.f : press[(size_t)sg->buf_of[(k - 1u) % LPS_MAX_ITER_SLOTS] * mk.stride + i];
        const float p = pold + dp;
        rho_pred[ob + i] = rp;
        press[ob + i] = p;
        // the pressure pass gathers the prediction and p / rho*^2 in one load
        xq[i] = make_float4(xi.x, xi.y, xi.z, p / (rp * rp));
        if (FEAT && k == 1u) {
            float* fr = feat + i;  // feature-major, as in k_forces
            const size_t fs = mk.stride;
            const float e1 = rp / c.rest_density - 1.f;
            fr[11 * fs] = e1;
            fr[12 * fs] = fmaxf(e1, 0.f);
            fr[13 * fs] = dp / (c.delta * c.rest_density);
            fr[14 * fs] = (rp - posrho[i].w) / c.rest_density;
            fr[15 * fs] = e1 * e1 * 100.f;
        }
        ebits = __float_as_uint(fabsf(e));
    }
    ebits = wave_max_u32(ebits);
    if (lane_id() == 0) s_max[threadIdx.x >> 6] = ebits;
    __syncthreads();
    // the scene's maximum error of this correction; the loop rule reads it
    // after the kernel boundary (no fence here: a device-scope fence writes
    // back and invalidates this XCD's L2 under every block still running)
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < WG / 64; w++) m = max(m, s_max[w]);
        if (active && m) atomicMax(&S.grid[B.s].err_bits[slot], m);
    }
}

// The loop rule of correction k (1-based), run by one whole block (block 0 of
// the pressure pass, before its own particles) once the predicted-density
// pass has ended: stop a scene when k >= min and its maximum error ratio is
// below the tolerance (converged), when k >= min and the maximum error has not
// improved on its best value for `stall_limit` corrections, or at k == max; a
// NaN error marks the step non-finite.  Builds the next correction's block
// list and writes {state, corrections, error bits, converged, particle blocks
// of the next correction} to the host slot (state 0 continue, 1 stopped, 2
// error).  The rest of the pressure pass reads
// none of what it writes except stop_iter (a scene stopping at k has
// stop_iter = k, which the pass treats like 0) and the step status.  A
// correction queued past the stop answers nothing: the host never waits on
// it, and it may run after the host has marked the same slot for the next
// step.  A dead step is answered (the host may be waiting on any correction of
// it) and the host drains the stream after it.
__device__ void loop_rule_block(DevGrid* g, const lps_dev::Scenes& S, uint32_t k, const LoopRule& rule,
                                float rest_density, uint32_t* __restrict__ host_flag) {
    if (step_dead(g) || g->running == 0u) {
        if (step_dead(g) && threadIdx.x == 0) {
            host_flag[1] = g->iterations;
            host_flag[2] = g->final_err;
            host_flag[3] = g->converged;
            // release only: the host polls word 0, then reads the others (an
            // acquire would also invalidate this XCD's L2 under the next pass)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            host_flag[0] = 2u;
        }
        return;
    }
    const uint32_t slot = (k - 1u) % LPS_MAX_ITER_SLOTS;
    const uint32_t L = g->list_of[k % LPS_MAX_ITER_SLOTS];
    __shared__ uint32_t s_run, s_err, s_conv, s_nan, s_nb;
    if (threadIdx.x == 0) {
        s_run = 0u;
        s_err = 0u;
        s_conv = 1u;
        s_nan = 0u;
    }
    __syncthreads();
    // one thread per scene
    for (uint32_t sidx = threadIdx.x; sidx < S.count; sidx += blockDim.x) {
        SceneGrid* sg = S.grid + sidx;
        if (sg->stop_iter != 0u) {
            atomicMax(&s_err, sg->final_bits);
            if (!sg->converged) atomicAnd(&s_conv, 0u);
            continue;
        }
        const uint32_t eb = sg->err_bits[slot];
        // the density pass of correction k + 1 (after this kernel) starts
        // from a cleared slot, also once the ring has wrapped
        sg->err_bits[k % LPS_MAX_ITER_SLOTS] = 0u;
        sg->final_bits = eb;
        if (eb >= 0x7F800000u) {
            atomicOr(&s_nan, 1u);
            continue;
        }
        const uint32_t cur = sg->buf_of[k % LPS_MAX_ITER_SLOTS];
        // stall bookkeeping from the second correction on (the first one's
        // predictions carry no pressure force); non-negative floats order
        // like their bit patterns.  Correction k's predictions came from the
        // buffer of correction k - 1.
        if (k >= 2u) {
            if (sg->best_k == 0u || eb < sg->best_bits) {
                sg->best_bits = eb;
                sg->best_k = k;
                sg->best_buf = sg->buf_of[(k - 1u) % LPS_MAX_ITER_SLOTS];
                sg->stall = 0u;
            } else {
                sg->stall += 1u;
            }
        }
        const float ratio = __uint_as_float(eb) / rest_density;
        bool stop = false, stalled = false;
        if (k >= rule.min_iter && ratio < rule.tol) {
            stop = true;
            sg->converged = 1u;
        } else if (k >= rule.min_iter && rule.stall_limit > 0u && sg->stall >= rule.stall_limit) {
            stop = stalled = true;
        } else if (k >= rule.max_iter) {
            stop = true;
        }
        uint32_t reported = eb;
        if (stop) {
            // a stall exit applies the pressure force whose predictions had
            // the smallest error (stall >= 1, so that correction is an
            // earlier one); any other stop applies this correction's force
            sg->stop_iter = k;
            sg->final_buf = stalled ? sg->best_buf : cur;
            if (stalled) reported = sg->best_bits;
            sg->final_bits = reported;
        } else {
            // correction k + 1 writes the buffer that is neither this
            // correction's nor the best one's
            const uint32_t keep = sg->best_k ? sg->best_buf : cur;
            uint32_t nb = 0u;
            while (nb == cur || nb == keep) nb++;
            sg->buf_of[(k + 1u) % LPS_MAX_ITER_SLOTS] = (uint8_t)nb;
            atomicAdd(&s_run, 1u);
        }
        atomicMax(&s_err, reported);
        if (!sg->converged) atomicAnd(&s_conv, 0u);
    }
    __syncthreads();
    // the pressure pass of this correction uses this list; the next
    // correction uses the blocks of the scenes still running
    uint32_t next = L;
    if (s_run != g->running) {
        const uint32_t N = 1u - L;
        if (threadIdx.x == 0) s_nb = 0u;
        __syncthreads();
        for (uint32_t sidx = 0; sidx < S.count; sidx++) {
            if (S.grid[sidx].stop_iter != 0u) continue;
            const SceneInfo& si = S.info[sidx];
            const uint32_t nb = si.padded / LPS_BLOCK, at = s_nb;
            for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) S.act_list[N * S.nblocks + at + b] = si.block0 + b;
            __syncthreads();
            if (threadIdx.x == 0) s_nb = at + nb;
            __syncthreads();
        }
        if (threadIdx.x == 0) g->act_count[N] = s_nb;
        next = N;
    }
    if (threadIdx.x != 0) return;
    g->list_of[(k + 1u) % LPS_MAX_ITER_SLOTS] = next;
    uint32_t state = 0u;
    if (s_nan) {
        g->status = LPS_ERR_NONFINITE;
        g->steps_ok -= 1u;  // counted by the grid setup; this step did not complete
        state = 2u;
    } else if (s_run == 0u) {
        state = 1u;
    }
    g->running = s_nan ? 0u : s_run;
    g->iterations = k;
    g->final_err = s_err;
    g->converged = s_conv;
    host_flag[1] = k;
    host_flag[2] = s_err;
    host_flag[3] = s_conv;
    host_flag[4] = g->act_count[next];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // as above: words 1-3 before word 0
    host_flag[0] = state;
}

// pciCalcPressureForce (SPHKernels.cl:806-878) followed by the next
// prediction (pciIntegrate + pciHandleCollisions) from f + fp.  Held to eight
// waves per SIMD (64 VGPRs): the compiler then spills a few values that live
// across the whole stencil, outside its loops, and the pass runs 61.7 µs per
// launch at 524k particles against 68.7 µs at the seven waves it would pick.
template <int WG>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(WG >= 128 ? 8 : 1))) void k_pci_pressure(
    const float4* __restrict__ pos, const float4* __restrict__ xq, const float* __restrict__ rho_pred, DevGrid* g,
    const uint32_t* __restrict__ cell_lb, SimConst c0,
    const float* __restrict__ tables, float4* __restrict__ fp, const float4* __restrict__ force,
    const float4* __restrict__ posrho, const float4* __restrict__ vel, const float4* __restrict__ vh,
    const uint8_t* __restrict__ first, const lps_box* __restrict__ boxes0, const lps_sphere* __restrict__ spheres0,
    const lps_plane* __restrict__ planes0, float4* __restrict__ xpred_next, uint32_t k, lps_dev::Scenes S,
    lps_dev::NbMasks mk, LoopRule rule, uint32_t* __restrict__ host_flag, uint32_t nlogical) {
    if (blockIdx.x == 0) loop_rule_block(g, S, k, rule, c0.rest_density, host_flag);
    if (step_dead(g)) return;
    // the blocks of the scenes that ran this correction's predicted-density pass
    uint32_t s0 = 0u;
    const uint32_t lb = active_piece<WG>(g, S, g->list_of[k % LPS_MAX_ITER_SLOTS], s0);
    if (lb == UINT_MAX) return;
    const uint32_t i = s0 + threadIdx.x;
    // the slot's own state loads while the tables are staged (every slot of
    // an active piece lies within the allocated slots); the predicted-density
    // pass's (prediction, p / rho*^2)
    const float4 xa = pos[i];
    const float4 xi = xq[i];
    const BlockScene B = block_scene(S, lb);
    const uint32_t stop = S.grid[B.s].stop_iter;
    if (stop != 0u && stop < k) return;
    // accepted pairs have d < 2^40; the predictions' coordinates were checked
    // when they were written; the common case also takes the table index as
    // the quotient d / div and needs no thresholds (one instantiation of the
    // stencil per case, so the pair loop carries no mode test)
    const bool fastdiv = c0.s_max < 0x1p78f;
    const bool quick = fastdiv && g->tiny_pred == 0u && index_q_ok(c0.div);
    __shared__ float sP1[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    stage_table<WG>(sP1, tables + 3 * LPS_TABLE_SIZE, sT, tables, !quick);
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const SimConst c = scene_const(c0, B.info);
    const lps_box* boxes = boxes0 + B.info.box0;
    const lps_sphere* spheres = spheres0 + B.info.sphere0;
    const lps_plane* planes = planes0 + B.info.plane0;
    const float tf = xi.w;
    float pfx = 0.f, pfy = 0.f, pfz = 0.f;
    auto pressure = [&](auto nocheck) {
        // the predicted-density pass of this correction recorded the masks
        for_neighbors_planes<2>(G, cell_lb, xa.x, xa.y, xa.z, xq, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                             [&](uint32_t j) { return i != j; }, [](uint32_t) { return 0; },
                             [&](uint32_t j, const float4& xqj, float s2, int, const float3& dl) {
            const float qj = xqj.w;
            const float d = sqrt_rn(s2);
            float ux, uy, uz;
            pair_dir<decltype(nocheck)::value>(dl, d, i, j, fastdiv, ux, uy, uz);
            const float w = sP1[decltype(nocheck)::value ? table_index_q(d, c.div, c.inv_div)
                                                         : table_index_d(s2, d, sT, c.inv_div)];
            const float sc = tf + qj;
            pfx += ux * sc * w;
            pfy += uy * sc * w;
            pfz += uz * sc * w;
        });
    };
    if (quick)
        pressure(std::true_type{});
    else
        pressure(std::false_type{});
    const size_t ob = (size_t)S.grid[B.s].buf_of[k % LPS_MAX_ITER_SLOTS] * mk.stride;
    const float ri = rho_pred[ob + i];
    const float4 f4 = make_float4((-pfx) * c.mass * ri, (-pfy) * c.mass * ri, (-pfz) * c.mass * ri, 0.f);
    fp[ob + i] = f4;
    const float4 f = force[i];
    const float4 xr = posrho[i];
    const float4 xn = predict(make_float4(xr.x, xr.y, xr.z, 0.f), vel[i], vh[i], first[i] != 0, f.x + f4.x,
                              f.y + f4.y, f.z + f4.z, xr.w, c, boxes, spheres, planes);
    xpred_next[i] = xn;
    mk.cq[i] = quantize(xn.x, xn.y, xn.z, G, compact_scale(G));
    if (lps_div::tiny_coord(xn.x, xn.y, xn.z)) g->tiny_pred = 1u;
}

// k_pci_density with a quad per particle: a workgroup of 256 threads takes a
// 64-slot piece of an active particle block.
template <bool FEAT>
__global__ __launch_bounds__(LPS_BLOCK) void k_pci_density_q(
    const float4* __restrict__ pos, const float4* __restrict__ xpred, DevGrid* g, const uint32_t* __restrict__ cell_lb,
    SimConst c, const float* __restrict__ tables, float* __restrict__ rho_pred, float* __restrict__ press,
    float4* __restrict__ xq, uint32_t k, float* __restrict__ feat, const float4* __restrict__ posrho,
    lps_dev::Scenes S, lps_dev::NbMasks mk, uint32_t nlogical) {
    const uint32_t slot = (k - 1u) % LPS_MAX_ITER_SLOTS;
    if (step_dead(g) || g->running == 0u) return;
    const uint32_t L = g->list_of[k % LPS_MAX_ITER_SLOTS];
    uint32_t s0 = 0u;
    const uint32_t lb = active_piece<LPS_BLOCK / LPS_QUAD>(g, S, L, s0);
    if (lb == UINT_MAX) return;
    __shared__ float sW[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    __shared__ uint32_t s_max[LPS_BLOCK / 64];
    // the slot's own state loads while the tables are staged (every slot of
    // an active piece lies within the allocated slots)
    const uint32_t i = s0 + threadIdx.x / LPS_QUAD;
    const float4 xa = pos[i];
    const float4 xi = xpred[i];
    const uint2 qi = mk.cq[i];
    const BlockScene B = block_scene(S, lb);
    stage_table<LPS_BLOCK>(sW, tables, sT, tables);
    __syncthreads();
    uint32_t ebits = 0;
    if (i < B.info.slot0 + B.info.count) {
        const GridView G = grid_view(g, S.grid + B.s);
        const float qs = compact_scale(G);
        CompactFilter cf{};
        if (compact_near(xi.x, xi.y, xi.z, G, qs)) cf = CompactFilter{mk.cq, qi, compact_bias(c.s_max, qs)};
        float acc[1] = {0.f};
        for_neighbors_quad<1>(G, cell_lb, xa.x, xa.y, xa.z, xpred, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i,
                              MC_RECORD, cf, [](uint32_t) { return true; }, [](uint32_t) { return 0; },
                              [&](bool ok, uint32_t, const float4&, float s2, int, const float3&, float (&t)[1]) {
                                  t[0] = ok ? sW[table_index(s2, sT, c.inv_div)] : 0.f;
                              },
                              acc);
        if (quad_sub() == 0u) {
            const float rp = c.mass * acc[0];
            float e = rp - c.rest_density;
            float dp = c.delta * e;
            if (dp < 0.f) {
                dp *= c.npf;
                e *= c.npf;
            }
            const SceneGrid* sg = S.grid + B.s;
            const size_t ob = (size_t)sg->buf_of[k % LPS_MAX_ITER_SLOTS] * mk.stride;
            const float pold =
                k == 1u ? 0.f : press[(size_t)sg->buf_of[(k - 1u) % LPS_MAX_ITER_SLOTS] * mk.stride + i];
            const float p = pold + dp;
            rho_pred[ob + i] = rp;
            press[ob + i] = p;
            xq[i] = make_float4(xi.x, xi.y, xi.z, p / (rp * rp));
            if (FEAT && k == 1u) {
                float* fr = feat + i;
                const size_t fs = mk.stride;
                const float e1 = rp / c.rest_density - 1.f;
                fr[11 * fs] = e1;
                fr[12 * fs] = fmaxf(e1, 0.f);
                fr[13 * fs] = dp / (c.delta * c.rest_density);
                fr[14 * fs] = (rp - posrho[i].w) / c.rest_density;
                fr[15 * fs] = e1 * e1 * 100.f;
            }
            ebits = __float_as_uint(fabsf(e));
        }
    }
    ebits = wave_max_u32(ebits);
    if (lane_id() == 0) s_max[threadIdx.x >> 6] = ebits;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < LPS_BLOCK / 64; w++) m = max(m, s_max[w]);
        if (m) atomicMax(&S.grid[B.s].err_bits[slot], m);
    }
}

// k_pci_pressure with a quad per particle (rows walked one by one).
__global__ __launch_bounds__(LPS_BLOCK) void k_pci_pressure_q(
    const float4* __restrict__ pos, const float4* __restrict__ xq, const float* __restrict__ rho_pred, DevGrid* g,
    const uint32_t* __restrict__ cell_lb, SimConst c0, const float* __restrict__ tables, float4* __restrict__ fp,
    const float4* __restrict__ force, const float4* __restrict__ posrho, const float4* __restrict__ vel,
    const float4* __restrict__ vh, const uint8_t* __restrict__ first, const lps_box* __restrict__ boxes0,
    const lps_sphere* __restrict__ spheres0, const lps_plane* __restrict__ planes0, float4* __restrict__ xpred_next,
    uint32_t k, lps_dev::Scenes S, lps_dev::NbMasks mk, LoopRule rule, uint32_t* __restrict__ host_flag,
    uint32_t nlogical) {
    if (blockIdx.x == 0) loop_rule_block(g, S, k, rule, c0.rest_density, host_flag);
    if (step_dead(g)) return;
    uint32_t s0 = 0u;
    const uint32_t lb = active_piece<LPS_BLOCK / LPS_QUAD>(g, S, g->list_of[k % LPS_MAX_ITER_SLOTS], s0);
    if (lb == UINT_MAX) return;
    // the slot's own state loads while the tables are staged; the
    // predicted-density pass's (prediction, p / rho*^2)
    const uint32_t i = s0 + threadIdx.x / LPS_QUAD;
    const float4 xa = pos[i];
    const float4 xi = xq[i];
    const BlockScene B = block_scene(S, lb);
    const uint32_t stop = S.grid[B.s].stop_iter;
    if (stop != 0u && stop < k) return;
    const bool fastdiv = c0.s_max < 0x1p78f;
    const bool nocheck = fastdiv && g->tiny_pred == 0u;
    const bool qidx = index_q_ok(c0.div);
    __shared__ float sP1[LPS_TABLE_SIZE], sT[LPS_THR_SIZE];
    // the thresholds serve only the pairs without the quotient index
    stage_table<LPS_BLOCK>(sP1, tables + 3 * LPS_TABLE_SIZE, sT, tables, !(nocheck && qidx));
    __syncthreads();
    if (i >= B.info.slot0 + B.info.count) return;
    const GridView G = grid_view(g, S.grid + B.s);
    const SimConst c = scene_const(c0, B.info);
    const float tf = xi.w;
    float acc[3] = {0.f, 0.f, 0.f};
    for_neighbors_quad<3>(G, cell_lb, xa.x, xa.y, xa.z, xq, xi.x, xi.y, xi.z, c.s_max, c.s_relax, mk, i, MC_REPLAY,
                          CompactFilter{}, [&](uint32_t j) { return i != j; }, [](uint32_t) { return 0; },
                          [&](bool ok, uint32_t j, const float4& xqj, float s2, int, const float3& dl, float (&t)[3]) {
        if (!ok) {
            t[0] = t[1] = t[2] = 0.f;
            return;
        }
        const float d = sqrt_rn(s2);
        float ux, uy, uz;
        if (nocheck)
            pair_dir<true>(dl, d, i, j, fastdiv, ux, uy, uz);
        else
            pair_dir<false>(dl, d, i, j, fastdiv, ux, uy, uz);
        const float w = sP1[nocheck && qidx ? table_index_q(d, c.div, c.inv_div) : table_index_d(s2, d, sT, c.inv_div)];
        const float sc = tf + xqj.w;
        t[0] = ux * sc * w;
        t[1] = uy * sc * w;
        t[2] = uz * sc * w;
    }, acc);
    if (quad_sub() != 0u) return;
    const size_t ob = (size_t)S.grid[B.s].buf_of[k % LPS_MAX_ITER_SLOTS] * mk.stride;
    const float ri = rho_pred[ob + i];
    const float4 f4 = make_float4((-acc[0]) * c.mass * ri, (-acc[1]) * c.mass * ri, (-acc[2]) * c.mass * ri, 0.f);
    fp[ob + i] = f4;
    const float4 f = force[i];
    const float4 xr = posrho[i];
    const float4 xn = predict(make_float4(xr.x, xr.y, xr.z, 0.f), vel[i], vh[i], first[i] != 0, f.x + f4.x,
                              f.y + f4.y, f.z + f4.z, xr.w, c, boxes0 + B.info.box0, spheres0 + B.info.sphere0,
                              planes0 + B.info.plane0);
    xpred_next[i] = xn;
    mk.cq[i] = quantize(xn.x, xn.y, xn.z, G, compact_scale(G));
    if (lps_div::tiny_coord(xn.x, xn.y, xn.z)) g->tiny_pred = 1u;
}

// integrate (SPHKernels.cl:424-455) + pciAddPressureForce (SPHKernels.cl:880-890)
// + handleCollisions (SPHKernels.cl:457-486); accumulates the bounding box of
// the new positions for the next step.
__global__ __launch_bounds__(LPS_BLOCK) void k_integrate(
    float4* __restrict__ pos, float4* __restrict__ vel, float4* __restrict__ vh, uint8_t* __restrict__ first,
    const float4* __restrict__ force, const float4* __restrict__ fp, const float4* __restrict__ posrho,
    float* __restrict__ pprev, const float* __restrict__ press, DevGrid* g, SimConst c0,
    const lps_box* __restrict__ boxes0, const lps_sphere* __restrict__ spheres0,
    const lps_plane* __restrict__ planes0, uint32_t* __restrict__ bbox_part, lps_dev::Scenes S, uint32_t stride,
    uint32_t nlogical) {
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const bool dead = step_dead(g);
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    // the slot's own state loads before its scene's constants arrive (every
    // slot of a block lies within the allocated slots)
    const float4 x4 = pos[i];
    const float4 f4 = force[i];
    const float rho = posrho[i].w;
    const float4 h4 = vh[i];
    const bool is_first = first[i] != 0;
    const BlockScene B = block_scene(S, lb);
    const bool valid = i < B.info.slot0 + B.info.count;
    const SimConst c = scene_const(c0, B.info);
    const lps_box* boxes = boxes0 + B.info.box0;
    const lps_sphere* spheres = spheres0 + B.info.sphere0;
    const lps_plane* planes = planes0 + B.info.plane0;
    float3 p = make_float3(0.f, 0.f, 0.f);
    float ke = 0.f, pe = 0.f;
    if (valid) {
        p = make_float3(x4.x, x4.y, x4.z);
        if (!dead) {
            // PCISPH: the pressure force and pressure of the buffer the
            // scene's loop chose; SPH: no pressure force here, buffer 0
            const size_t ob = fp ? (size_t)S.grid[B.s].final_buf * stride : 0u;
            float4 f = f4;
            if (fp) {
                const float4 a = fp[ob + i];
                f.x += a.x; f.y += a.y; f.z += a.z;
            }
            const float ax = f.x / rho, ay = f.y / rho, az = f.z / rho;
            float3 hv = make_float3(h4.x, h4.y, h4.z);
            if (is_first) {
                const float4 v4 = vel[i];
                hv.x = v4.x - ax * c.dt / 2.f;
                hv.y = v4.y - ay * c.dt / 2.f;
                hv.z = v4.z - az * c.dt / 2.f;
                first[i] = 0;
            }
            const float3 old = hv;
            hv.x += ax * c.dt; hv.y += ay * c.dt; hv.z += az * c.dt;
            p.x += hv.x * c.dt; p.y += hv.y * c.dt; p.z += hv.z * c.dt;
            collide_all(p, hv, c, boxes, spheres, planes);
            vh[i] = make_float4(hv.x, hv.y, hv.z, 0.f);
            pos[i] = make_float4(p.x, p.y, p.z, 0.f);
            const float4 vn = make_float4((old.x + hv.x) / 2.f, (old.y + hv.y) / 2.f, (old.z + hv.z) / 2.f, 0.f);
            vel[i] = vn;
            particle_energy(c.mass, c.gravity, p.x, p.y, p.z, vn.x, vn.y, vn.z, ke, pe);
            if (pprev) pprev[i] = press[ob + i];
        } else {
            const float4 vo = vel[i];
            particle_energy(c.mass, c.gravity, p.x, p.y, p.z, vo.x, vo.y, vo.z, ke, pe);
        }
    }
    bbox_partial(bbox_part, lb, valid, p.x, p.y, p.z, ke, pe);
}

// Render interop: the real particles' positions, in readback order (scenes
// back to back, each in its sorted order), into a caller's device buffer
// (a mapped vertex buffer, a tensor) at a stride of `stride` floats; the
// readback of SPHSolver::onEndUpdate (SPHSolver.cpp:398-428) without a host copy.
__global__ __launch_bounds__(LPS_BLOCK) void k_export_positions(const float4* __restrict__ pos, lps_dev::Scenes S,
                                                                float* __restrict__ dst, uint32_t stride,
                                                                uint32_t nlogical) {
    const uint32_t lb = xcd_block(nlogical);
    if (lb == UINT_MAX) return;
    const BlockScene B = block_scene(S, lb);
    const uint32_t i = lb * LPS_BLOCK + threadIdx.x;
    if (i >= B.info.slot0 + B.info.count) return;
    const float4 p = pos[i];
    float* o = dst + (size_t)(B.info.real0 + (i - B.info.slot0)) * stride;
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
}

// ====================================================== host launchers ====
namespace lps_dev {

static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }
static inline uint32_t xcd_grid(uint32_t nlogical) { return cdiv(nlogical, LPS_XCDS) * LPS_XCDS; }

void launch_bbox(hipStream_t s, const float4* pos, const float4* vel, float mass, uint32_t n, const Scenes& sc,
                 uint32_t* part) {
    const uint32_t nl = cdiv(n, LPS_BLOCK);
    k_bbox<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(pos, vel, mass, sc, part, nl);
}

void launch_scene_energy(hipStream_t s, const Scenes& sc, const uint32_t* part, double* out) {
    k_scene_energy<<<sc.count, LPS_BLOCK, 0, s>>>(sc, part, out);
}

uint32_t bbox_parts(uint32_t n) { return cdiv(n, LPS_BLOCK); }

void launch_grid_setup(hipStream_t s, DevGrid* g, const Scenes& sc, const uint32_t* part, float h, uint32_t cell_cap,
                       double* energy) {
    if (sc.count == 1u) {
        k_grid_one<<<1, LPS_BLOCK, 0, s>>>(g, sc, part, h, cell_cap, energy);
    } else {
        k_scene_grid<<<sc.count, LPS_BLOCK, 0, s>>>(sc, part, h, energy);
        k_grid_finish<<<1, LPS_BLOCK, 0, s>>>(g, sc, h, cell_cap);
    }
}

uint32_t radix_tiles(uint32_t n) { return cdiv(n, LPS_RADIX_TILE); }

void launch_keys(hipStream_t s, const float4* pos, uint32_t n, const DevGrid* g, const Scenes& sc, uint32_t* keys,
                 uint32_t* vals, uint32_t* tile_max, uint32_t* hist) {
    const uint32_t nt = radix_tiles(n);
    k_keys<<<nt, LPS_BLOCK, 0, s>>>(pos, n, g, sc, keys, vals, tile_max, hist, nt);
}

void launch_radix_pass(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       uint32_t n, uint32_t shift, DevGrid* g, uint32_t* hist, uint32_t* digit_total) {
    const uint32_t nt = radix_tiles(n);
    // the first digit's histogram comes from the key kernel
    if (shift != 0u) k_radix_hist<<<nt, LPS_BLOCK, 0, s>>>(kin, n, shift, g, hist, nt);
    k_radix_scan<<<256, LPS_BLOCK, 0, s>>>(hist, nt, digit_total, g, shift);
    k_radix_scatter<<<nt, LPS_BLOCK, 0, s>>>(kin, vin, kout, vout, n, shift, g, hist, nt, digit_total);
}

void launch_reorder(hipStream_t s, const ReorderArgs& a) {
    const uint32_t nl = cdiv(a.n, LPS_BLOCK);
    k_reorder<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(a.keysA, a.valsA, a.keysB, a.valsB, a.n, a.g, a.pos_in, a.vel_in,
                                                 a.vh_in, a.first_in, a.id_in, a.pprev_in, a.pos_out, a.vel_out,
                                                 a.vh_out, a.first_out, a.id_out, a.pprev_out, a.run_end,
                                                 a.skeys, a.perm, a.sc, a.cq, a.tile_max, nl);
}

uint32_t cell_blocks(uint32_t cell_cap) { return cdiv(cell_cap + 1u, LPS_SCAN_TILE); }

void launch_cell_table(hipStream_t s, uint32_t* run_end, const DevGrid* g, uint32_t* block_max, uint32_t* cell_lb,
                       uint32_t cell_cap) {
    const uint32_t nb = cell_blocks(cell_cap);
    k_cell_scan<<<nb, LPS_BLOCK, 0, s>>>(run_end, g, block_max, cell_lb);
}

void launch_density(hipStream_t s, const PhysArgs& a, float* press, float* q) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    if (a.quad) {
        const uint32_t nq = nl * LPS_QUAD;
        k_density_q<<<xcd_grid(nq), LPS_BLOCK, 0, s>>>(a.pos, a.g, a.sc, a.cell_lb, a.c, a.tables, a.posrho, press, q,
                                                      a.mk, nq);
        return;
    }
    k_density<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(a.pos, a.g, a.sc, a.cell_lb, a.c, a.tables, a.posrho, press, q, a.mk,
                                                nl);
}

void launch_forces(hipStream_t s, const PhysArgs& a, float4* xpred, const float* pprev, float* feat) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    if (a.quad) {
        const uint32_t nq = nl * LPS_QUAD;
        auto kern = feat ? k_forces_q<true> : k_forces_q<false>;
        kern<<<xcd_grid(nq), LPS_BLOCK, 0, s>>>(a.posrho, a.vel, a.vh, a.first, a.g, a.cell_lb, a.c, a.tables,
                                                a.boxes, a.spheres, a.planes, a.force, xpred, pprev, feat, a.sc, a.mk,
                                                nq);
        return;
    }
    auto kern = feat ? k_forces<true> : k_forces<false>;
    kern<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(a.posrho, a.vel, a.vh, a.first, a.g, a.cell_lb, a.c, a.tables, a.boxes,
                                            a.spheres, a.planes, a.force, xpred, pprev, feat, a.sc, a.mk, nl);
}

void launch_pressure_sph(hipStream_t s, const PhysArgs& a, const float* q) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    if (a.quad) {
        const uint32_t nq = nl * LPS_QUAD;
        k_pressure_sph_q<<<xcd_grid(nq), LPS_BLOCK, 0, s>>>(a.posrho, q, a.g, a.cell_lb, a.c, a.tables, a.force, a.sc,
                                                           a.mk, nq);
        return;
    }
    k_pressure_sph<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(a.posrho, q, a.g, a.cell_lb, a.c, a.tables, a.force, a.sc, a.mk,
                                                     nl);
}

template <int WG>
static void pci_density_wg(hipStream_t s, const PhysArgs& a, const float4* xpred, float* rho_pred, float* press,
                           float4* xq, uint32_t k, float* feat) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    auto kern = (feat && k == 1u) ? k_pci_density<true, WG> : k_pci_density<false, WG>;
    kern<<<xcd_grid(nl * (LPS_BLOCK / WG)), WG, 0, s>>>(a.pos, xpred, a.g, a.cell_lb, a.c, a.tables, rho_pred, press,
                                                       xq, k, feat, a.posrho, a.sc, a.mk, nl);
}

void launch_pci_density(hipStream_t s, const PhysArgs& a, const float4* xpred, float* rho_pred, float* press,
                        float4* xq, uint32_t k, float* feat, uint32_t threads) {
    if (a.quad) {
        const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
        auto kern = (feat && k == 1u) ? k_pci_density_q<true> : k_pci_density_q<false>;
        kern<<<xcd_grid(nl * LPS_QUAD), LPS_BLOCK, 0, s>>>(a.pos, xpred, a.g, a.cell_lb, a.c, a.tables, rho_pred, press,
                                                          xq, k, feat, a.posrho, a.sc, a.mk, nl);
        return;
    }
    if (threads == 64u) pci_density_wg<64>(s, a, xpred, rho_pred, press, xq, k, feat);
    else if (threads == 128u) pci_density_wg<128>(s, a, xpred, rho_pred, press, xq, k, feat);
    else pci_density_wg<LPS_BLOCK>(s, a, xpred, rho_pred, press, xq, k, feat);
}

template <int WG>
static void pci_pressure_wg(hipStream_t s, const PhysArgs& a, const float4* xq, const float* rho_pred, float4* fp,
                            float4* xpred_next, uint32_t k, const LoopRule& rule, uint32_t* host_flag) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    k_pci_pressure<WG><<<xcd_grid(nl * (LPS_BLOCK / WG)), WG, 0, s>>>(
        a.pos, xq, rho_pred, a.g, a.cell_lb, a.c, a.tables, fp, a.force, a.posrho, a.vel, a.vh, a.first, a.boxes,
        a.spheres, a.planes, xpred_next, k, a.sc, a.mk, rule, host_flag, nl);
}

void launch_pci_pressure(hipStream_t s, const PhysArgs& a, const float4* xq, const float* rho_pred, float4* fp,
                         float4* xpred_next, uint32_t k, const LoopRule& rule, uint32_t* host_flag,
                         uint32_t threads) {
    if (a.quad) {
        const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
        k_pci_pressure_q<<<xcd_grid(nl * LPS_QUAD), LPS_BLOCK, 0, s>>>(
            a.pos, xq, rho_pred, a.g, a.cell_lb, a.c, a.tables, fp, a.force, a.posrho, a.vel, a.vh, a.first, a.boxes,
            a.spheres, a.planes, xpred_next, k, a.sc, a.mk, rule, host_flag, nl);
        return;
    }
    if (threads == 64u) pci_pressure_wg<64>(s, a, xq, rho_pred, fp, xpred_next, k, rule, host_flag);
    else if (threads == 128u) pci_pressure_wg<128>(s, a, xq, rho_pred, fp, xpred_next, k, rule, host_flag);
    else pci_pressure_wg<LPS_BLOCK>(s, a, xq, rho_pred, fp, xpred_next, k, rule, host_flag);
}

void launch_integrate(hipStream_t s, const PhysArgs& a, const float4* fp, float* pprev, const float* press,
                      uint32_t* bbox_part) {
    const uint32_t nl = cdiv(a.c.n, LPS_BLOCK);
    k_integrate<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(a.pos, a.vel, a.vh, a.first, a.force, fp, a.posrho, pprev, press,
                                                   a.g, a.c, a.boxes, a.spheres, a.planes, bbox_part, a.sc,
                                                   a.mk.stride, nl);
}

void launch_export_positions(hipStream_t s, const float4* pos, uint32_t n, const Scenes& sc, float* dst,
                             uint32_t stride) {
    const uint32_t nl = cdiv(n, LPS_BLOCK);
    k_export_positions<<<xcd_grid(nl), LPS_BLOCK, 0, s>>>(pos, sc, dst, stride, nl);
}

}  // namespace lps_dev
