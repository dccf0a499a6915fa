// lps_common.h -- types shared by the host orchestration and the gfx950 kernels.
//
// Floating-point rule for every kernel: no FMA contraction (the library is
// compiled with -ffp-contract=off), IEEE division and square root (HIP's
// default correctly rounded fp32 div/sqrt), and the evaluation orders of the
// reference's OpenCL kernels (LiquidPhysics/cl_kernels/SPHKernels.cl), so the
// device results equal an IEEE single-precision evaluation of those orders.
#pragma once
#include <stdint.h>

#include "lps.h"

#define LPS_BLOCK 256           // threads per block for particle kernels (4 wave64)
#define LPS_RADIX_ITEMS 4       // keys per thread per radix tile
#define LPS_RADIX_TILE (LPS_BLOCK * LPS_RADIX_ITEMS)
#define LPS_SCAN_ITEMS 4       // cells per thread in the cell-table scan
#define LPS_SCAN_TILE (LPS_BLOCK * LPS_SCAN_ITEMS)
#define LPS_MAX_ITER_SLOTS 64   // per-correction error slots, used as a ring: correction k owns slot (k - 1) % 64
// host flag words per correction slot: state, corrections, error bits,
// converged, particle blocks the next correction runs (then unused)
#define LPS_FLAG_WORDS 8
#define LPS_MAX_COLLIDERS 1024    // per context (all scenes)
#define LPS_FEAT 16              // warm-start features per particle (LPS_MLP_IN)

// A context holds one or more independent scenes (an ensemble) that share
// the fluid parameters.  Scene s owns the slot range [slot0, slot0 + padded)
// of every particle array, padded to whole blocks of LPS_BLOCK slots, so each
// particle block belongs to exactly one scene; its real particles always sit
// in the first `count` slots (empty slots sort last, see below).
#define LPS_MAX_SCENES 256

// Per-scene constants, written by the host.
struct SceneInfo {
    uint32_t slot0;        // first slot (a multiple of LPS_BLOCK)
    uint32_t count;        // real particles
    uint32_t padded;       // slots reserved (count rounded up to LPS_BLOCK)
    uint32_t block0;       // slot0 / LPS_BLOCK
    float gravity[3];
    uint32_t box0, nbox, sphere0, nsphere, plane0, nplane;  // ranges of the collider arrays
    uint32_t real0;        // real particles of the scenes before this one (readback position)
};

// Per-scene grid and PCISPH loop state of the current step, written on the
// device (grid from the scene's bounding box, SPHSolver.cpp:1011-1040).  The
// scene's cells are global cells [cell_base, cell_base + cells); global cell
// cell_base + cells holds its empty slots.
struct SceneGrid {
    float bbox_min[3];
    float bbox_max[3];
    float offset[3];
    uint32_t gsize[3];
    uint32_t cells;
    uint32_t cell_base;
    uint32_t bad;            // 1: non-finite position, 2: extent too large for a grid
    uint32_t stop_iter;      // correction after which this scene's loop stopped (0: running)
    uint32_t best_bits;      // smallest max |e| of this step's corrections from the second on
    uint32_t best_k;         // the correction that had it (0: none yet)
    uint32_t best_buf;       // pressure buffer of correction best_k - 1 (its predictions came from it)
    uint32_t stall;          // corrections since the best value last improved
    uint32_t converged;      // the tolerance was met
    uint32_t final_bits;     // max |e_i| (float bits) the loop reports: the last correction's, or
                             // the best one's after a stall exit
    uint32_t final_buf;      // pressure buffer the step applies (read by the integration)
    // Pressure, predicted density and pressure force of a correction live in
    // one of three buffers: correction k writes buffer buf_of[k % 64], which
    // the loop rule of correction k - 1 chose as neither that correction's nor
    // the best one's, so the best correction's inputs survive until the stop.
    uint8_t buf_of[LPS_MAX_ITER_SLOTS];
    // max |e_i| (float bits) of correction k in slot (k - 1) % 64; the loop
    // rule of correction k clears the slot of correction k + 1
    uint32_t err_bits[LPS_MAX_ITER_SLOTS];
};

// Global state of the current step.
struct DevGrid {
    float spacing;           // cell size h (SPHSolver.cpp:1031)
    uint32_t nscenes;
    uint32_t total_cells;    // sum over scenes of cells + 1 (the empty-slot cell)
    uint32_t key_bits;       // bits of (total_cells - 1): radix passes beyond are skipped
    uint32_t passes_done;    // radix passes executed this step
    uint32_t status;         // lps_status of the step; kernels no-op when non-zero
    uint32_t cells_needed;   // total cells requested when status == GRID_OVERFLOW
    uint32_t steps_ok;       // steps whose grid set-up succeeded, minus failed loops (monotone)
    uint32_t running;        // scenes whose correction loop has not stopped
    uint32_t iterations;     // corrections run this step (max over scenes)
    uint32_t final_err;      // max over scenes of the last correction's max |e| (float bits)
    uint32_t converged;      // every scene met the tolerance
    uint32_t tiny_pred;      // a prediction of this step has a coordinate in (0, 2^-56)
    // Correction kernels run only the particle blocks of scenes whose loop
    // is still running.  Two block lists alternate; list_of[k % 64] names the
    // list of correction k (both of its passes).  Entries are written once per
    // correction, ahead of it, so a workgroup that starts late never sees a
    // list change under it.
    uint32_t act_count[2];   // blocks in list 0 / 1
    uint32_t list_of[LPS_MAX_ITER_SLOTS];
};

// Per-block partial record: orderable-uint min xyz, max xyz, non-finite flag,
// then the block's kinetic and potential energy (float bits), one spare word.
#define LPS_BBOX_PART 10

// Per-step constants passed by value (the role of the 128-byte
// ParallelSPHParameters block, ParallelSPHStructs.h:17-40).
struct SimConst {
    float gravity[3];
    float particle_radius;
    float h;            // kernel radius
    float div;          // table step h / 1000
    float s_max;        // largest float s with sqrtf(s) < h: d < h  <=>  d^2 <= s_max
    float s_relax;      // s_max widened by 2^-18: bound for a candidate filter evaluated with FMA
    float inv_div;      // 1 / div, for the first guess of a table index
    float mass;
    float rest_density;
    float stiffness;
    float npf;          // negative pressure factor
    float viscosity;
    float surface_tension;
    float st_threshold;
    float restitution;
    float friction;
    float dt;
    float delta;        // PCISPH delta (PCISPHSolver.cpp:286-328)
    uint32_t n;
    uint32_t nbox, nsphere, nplane;
    uint32_t cell_cap;  // allocated cell-table entries (excluding the sentinel)
};

// orderable encoding of floats for atomic min/max (monotone for non-NaN)
static inline __host__ __device__ uint32_t f2o(float f) {
    union { float f; uint32_t u; } c;
    c.f = f;
    return (c.u & 0x80000000u) ? ~c.u : (c.u | 0x80000000u);
}
static inline __host__ __device__ float o2f(uint32_t u) {
    union { float f; uint32_t u; } c;
    c.u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return c.f;
}

// Accept masks recorded per particle by a neighbour pass: one 32-bit mask
// per row of the 27-cell stencil (9 rows), covering a row's first 32
// candidates; a replaying pass filters any further candidates live.
#define LPS_MC_ROWS 9

// Position arrays scanned by the stencils carry this many padding entries:
// the candidate filter reads whole groups of eight past a row's end.
#define LPS_STENCIL_PAD 8

// Device table block: the five lookup tables, then the index thresholds.
#define LPS_THR_SIZE (LPS_TABLE_SIZE + 1)
#define LPS_TABLE_FLOATS (5 * LPS_TABLE_SIZE + LPS_THR_SIZE)

// Host-side derived quantities (lps_host.cpp).
struct Derived {
    float h, div, mass, s_max, s_relax, inv_div;
    float tables[5][LPS_TABLE_SIZE];  // W, W', W'' (poly6), spiky', viscosity''
    // thr[k] (k = 1..1000): smallest float s >= 0 whose lookup index
    // min(trunc(sqrtf(s) / div), 1000) is >= k; thr[0] = 0, thr[1001] = +inf.
    // The index of a squared distance s is then the number of k >= 1 with
    // thr[k] <= s, so stencils need neither the square root nor the division
    // to find it.
    float thr[LPS_THR_SIZE];
};
void lps_index_thresholds(float h, float div, float* thr);
void lps_compute_derived(const lps_params* p, Derived* d);
float lps_compute_delta(const lps_params* p, const Derived* d, float dt);
