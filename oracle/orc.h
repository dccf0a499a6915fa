/*
 * oracle/orc.h -- C interface of the CPU oracle for the SPH/PCISPH step.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the
 * reference solver (GitDaroth/LiquidSimulation, LiquidPhysics/) used as the
 * checker for the HIP library and as the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Two step flavours are restated:
 *   - "ocl": LiquidPhysics/cl_kernels/SPHKernels.cl plus the host grid/sort
 *     logic of SPHSolver::buildParallelGrid (LiquidPhysics/src/SPHSolver.cpp:1008-1132)
 *     and the OpenCL branch of PCISPHSolver::accumulatePressureForces
 *     (LiquidPhysics/src/PCISPHSolver.cpp:171-239).  This is the semantics the
 *     GPU path implements.
 *   - "seq": the ParallelizationType::NONE branches of SPHSolver/PCISPHSolver
 *     (SPHSolver.cpp:188-464, PCISPHSolver.cpp:55-170) with SPHSpatialGrid
 *     neighbour lists (SPHSpatialGrid.cpp:19-90).
 * Physics is templated on float/double; the grid (keys, permutation, cell
 * table) is always computed from FP32 positions.
 *
 * Parity pins (what checks each function):
 *   tables, mass, delta     -> closed-form known answers + oracle/_ref (reference
 *                              SPHKernel/DefaultKernel/PressureKernel/ViscosityKernel)
 *   keys                    -> closed-form integer formula + brute force in tests
 *   permutation             -> std::stable_sort and the reference-style LSD radix
 *   cell table              -> lower_bound on sorted keys
 *   density / forces        -> O(N^2) brute force, rest-lattice density, momentum
 *                              conservation of pair terms, seq flavour vs oracle/_ref
 *   integrate / collisions  -> free-fall closed form, per-case hand computations,
 *                              oracle/_ref (reference SPHParticle / StaticCollision*)
 *   PCISPH loop             -> iteration-count rule, monotone error on a compressed block
 *   PCISPH corrections      -> prediction, predicted density, pressure update and
 *                              pressure force against FP64 brute force over all pairs
 *                              with the 27-cell rule of the current cells (including
 *                              a state where that anchor decides pairs), momentum of
 *                              the pressure pair terms, OpenCL vs sequential form once
 *                              the rho*_i / rho_i prefactor (H2) is divided out
 *                              (tests/test_oracle_pcisph.py)
 *   stall rule and its exit -> energy of the 16k dam break with the best-correction
 *                              exit vs the last-correction exit (test hook), and the
 *                              524k dam break's error floor at the walls
 *                              (tests/test_oracle_stability.py)
 *   plane collider          -> no reference counterpart (new collider), pinned by
 *                              hand-computed cases only
 */
#ifndef LPS_ORACLE_H
#define LPS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TABLE_SIZE 1001

/* Same field order as lps_params in include/lps.h. */
typedef struct orc_params {
    uint32_t abi_version;
    int32_t method; /* 0 SPH, 1 PCISPH */
    float gravity[3];
    float particle_radius;
    float kernel_radius_factor;
    float rest_density;
    float stiffness;
    float negative_pressure_factor;
    float viscosity;
    float surface_tension;
    float surface_tension_threshold;
    float restitution;
    float friction;
    float mass_override;
    int32_t min_iterations;
    int32_t max_iterations;
    float max_density_error_ratio;
    int32_t warm_start;
    int32_t stall_iterations;
} orc_params;

typedef struct orc_box { float center[3]; float half[3]; int32_t boundary; } orc_box;
typedef struct orc_sphere { float center[3]; float radius; int32_t boundary; } orc_sphere;
typedef struct orc_plane { float normal[3]; float offset; } orc_plane;

/* Optional outputs of one step; every pointer may be NULL. */
typedef struct orc_trace {
    /* grid (sorted order) */
    float bbox_min[3];
    float bbox_max[3];
    float grid_offset[3];
    float grid_spacing;
    uint32_t grid_size[3];
    uint32_t cell_count;
    uint32_t* keys;        /* [n] sorted keys */
    uint32_t* perm;        /* [n] sorted slot -> incoming index */
    uint32_t* cell_start;  /* [cell_cap] reference cellList (lower_bound) */
    uint32_t cell_cap;
    /* fields in sorted order; float* for f32 steps, double* for f64 steps */
    void* rho;             /* [n] density after the density phase */
    void* pressure;        /* [n] SPH pressure, or final PCISPH pressure */
    void* force;           /* [3n] total force before integration */
    void* rho_pred;        /* [n] last PCISPH predicted density */
    /* PCISPH */
    int32_t iterations;
    float delta;
    float err_ratio[64];   /* max density error ratio per iteration (first 64) */
    int32_t converged;     /* PCISPH: tolerance met */
    int32_t status;
    float final_err_ratio; /* max density error ratio of the last correction */
    void* xpred;           /* [3n] PCISPH: predictions of the last correction */
    void* fp;              /* [3n] PCISPH: the pressure force the step applied */
} orc_trace;

/* derived quantities (SPHSolver::recalcKernelRadius / recalcParticleMass) */
int orc_derived_f32(const orc_params* p, float* h, float* div_step, float* mass, float* tables5x1001);
int orc_derived_f64(const orc_params* p, double* h, double* div_step, double* mass, double* tables5x1001);
/* PCISPHSolver::calcDelta in FP32 */
float orc_delta_f32(const orc_params* p, float dt);

/* grid only (bbox, keys, stable sort, cell table); x is [3n] */
int orc_grid(const orc_params* p, int32_t n, const float* x, orc_trace* tr, int32_t use_radix);

/* one "ocl" step, state updated in place and left in sorted order */
int orc_step_ocl_f32(const orc_params* p, const orc_box* boxes, int32_t nbox,
                     const orc_sphere* spheres, int32_t nsphere,
                     const orc_plane* planes, int32_t nplane,
                     float dt, int32_t n, float* x, float* v, float* vh, uint8_t* first,
                     orc_trace* tr, int32_t nthreads);
/* the same step with the first correction's pressure replaced by seed[n]
 * (sorted order of this step) unless the loop stops at the first correction:
 * the learned warm start's loop, with the network's output as given input */
int orc_step_ocl_seeded_f32(const orc_params* p, const orc_box* boxes, int32_t nbox,
                            const orc_sphere* spheres, int32_t nsphere,
                            const orc_plane* planes, int32_t nplane,
                            float dt, int32_t n, float* x, float* v, float* vh, uint8_t* first,
                            const float* seed, orc_trace* tr, int32_t nthreads);
/* the learned warm start's seed rule: with negative pressures disabled, the
 * first correction's pressure becomes max(p1, min(seed, kappa * pprev)),
 * p1 the zero start's first-correction pressure, seed[n] the network's
 * pressure in this step's sorted order (NULL: unbounded) and pprev[n] the
 * previous step's applied pressure in the incoming particle order */
int orc_step_ocl_warm_f32(const orc_params* p, const orc_box* boxes, int32_t nbox,
                          const orc_sphere* spheres, int32_t nsphere,
                          const orc_plane* planes, int32_t nplane,
                          float dt, int32_t n, float* x, float* v, float* vh, uint8_t* first,
                          const float* seed, const float* pprev, float kappa, orc_trace* tr, int32_t nthreads);
int orc_step_ocl_f64(const orc_params* p, const orc_box* boxes, int32_t nbox,
                     const orc_sphere* spheres, int32_t nsphere,
                     const orc_plane* planes, int32_t nplane,
                     float dt, int32_t n, double* x, double* v, double* vh, uint8_t* first,
                     orc_trace* tr, int32_t nthreads);
/* one "seq" (NONE) step, particle order unchanged; colliders in insertion order
 * given by kinds[] (0 box, 1 sphere, 2 plane) indexing into the arrays */
int orc_step_seq_f32(const orc_params* p, const orc_box* boxes, const orc_sphere* spheres,
                     const orc_plane* planes, const int32_t* kinds, const int32_t* index, int32_t ncoll,
                     float dt, int32_t n, float* x, float* v, float* vh, uint8_t* first,
                     orc_trace* tr);

/* helpers exposed for unit tests */
int orc_collide_ocl(const orc_params* p, const orc_box* boxes, int32_t nbox,
                    const orc_sphere* spheres, int32_t nsphere,
                    const orc_plane* planes, int32_t nplane,
                    int32_t n, float* x, float* vel);
int orc_neighbors_seq(float h, int32_t n, const float* x, int32_t i, int32_t* out, int32_t cap);
int orc_brute_density(const orc_params* p, int32_t n, const float* x, double* rho);

/* test hook: 0 makes the stall exit apply the last correction instead of the
 * best one (the stall rule's earlier form); default 1 */
void orc_set_best_exit(int32_t on);

/* multi-step drivers for the CPU baseline (ocl flavour, OpenMP) */
int orc_run_ocl_f32(const orc_params* p, const orc_box* boxes, int32_t nbox,
                    const orc_sphere* spheres, int32_t nsphere,
                    const orc_plane* planes, int32_t nplane,
                    float dt, int32_t nsteps, int32_t n, float* x, float* v, float* vh, uint8_t* first,
                    int32_t* iterations_out, float* err_out, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
