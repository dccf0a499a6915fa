/*
 * include/lps.h -- C ABI of the MI355X SPH/PCISPH step library (liblps.so).
 *
 * The library replaces the step path of the reference's LiquidPhysics solver:
 *   PhysicSolver::update(dt)          LiquidPhysics/src/PhysicSolver.cpp:12-21
 * whose phases branch on ParallelizationType {NONE, CPU, GPU}
 *   (LiquidPhysics/include/SPHSolver.h:19-23) and reach the OpenCL kernels in
 *   LiquidPhysics/cl_kernels/SPHKernels.cl through ParallelComputationInterface
 *   (LiquidPhysics/include/Parallelization/ParallelComputationInterface.h:35-58).
 * A solver adapter (a fourth ParallelizationType) calls these entry points in
 * place of the OpenCL launch sequence (include/lps_liphen_adapter.hpp is one).
 *
 * Conventions: plain C types, host arrays owned by the caller and copied in
 * or out; device memory owned by the context.  Every function returns an
 * lps_status.  One HIP stream per context; a context is thread-compatible but
 * not thread-safe (the reference solver is single-threaded too,
 * LiquidPhysics/src/Parallelization/OpenCLInterface.cpp:121,131).
 * Particle state is kept in the reference's sorted order between steps: each
 * step sorts by grid cell, stable with respect to the incoming order, exactly
 * as the OpenCL path does (SPHSolver.cpp:1049-1119, readback order 412-422).
 */
#ifndef LPS_H
#define LPS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPS_ABI_VERSION 1u
#define LPS_TABLE_SIZE 1001 /* SPHKernel subdivision 1000 + 1 (Kernels/SPHKernel.cpp:4-7,124) */

typedef enum lps_status {
    LPS_OK = 0,
    LPS_ERR_INVALID = 1,       /* bad argument, size or ABI version */
    LPS_ERR_NO_DEVICE = 2,     /* no usable gfx950 device (caller falls back) */
    LPS_ERR_OOM = 3,           /* device allocation failed */
    LPS_ERR_NONFINITE = 4,     /* NaN/inf in the particle state or a density error */
    LPS_ERR_GRID_OVERFLOW = 5, /* bounding-box cell count beyond the supported maximum */
    LPS_ERR_HIP = 6,           /* other HIP runtime error */
    LPS_ERR_STATE = 7          /* call out of order (e.g. step before particles) */
} lps_status;

typedef enum lps_method { LPS_METHOD_SPH = 0, LPS_METHOD_PCISPH = 1 } lps_method;

typedef enum lps_warm_start {
    LPS_WARM_NONE = 0,    /* pciInit: p = 0, fp = 0 (SPHKernels.cl:661-672) */
    LPS_WARM_LEARNED = 1, /* MLP-seeded pressure, falls back to NONE when it does not help */
    LPS_WARM_LEARNED_ALWAYS = 2 /* MLP-seeded pressure on every step, no fallback (measurement) */
} lps_warm_start;

/* Fluid and solver parameters.  Field meanings follow the reference setters
 * (SPHSolver.h:57-69, PCISPHSolver.h:19-20); derived quantities (kernel radius,
 * lookup tables, particle mass, PCISPH delta) follow SPHSolver.cpp:480-523 and
 * PCISPHSolver.cpp:286-328.  Every float field must be finite, and the radius,
 * kernel radius factor and rest density positive (else LPS_ERR_INVALID). */
typedef struct lps_params {
    uint32_t abi_version;            /* LPS_ABI_VERSION */
    int32_t method;                  /* lps_method */
    float gravity[3];                /* setGravity */
    float particle_radius;           /* setParticleRadius */
    float kernel_radius_factor;      /* setKernelRadiusFactor: h = factor * r_p */
    float rest_density;              /* setRestDensity */
    float stiffness;                 /* setPressureStiffnessCoefficient (SPH) */
    float negative_pressure_factor;  /* setNegativePressureFactor */
    float viscosity;                 /* setViscosityCoefficient */
    float surface_tension;           /* setSurfaceTensionCoefficient */
    float surface_tension_threshold; /* setSurfaceTensionThreshold */
    float restitution;               /* setRestitutionCoefficient */
    float friction;                  /* setFrictionCoefficient */
    float mass_override;             /* <= 0: the reference lattice rule */
    int32_t min_iterations;          /* PCISPH corrections always run (>= 1) */
    int32_t max_iterations;          /* PCISPH cap */
    float max_density_error_ratio;   /* setMaxDensityErrorRatio: stop when max|e|/rho0 < this */
    int32_t warm_start;              /* lps_warm_start */
    int32_t stall_iterations;        /* also stop once the error has not improved on its best value
                                        (from the second correction on) for this many corrections
                                        (0: off); such a step applies the pressure force whose
                                        predictions had that best error, reports that error and is
                                        reported as not converged */
} lps_params;

/* Static colliders (StaticCollisionBox / StaticCollisionSphere, include/Collision/),
 * plus a plane that the reference lacks. */
typedef struct lps_box { float center[3]; float half[3]; int32_t boundary; } lps_box;
typedef struct lps_sphere { float center[3]; float radius; int32_t boundary; } lps_sphere;
typedef struct lps_plane { float normal[3]; float offset; } lps_plane; /* fluid side: n.x >= offset */

typedef struct lps_step_stats {
    int32_t iterations;              /* PCISPH corrections run this step (0 for SPH) */
    float max_density_error_ratio;   /* max_i |e_i| / rho0 of the applied pressure force's
                                        predictions (the last correction's; the best one's after a
                                        stall stop) */
    int32_t warm_start_used;         /* 1 if the learned seed was applied */
    int32_t status;                  /* lps_status of this step */
    int32_t converged;               /* 1: tolerance met; 0: stopped at the cap or by the stall rule */
} lps_step_stats;

typedef struct lps_grid_info {
    float bbox_min[3];
    float bbox_max[3];
    float grid_offset[3];
    float grid_spacing;
    uint32_t grid_size[3];
    uint32_t cell_count;
} lps_grid_info;

typedef struct lps_ctx lps_ctx;

/* Context lifetime.  Replaces OpenCLInterface::initialize/reinitContext
 * (OpenCLInterface.cpp:65-135) and the buffer set-up of
 * SPHSolver::reinitParallelContext (SPHSolver.cpp:713-786). */
int lps_create(const lps_params* params, lps_ctx** out);
int lps_destroy(lps_ctx* ctx);
int lps_device_info(char* name, size_t cap, int32_t* cu_count);

/* Parameters and derived quantities (SPHSolver setters, SPHSolver.cpp:628-710). */
int lps_set_params(lps_ctx* ctx, const lps_params* params);
int lps_get_derived(lps_ctx* ctx, float* kernel_radius, float* division_step, float* mass,
                    float* tables /* 5 x LPS_TABLE_SIZE: W, W', W'' (poly6), spiky', visc'' */);
int lps_get_delta(lps_ctx* ctx, float dt, float* delta);
/* Lookup-index thresholds (no context, no device): thr[k], k = 1..1000, is
 * the smallest float s >= 0 with min(trunc(sqrtf(s) / (h/1000)), 1000) >= k,
 * thr[0] = 0 and thr[1001] = +inf, so the table index of a squared distance s
 * is the number of k >= 1 with thr[k] <= s.  The stencils use this in place
 * of the square root and division of SPHKernels.cl:237. */
int lps_index_thresholds(const lps_params* params, float* thr /* [LPS_TABLE_SIZE + 1] */);

/* Scene.  lps_set_particles replaces the upload of SPHSolver::initParallelBuffers
 * (SPHSolver.cpp:788-886); arrays are xyz-interleaved [3n]; vhalf and first may
 * be NULL (zero half velocity, first step set).  n >= 1, and n rounded up to
 * whole particle blocks of 256 slots is at most LPS_MAX_PARTICLES (so
 * n <= 2^28 - 256); other counts are refused with LPS_ERR_INVALID before any
 * array is read.  The same bound holds for the slots of all scenes together
 * (lps_set_scenes) and after lps_add_particles. */
#define LPS_MAX_PARTICLES ((1u << 28) - 1u)
int lps_set_particles(lps_ctx* ctx, const float* x, const float* v, const float* vhalf,
                      const uint8_t* first, uint32_t n);
/* Particles added mid-run (SPHSolver::addParticle between steps, as the
 * viewer's emitting scenarios do, e.g. WaterfallScenario.cpp:55-74): appended
 * after the current particles, which keep their state and sorted order; the
 * reference re-uploads all particles instead (m_hasParticleDataChanged,
 * SPHSolver.cpp:791).  Device storage grows by half again when full.  One
 * scene only (LPS_ERR_STATE for ensembles or an empty context); vhalf and
 * first as in lps_set_particles; new particles get ids counting on from the
 * last.  Moving colliders need only lps_set_colliders between steps. */
int lps_add_particles(lps_ctx* ctx, const float* x, const float* v, const float* vhalf, const uint8_t* first,
                      uint32_t n);
/* Readback of SPHSolver::onEndUpdate (SPHSolver.cpp:398-428), in sorted order.
 * Any pointer may be NULL; id returns each particle's index at lps_set_particles
 * (particles added later count on from there). */
int lps_get_particles(lps_ctx* ctx, float* x, float* v, float* vhalf, uint8_t* first, uint32_t* id,
                      uint32_t n);
int lps_get_count(lps_ctx* ctx, uint32_t* n);
/* Render interop (no counterpart in the reference, whose viewer reads every
 * particle back to the host each frame, SPHSolver.cpp:398-428): writes the
 * positions in lps_get_particles order into device memory, particle i's x, y,
 * z at dst[i * stride + 0..2] (stride >= 3 floats; 4 for a vec4 instance
 * buffer).  dst is device memory of capacity_bytes, e.g. a vertex buffer
 * mapped with hipGraphicsGLRegisterBuffer / hipGraphicsResourceGetMappedPointer;
 * a host pointer or a buffer too small for n particles is rejected
 * (LPS_ERR_INVALID).  n must equal the particle count.  Returns when the copy
 * is complete. */
int lps_copy_positions_device(lps_ctx* ctx, float* dst, uint32_t stride, uint32_t n, uint64_t capacity_bytes);
/* Collider upload of SPHSolver::initParallelBuffers (SPHSolver.cpp:932-1003). */
int lps_set_colliders(lps_ctx* ctx, const lps_box* boxes, uint32_t nbox, const lps_sphere* spheres,
                      uint32_t nsphere, const lps_plane* planes, uint32_t nplane);

/* Ensembles (no counterpart in the reference): one context can hold up to
 * LPS_MAX_SCENES independent scenes sharing the fluid parameters, each with
 * its own gravity and colliders, stepped together by lps_step (one launch per
 * phase for all scenes, one PCISPH loop rule per scene; a scene whose loop
 * has stopped skips the remaining corrections).  Every scene's results equal
 * those of the scene run alone.  Arrays are the scenes' particles
 * concatenated in scene order; gravity is [3 * nscenes] or NULL (the
 * parameters' gravity for all).  lps_set_particles is the single-scene case;
 * lps_set_colliders gives every scene the same colliders. */
#define LPS_MAX_SCENES 256
int lps_set_scenes(lps_ctx* ctx, uint32_t nscenes, const uint32_t* counts, const float* gravity,
                   const float* x, const float* v, const float* vhalf, const uint8_t* first);
int lps_set_scene_colliders(lps_ctx* ctx, uint32_t scene, const lps_box* boxes, uint32_t nbox,
                            const lps_sphere* spheres, uint32_t nsphere, const lps_plane* planes, uint32_t nplane);
int lps_get_scene_count(lps_ctx* ctx, uint32_t* nscenes);
/* Per-scene statistics of the last step (iterations, final error ratio, converged). */
int lps_get_scene_stats(lps_ctx* ctx, lps_step_stats* stats /* [nscenes] */, uint32_t nscenes);

/* Launch tuning (no counterpart in the reference, which fixes its OpenCL
 * workgroup size, SPHSolver.cpp:25).  use_graphs: replay the fixed part of a
 * PCISPH step (grid build, density, forces, corrections 1..min) as a captured
 * HIP graph; lookahead: corrections kept queued beyond the one whose
 * convergence result the host waits for; correction_threads: workgroup size
 * of the two correction kernels (256 runs a 256-slot particle block per
 * workgroup, 128 and 64 split it over 2 or 4); lanes_per_particle: lanes of
 * a wave that share one particle's neighbour walk in the stencil kernels (1,
 * or 4: a quad splits the candidates and adds the terms in the one-lane
 * order; 0 chooses 4 below 512 particle blocks of 256 slots, else 1; the
 * correction workgroup size then applies only with 1; with several scenes, 0
 * also gives quads to the corrections that run fewer than 512 blocks).
 * lps_autotune times `steps` steps of every combination from the current
 * state (4 lanes only below 2048 particle blocks, 0 only with several scenes
 * and at least 512 blocks), restores the state (the run continues exactly as
 * without tuning) and keeps the fastest; every setting gives bit-identical
 * results. */
typedef struct lps_tuning {
    int32_t use_graphs;            /* 0 or 1 */
    uint32_t lookahead;            /* 1..4 */
    float ms_per_step;             /* lps_autotune: measured time of the chosen setting */
    uint32_t correction_threads;   /* 64, 128 or 256 */
    uint32_t lanes_per_particle;   /* 0 (automatic), 1 or 4; lps_get_tuning reports the step's static choice */
} lps_tuning;
int lps_set_tuning(lps_ctx* ctx, const lps_tuning* tuning);
int lps_get_tuning(lps_ctx* ctx, lps_tuning* tuning);
int lps_autotune(lps_ctx* ctx, float dt, uint32_t steps, lps_tuning* chosen);

/* One or more steps of PhysicSolver::update (PhysicSolver.cpp:12-21) on the device.
 * stats may be NULL or hold nsteps entries.  dt must be positive and finite
 * (else LPS_ERR_INVALID).  Returns the first failing status. */
int lps_step(lps_ctx* ctx, float dt, uint32_t nsteps, lps_step_stats* stats);

/* Parity hooks: the grid and fields of the last step (sorted order).
 * keys/perm [n]; cell_start/cell_end [cap] follow buildCellList's rule
 * (SPHKernels.cl:148-186, end rule :230).  Any pointer may be NULL.  With
 * several scenes, info describes scene 0 and keys/cells are the global
 * ones (scene s's cells follow scene s-1's and its one empty-slot cell). */
int lps_get_grid(lps_ctx* ctx, lps_grid_info* info, uint32_t* keys, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end, uint32_t cap);
int lps_get_fields(lps_ctx* ctx, float* rho, float* pressure, float* force /* [3n] */,
                   float* rho_pred);

/* Learned warm start (PCISPH only; no counterpart in the reference, whose
 * pciInit starts every step from p = 0, SPHKernels.cl:661-672).
 * A per-particle MLP 16 -> 64 -> 64 -> 1 (ReLU), BF16 weights/activations on
 * the matrix cores with FP32 accumulation, predicts each particle's converged
 * pressure from features of the step's first correction; the prediction
 * replaces the first correction's pressure and the unchanged loop runs on to
 * the same tolerance.  Features (per particle, sorted order):
 *   0 rho/rho0 - 1   1 neighbours/32   2 velocity divergence*m*dt
 *   3-5 colour-field normal*h   6 its length*h   7 previous step's pressure/P
 *   8 |v| dt/h   9 first-step flag   10 |first prediction - x|/h
 *   11 e1 = rho*_1/rho0 - 1   12 max(e1, 0)   13 first pressure update/P
 *   14 (rho*_1 - rho)/rho0   15 100 e1^2
 * with P = delta * rho0 (PCISPH's delta of the step), so features and target
 * do not depend on particle size or time step.
 * Parameters (fp32 master copy): W1[64][16], b1[64], W2[64][64], b2[64],
 * W3[64], b3.  Inputs are normalised as (f - mean) * inv_std; the seed is the
 * output times target_scale times P (0 when not finite).  With negative
 * pressures disabled the loop never lowers a pressure, so the first
 * correction's pressure becomes max(p1, min(seed, LPS_WARM_KAPPA * p_prev))
 * (p1: the zero start's own first-correction pressure; p_prev: the
 * particle's applied pressure of the previous step): an over-prediction
 * shrinks by LPS_WARM_KAPPA per step instead of feeding itself through
 * feature 7.  With negative pressures enabled it becomes the seed.
 * With warm_start = LPS_WARM_LEARNED and weights loaded the seed is judged on
 * paired samples: wherever two consecutive steps ran in different modes, the
 * seeded step's cost minus the zero start's, a step's cost being its
 * corrections if it met the tolerance and max_iterations if it did not.
 * Every 16th step runs the mode not in use (giving two samples), and the
 * seed is used while the sum of the last 8 samples is not positive, i.e.
 * while its iteration count to the tolerance is not worse.  A seeded step
 * that turns non-finite is re-run unseeded, and it, like a seeded step that
 * adds mechanical energy (LPS_WARM_ENERGY_TOL), enters as a sample of
 * max_iterations. */
/* Seed envelope factor: with negative pressures disabled the first
 * correction's pressure is max(p1, min(seed, LPS_WARM_KAPPA * p_prev)), p1
 * the zero start's own first-correction pressure and p_prev the particle's
 * applied pressure of the previous step. */
#define LPS_WARM_KAPPA 0.85f
/* Energy tolerance of the fallback: a seeded step whose relative change of
 * kinetic + potential energy (per scene, relative to |KE| + |PE|) exceeds
 * the zero start's running mean (at least 0) by this much is a failed seed. */
#define LPS_WARM_ENERGY_TOL 1e-4
#define LPS_MLP_IN 16
#define LPS_MLP_HIDDEN 64
#define LPS_MLP_PARAMS (LPS_MLP_IN * LPS_MLP_HIDDEN + LPS_MLP_HIDDEN + LPS_MLP_HIDDEN * LPS_MLP_HIDDEN + LPS_MLP_HIDDEN + LPS_MLP_HIDDEN + 1)
int lps_mlp_set_weights(lps_ctx* ctx, const float* params /* LPS_MLP_PARAMS, fp32 master copy */,
                        const float* feature_scale /* [2*LPS_MLP_IN]: mean, inv_std */, float target_scale);
int lps_mlp_get_weights(lps_ctx* ctx, float* params, float* feature_scale, float* target_scale);
/* Compute the features on every PCISPH step (for lps_dump_pairs) when on. */
int lps_mlp_capture(lps_ctx* ctx, int32_t on);
/* Features, applied pressure and (if the step was seeded) the network's seed
 * (before the envelope) of the last PCISPH step, sorted order, pressures in
 * Pa.  Any pointer may be NULL. */
int lps_dump_pairs(lps_ctx* ctx, float* features /* [n*16] */, float* target /* [n] */,
                   float* seed /* [n] */, uint32_t n);
/* Device-resident training set: while capacity > 0 and stride > 0, every
 * PCISPH step that ends without error appends (features, final pressure / P)
 * of every stride-th particle (offset rotating per step) until capacity rows
 * are held.  A call with the current capacity keeps the rows and only sets
 * the stride (0 pauses); another capacity starts an empty set; 0 frees it. */
int lps_mlp_collect(lps_ctx* ctx, uint64_t capacity, uint32_t stride);
int lps_mlp_collected(lps_ctx* ctx, uint64_t* rows);
typedef struct lps_mlp_train_cfg {
    uint32_t epochs;
    uint32_t batch;        /* rows per Adam step */
    float lr, lr_final;    /* cosine decay from lr to lr_final over all steps */
    float beta1, beta2, eps;
    int32_t reinit;        /* 1: fresh He-normal weights; 0: continue from the loaded ones */
    int32_t refit_scale;   /* 1: feature normalisation and target scale from the rows */
    float over_weight;     /* loss weight of over-predictions (1: plain MSE) */
    uint64_t seed;         /* initialisation and per-epoch row permutation */
} lps_mlp_train_cfg;
/* Train on the collected rows on the device (loss mean(w d^2) with
 * d = y - target/target_scale, w = over_weight where d > 0, else 1; Adam);
 * epoch_loss [epochs] may be NULL.  The trained weights become the loaded
 * ones.  With negative pressures disabled the correction loop never lowers a
 * pressure, so an over-predicted seed is never corrected back; weighting
 * over-predictions more keeps the seed below the converged pressure. */
int lps_mlp_train(lps_ctx* ctx, const lps_mlp_train_cfg* cfg, float* epoch_loss);
/* Batch inference and gradient with the loaded weights (training checks):
 * y [n] is the raw network output (before target_scale and P); grad
 * [LPS_MLP_PARAMS] and loss are those of the training loss above. */
int lps_mlp_forward(lps_ctx* ctx, const float* x /* [n*16] */, uint32_t n, float* y);
int lps_mlp_grad(lps_ctx* ctx, const float* x /* [n*16] */, const float* t /* [n] */, uint32_t n,
                 float over_weight, float* grad, float* loss);

/* lps_mlp_grad without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight training or serving loops, harnesses that forbid hipMalloc).
 * Same inputs, outputs and result as lps_mlp_grad, for n <= LPS_NOALLOC_MAX_ROWS
 * rows (larger n: LPS_ERR_INVALID).  The staged inputs and the gradient
 * accumulators live in one static device array per GPU (about 4.3 MB, held
 * until the process ends) shared by every call on that GPU: calls that
 * overlap on one GPU (two threads, two contexts) race.  Needs weights loaded
 * (LPS_ERR_STATE otherwise); the context's stream is synchronised on return. */
#define LPS_NOALLOC_MAX_ROWS 65536
int lps_mlp_grad_noalloc(lps_ctx* ctx, const float* x /* [n*16] */, const float* t /* [n] */, uint32_t n,
                         float over_weight, float* grad, float* loss);

/* Kinetic (sum m|v|^2/2) and potential (sum -m g.x) energy of every scene's
 * current state, reduced on the device in double from per-block sums; no
 * counterpart in the reference (long-run statistics, warm-start fallback). */
int lps_get_energy(lps_ctx* ctx, double* kinetic /* [nscenes] */, double* potential /* [nscenes] */,
                   uint32_t nscenes);

/* Kernel timing (HIP events on the context stream), for measurement only. */
int lps_profile_enable(lps_ctx* ctx, int32_t on);
int lps_profile_query(lps_ctx* ctx, const char* kernel, uint32_t* launches, double* total_ms);
int lps_profile_reset(lps_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that record their own events. */
int lps_get_stream(lps_ctx* ctx, void** stream);
const char* lps_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* LPS_H */
