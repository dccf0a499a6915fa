// lps_api.hip -- C ABI (include/lps.h) and the host orchestration of one step.
//
// The reference's OpenCL branch drives a step from the host with blocking
// transfers in the middle (host bounding box SPHSolver.cpp:1011-1029, host
// radix scan :1066-1078, per-step buffer re-creation :1123-1125,
// PCISPHSolver.cpp:271-283, readback :398-428).  Here the whole step stays on
// the device: bounding box, grid parameters, sort, cell table and the PCISPH
// convergence test are kernels, and the host only enqueues.  The PCISPH loop
// keeps one correction queued ahead of the convergence result it waits for;
// corrections issued after convergence see the device "done" flag and return
// immediately.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "lps_common.h"
#include "lps_kernels.h"

// Scratch memory of lps_mlp_grad_noalloc: one array per GPU, held until the
// process ends.  Layout (bytes): staged features [LPS_NOALLOC_MAX_ROWS][16]
// floats, staged targets [LPS_NOALLOC_MAX_ROWS] floats, gradient
// accumulators [LPS_MLP_PARAMS] floats (padded to 256 bytes), loss [1].
#define LPS_NOALLOC_X_BYTES ((size_t)LPS_NOALLOC_MAX_ROWS * LPS_MLP_IN * sizeof(float))
#define LPS_NOALLOC_T_BYTES ((size_t)LPS_NOALLOC_MAX_ROWS * sizeof(float))
#define LPS_NOALLOC_G_BYTES ((((size_t)LPS_MLP_PARAMS * sizeof(float)) + 255u) / 256u * 256u)
#define LPS_NOALLOC_BYTES (LPS_NOALLOC_X_BYTES + LPS_NOALLOC_T_BYTES + LPS_NOALLOC_G_BYTES + 256u)
static __device__ __align__(256) unsigned char scratch[LPS_NOALLOC_BYTES];

namespace {

// Device address of `scratch` per device ordinal, resolved once per device
// with that device current (the host-side name is only a stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_status[kMaxDevices] = {};  // 0 unresolved, 1 resolved, 2 failed

// Called by lps_create with the context's device current.
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_status[device] != 0) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_scratch_status[device] = (e == hipSuccess && ptr) ? 1 : 2;
    if (e != hipSuccess) (void)hipGetLastError();
}

unsigned char* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_status[device] == 1 ? static_cast<unsigned char*>(g_scratch_ptr[device]) : nullptr;
}

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

// The learned seed's fallback (LPS_WARM_LEARNED).  The corrections a step
// needs vary several-fold from one step to the next and drift over a run, so
// the seed is judged on paired samples: wherever two consecutive steps ran in
// different modes, the seeded one's cost minus the zero start's.  A step's
// cost is its corrections if it met the tolerance and max_iterations if it
// did not (a loop the stall rule ended early has not reached the tolerance,
// whatever its count).  The seed is used while the sum of the last WS_WINDOW
// samples is not positive; every WS_PROBE_PERIOD-th step runs the other mode,
// which gives two samples (the probe against the step before it and the step
// after it).  A seeded step that fails (non-finite, or adding energy) enters
// as a sample of max_iterations.
constexpr uint64_t WS_PROBE_PERIOD = 16;
constexpr uint32_t WS_WINDOW = 8;

struct WarmFallback {
    double win[WS_WINDOW] = {};
    uint32_t n = 0, head = 0;
    int32_t prev_its = 0;
    int8_t prev_seeded = -1;     // mode of the last recorded step; -1 after new weights or a new state
    bool on = true;
    uint64_t steps = 0;          // steps recorded since the weights changed
    double ema_de_plain = 0.0;   // running relative energy change of zero-start steps

    void push(double d) {
        win[head] = d;
        head = (head + 1) % WS_WINDOW;
        if (n < WS_WINDOW) n++;
        double sum = 0.0;
        for (uint32_t i = 0; i < n; i++) sum += win[i];
        on = sum <= 0.0;
    }
    bool probe() const { return steps % WS_PROBE_PERIOD == WS_PROBE_PERIOD - 1u; }
    // new weights: no samples, seed on (the zero start's energy mean stays)
    void restart() {
        const double de = ema_de_plain;
        *this = WarmFallback{};
        ema_de_plain = de;
    }
    void record(bool seeded, int32_t its) {
        if (prev_seeded >= 0 && (prev_seeded != 0) != seeded)
            push(seeded ? (double)(its - prev_its) : (double)(prev_its - its));
        prev_seeded = seeded ? 1 : 0;
        prev_its = its;
        steps++;
    }
};

}  // namespace

struct lps_ctx {
    lps_params params{};
    Derived der{};
    hipStream_t stream = nullptr;
    int device = 0;

    uint32_t n = 0;         // particle slots: all scenes, each padded to whole blocks
    uint32_t n_real = 0;    // particles
    uint32_t next_id = 0;   // id of the next particle added (ids number particles in order of arrival)
    uint32_t cap_n = 0;     // allocated slot capacity
    uint32_t cell_cap = 0;  // allocated cell-table entries (plus one sentinel)
    uint32_t passes = 0;    // radix passes for cell_cap
    int cur = 0;            // which state buffer holds the sorted state
    uint64_t steps = 0;
    bool have_step = false; // a step has produced grid/fields

    // device state
    float4 *pos[2] = {}, *vel[2] = {}, *vh[2] = {};
    uint8_t* first[2] = {};
    uint32_t* id[2] = {};
    float* pprev[2] = {};
    uint32_t *keysA = nullptr, *valsA = nullptr, *keysB = nullptr, *valsB = nullptr;
    uint32_t *skeys = nullptr, *perm = nullptr;
    uint32_t* hist = nullptr;
    uint32_t* digit_total = nullptr;
    uint32_t* bbox_part = nullptr;
    float4 *posrho = nullptr, *force = nullptr, *fp = nullptr, *xpred[2] = {};
    float4* xq = nullptr;  // predicted-density pass: (prediction, p / rho*^2) [cap_n + LPS_STENCIL_PAD]
    float *press = nullptr, *q = nullptr, *rho_pred = nullptr;
    uint32_t *mk_mask = nullptr, *mk_ok = nullptr;  // recorded candidate masks [LPS_MC_ROWS][cap_n], flags
    uint2* mk_cq = nullptr;                         // compact predictions [cap_n + LPS_STENCIL_PAD]
    uint32_t *run_end = nullptr, *cell_lb = nullptr, *block_max = nullptr;
    float* tables = nullptr;
    lps_box* boxes = nullptr;
    lps_sphere* spheres = nullptr;
    lps_plane* planes = nullptr;
    DevGrid* grid = nullptr;

    // scenes (one for the single-scene API): layout, gravity, colliders
    std::vector<SceneInfo> h_sinfo{SceneInfo{}};
    std::vector<std::vector<lps_box>> s_boxes{{}};
    std::vector<std::vector<lps_sphere>> s_spheres{{}};
    std::vector<std::vector<lps_plane>> s_planes{{}};
    std::vector<float> s_gravity;          // [3 * nscenes] when set per scene, else empty
    SceneInfo* d_sinfo = nullptr;
    SceneGrid* d_sgrid = nullptr;
    SceneGrid* h_sgrid = nullptr;          // pinned copy for per-scene statistics
    uint8_t* d_bscene = nullptr;           // scene of every particle block
    uint32_t* d_actlist = nullptr;         // [2][blocks] active-block lists of the correction kernels

    // host
    uint32_t* flags = nullptr;  // pinned, coherent: LPS_FLAG_WORDS words per iteration slot
    DevGrid* h_grid = nullptr;  // pinned copy
    // coherent: per scene kinetic and potential energy of the state entering
    // the last step (written by its grid set-up) or of the current state
    // (lps_get_energy); the warm start's fallback compares consecutive steps
    double* h_energy = nullptr;
    std::vector<double> e_prev;  // per scene energy entering the previous step
    bool e_prev_valid = false;
    bool e_prev_seeded = false;
    int last_fp_valid = 0;      // PCISPH fields available

    // learned warm start: fp32 master weights, feature normalisation, target scale
    std::vector<float> mlp_w;
    std::vector<float> mlp_fscale;
    float mlp_tscale = 1.f;
    bool mlp_loaded = false;
    bool capture = false;         // compute features every PCISPH step (for lps_dump_pairs)
    float* d_mlp = nullptr;       // params [LPS_MLP_PARAMS] then fscale [2 * LPS_MLP_IN]
    uint8_t* d_mlp_packed = nullptr;  // BF16 inference copy of d_mlp (lps_dev::mlp_packed_bytes())
    float* feat = nullptr;        // [LPS_MLP_IN][cap_n] features of the current step (feature-major)
    float* seed = nullptr;        // [cap_n] seeded pressure of the current step
    bool feat_valid = false;      // the last step wrote features (and final pressure)
    bool seed_valid = false;      // the last step applied the seed
    WarmFallback wf;              // the seed's fallback rule (seed_this_step)
    // training-set collection and training
    float *coll_x = nullptr, *coll_t = nullptr;
    uint64_t coll_cap = 0, coll_rows = 0;
    uint32_t coll_stride = 1, coll_offset = 0;
    float *d_grad = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_loss = nullptr;
    double* d_stats = nullptr;

    // captured step prefixes (grid build, density, forces and corrections
    // 1..min of a PCISPH step) keyed by state-buffer parity, seeding and
    // feature capture; valid while `epoch` (bumped by every change of
    // buffers or parameters) and dt match
    struct PrefixGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t epoch = 0;
        float dt = 0.f;
    };
    PrefixGraph graphs[2][2][2];
    uint64_t epoch = 1;
    bool graphs_on = true;
    uint32_t lookahead = 1;   // corrections kept queued ahead of the one waited for
    uint32_t corr_threads = LPS_BLOCK;  // workgroup size of the correction kernels
    uint32_t lanes = 0;                 // lanes per particle in the stencil kernels: 0 automatic, 1 or 4

    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    std::vector<ProfRec> pending;
    std::map<std::string, std::pair<uint32_t, double>> prof_acc;
};

namespace {

int hip_status(hipError_t e) {
    if (e == hipSuccess) return LPS_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return LPS_ERR_OOM;
    return LPS_ERR_HIP;
}

#define HIPC(expr)                              \
    do {                                        \
        hipError_t e_ = (expr);                 \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)

template <class T>
int dalloc(T** p, size_t count) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    HIPC(hipMalloc((void**)p, sizeof(T) * count));
    return LPS_OK;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

hipEvent_t prof_event(lps_ctx* c) {
    if (c->ev_next == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_next++];
}

struct Timed {
    lps_ctx* c;
    const char* name;
    hipEvent_t a = nullptr;
    Timed(lps_ctx* ctx, const char* nm) : c(ctx), name(nm) {
        if (c->prof) {
            a = prof_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {
        if (c->prof && a) {
            hipEvent_t b = prof_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->pending.push_back({name, a, b});
            }
        }
    }
};

void prof_collect(lps_ctx* c) {
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (const ProfRec& r : c->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        auto& acc = c->prof_acc[r.name];
        acc.first += 1;
        acc.second += (double)ms;
    }
    c->pending.clear();
    c->ev_next = 0;
}

SimConst make_const(const lps_ctx* c, float dt, float delta) {
    SimConst k{};
    for (int a = 0; a < 3; a++) k.gravity[a] = c->params.gravity[a];
    k.particle_radius = c->params.particle_radius;
    k.h = c->der.h;
    k.div = c->der.div;
    k.s_max = c->der.s_max;
    k.s_relax = c->der.s_relax;
    k.inv_div = c->der.inv_div;
    k.mass = c->der.mass;
    k.rest_density = c->params.rest_density;
    k.stiffness = c->params.stiffness;
    k.npf = c->params.negative_pressure_factor;
    k.viscosity = c->params.viscosity;
    k.surface_tension = c->params.surface_tension;
    k.st_threshold = c->params.surface_tension_threshold;
    k.restitution = c->params.restitution;
    k.friction = c->params.friction;
    k.dt = dt;
    k.delta = delta;
    k.n = c->n;
    k.nbox = 0;     // per scene (SceneInfo)
    k.nsphere = 0;
    k.nplane = 0;
    k.cell_cap = c->cell_cap;
    return k;
}

int check_params(const lps_params* p) {
    if (!p || p->abi_version != LPS_ABI_VERSION) return LPS_ERR_INVALID;
    if (p->method != LPS_METHOD_SPH && p->method != LPS_METHOD_PCISPH) return LPS_ERR_INVALID;
    if (!(p->particle_radius > 0.f) || !(p->kernel_radius_factor > 0.f) || !(p->rest_density > 0.f))
        return LPS_ERR_INVALID;
    if (p->method == LPS_METHOD_PCISPH && (p->max_iterations < 1 || p->max_iterations > 10000)) return LPS_ERR_INVALID;
    if (p->warm_start < LPS_WARM_NONE || p->warm_start > LPS_WARM_LEARNED_ALWAYS) return LPS_ERR_INVALID;
    // every other float parameter must be finite (a non-finite one would
    // only show later, as a failed step)
    const float fs[] = {p->gravity[0], p->gravity[1], p->gravity[2], p->particle_radius, p->kernel_radius_factor,
                        p->rest_density, p->stiffness, p->negative_pressure_factor, p->viscosity,
                        p->surface_tension, p->surface_tension_threshold, p->restitution, p->friction,
                        p->mass_override, p->max_density_error_ratio};
    for (float f : fs)
        if (!std::isfinite(f)) return LPS_ERR_INVALID;
    return LPS_OK;
}

int upload_tables(lps_ctx* c) {
    if (!c->tables) HIPC(hipMalloc((void**)&c->tables, sizeof(float) * LPS_TABLE_FLOATS));
    HIPC(hipMemcpyAsync(c->tables, c->der.tables, sizeof(float) * 5 * LPS_TABLE_SIZE, hipMemcpyHostToDevice,
                        c->stream));
    HIPC(hipMemcpyAsync(c->tables + 5 * LPS_TABLE_SIZE, c->der.thr, sizeof(float) * LPS_THR_SIZE,
                        hipMemcpyHostToDevice, c->stream));
    return LPS_OK;
}

// Cell-count bound of one scene: particles end every step inside each
// BOUNDARY box shrunk by the particle radius, so that box bounds the grid; the
// first step's grid comes from the initial positions.
uint64_t scene_cell_bound(const lps_ctx* c, const float* x, uint32_t n, const std::vector<lps_box>& boxes) {
    const float h = c->der.h;
    uint64_t init = 1, bound = 0;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            if (!std::isfinite(x[3 * i + a])) continue;
            mn[a] = std::min(mn[a], x[3 * i + a]);
            mx[a] = std::max(mx[a], x[3 * i + a]);
        }
    for (int a = 0; a < 3; a++) {
        const double ex = mx[a] >= mn[a] ? ((double)mx[a] - (double)mn[a]) / (double)h : 0.0;
        init *= (uint64_t)std::min(ex, 4.0e9) + 2;
    }
    init = std::min<uint64_t>(init, 0x7FFFFFF0ull);
    for (const lps_box& b : boxes) {
        if (!b.boundary) continue;
        uint64_t v = 1;
        for (int a = 0; a < 3; a++) v *= (uint64_t)(2.f * std::max(b.half[a] - c->params.particle_radius, 0.f) / h) + 2;
        bound = bound ? std::min(bound, v) : v;
    }
    uint64_t cap = bound ? std::max(bound, init) : init * 8;
    return cap + cap / 8 + 64;
}

int ensure_cells(lps_ctx* c, uint32_t cap) {
    if (cap <= c->cell_cap && c->cell_lb) return LPS_OK;
    c->epoch++;
    int st;
    if ((st = dalloc(&c->cell_lb, (size_t)cap + 1))) return st;
    if ((st = dalloc(&c->run_end, (size_t)cap + 1))) return st;
    if ((st = dalloc(&c->block_max, lps_dev::cell_blocks(cap) + 1))) return st;
    HIPC(hipMemsetAsync(c->run_end, 0, sizeof(uint32_t) * ((size_t)cap + 1), c->stream));
    c->cell_cap = cap;
    const uint32_t bits = cap > 1 ? 32u - (uint32_t)__builtin_clz(cap - 1) : 1u;
    c->passes = (bits + 7) / 8;
    return LPS_OK;
}

int ensure_particles(lps_ctx* c, uint32_t n) {
    if (n <= c->cap_n && c->pos[0]) return LPS_OK;
    c->epoch++;
    int st;
    for (int b = 0; b < 2; b++) {
        if ((st = dalloc(&c->pos[b], (size_t)n + LPS_STENCIL_PAD))) return st;
        if ((st = dalloc(&c->vel[b], n))) return st;
        if ((st = dalloc(&c->vh[b], n))) return st;
        if ((st = dalloc(&c->first[b], n))) return st;
        if ((st = dalloc(&c->id[b], n))) return st;
        if ((st = dalloc(&c->pprev[b], n))) return st;
        if ((st = dalloc(&c->xpred[b], (size_t)n + LPS_STENCIL_PAD))) return st;
    }
    if ((st = dalloc(&c->keysA, n))) return st;
    if ((st = dalloc(&c->valsA, n))) return st;
    if ((st = dalloc(&c->keysB, n))) return st;
    if ((st = dalloc(&c->valsB, n))) return st;
    if ((st = dalloc(&c->skeys, n))) return st;
    if ((st = dalloc(&c->perm, n))) return st;
    if ((st = dalloc(&c->hist, (size_t)256 * lps_dev::radix_tiles(n)))) return st;
    if ((st = dalloc(&c->digit_total, 256))) return st;
    if ((st = dalloc(&c->bbox_part, (size_t)LPS_BBOX_PART * lps_dev::bbox_parts(n)))) return st;
    if ((st = dalloc(&c->posrho, (size_t)n + LPS_STENCIL_PAD))) return st;
    if ((st = dalloc(&c->force, n))) return st;
    // pressure force, pressure and predicted density: three buffers each
    // (a correction's own, the previous one's, the best one's)
    if ((st = dalloc(&c->fp, 3 * n))) return st;
    if ((st = dalloc(&c->press, 3 * n))) return st;
    if ((st = dalloc(&c->q, n))) return st;
    if ((st = dalloc(&c->xq, (size_t)n + LPS_STENCIL_PAD))) return st;
    if ((st = dalloc(&c->rho_pred, 3 * n))) return st;
    if ((st = dalloc(&c->mk_mask, (size_t)LPS_MC_ROWS * n))) return st;
    if ((st = dalloc(&c->mk_ok, n))) return st;
    if ((st = dalloc(&c->mk_cq, (size_t)n + LPS_STENCIL_PAD))) return st;
    // every replaying pass follows its recording pass; until then no slot
    // counts as recorded
    HIPC(hipMemsetAsync(c->mk_ok, 0, sizeof(uint32_t) * n, c->stream));
    dfree(c->feat);
    dfree(c->seed);
    c->cap_n = n;
    return LPS_OK;
}

void free_all(lps_ctx* c) {
    for (int b = 0; b < 2; b++) {
        dfree(c->pos[b]); dfree(c->vel[b]); dfree(c->vh[b]); dfree(c->first[b]);
        dfree(c->id[b]); dfree(c->pprev[b]); dfree(c->xpred[b]);
    }
    dfree(c->keysA); dfree(c->valsA); dfree(c->keysB); dfree(c->valsB);
    dfree(c->skeys); dfree(c->perm); dfree(c->hist); dfree(c->digit_total); dfree(c->bbox_part);
    dfree(c->posrho); dfree(c->force); dfree(c->fp); dfree(c->press); dfree(c->q); dfree(c->xq); dfree(c->rho_pred);
    dfree(c->mk_mask); dfree(c->mk_ok); dfree(c->mk_cq);
    dfree(c->run_end); dfree(c->cell_lb); dfree(c->block_max);
    dfree(c->tables); dfree(c->boxes); dfree(c->spheres); dfree(c->planes); dfree(c->grid);
    dfree(c->d_mlp); dfree(c->d_mlp_packed); dfree(c->feat); dfree(c->seed); dfree(c->coll_x); dfree(c->coll_t);
    dfree(c->d_grad); dfree(c->d_m1); dfree(c->d_m2); dfree(c->d_loss); dfree(c->d_stats);
    dfree(c->d_sinfo); dfree(c->d_sgrid); dfree(c->d_bscene); dfree(c->d_actlist);
    if (c->h_sgrid) (void)hipHostFree(c->h_sgrid);
    c->h_sgrid = nullptr;
    if (c->flags) (void)hipHostFree(c->flags);
    if (c->h_grid) (void)hipHostFree(c->h_grid);
    if (c->h_energy) (void)hipHostFree(c->h_energy);
    c->flags = nullptr;
    c->h_grid = nullptr;
    c->h_energy = nullptr;
}

lps_dev::Scenes scenes(const lps_ctx* c) {
    return lps_dev::Scenes{c->d_sinfo, c->d_sgrid, c->d_bscene, (uint32_t)c->h_sinfo.size(), c->n / LPS_BLOCK,
                           c->d_actlist};
}

// Whether the stencil kernels give each particle a quad of lanes.  By
// default below 512 particle blocks (131,072 slots): one lane per particle
// then fills at most two waves per SIMD.
bool quad_lanes(const lps_ctx* c) { return c->lanes == 4u || (c->lanes == 0u && c->n / LPS_BLOCK < 512u); }

lps_dev::PhysArgs phys_args(lps_ctx* c, const SimConst& k) {
    lps_dev::PhysArgs a;
    a.sc = scenes(c);
    const int s = c->cur;
    a.pos = c->pos[s];
    a.vel = c->vel[s];
    a.vh = c->vh[s];
    a.first = c->first[s];
    a.posrho = c->posrho;
    a.force = c->force;
    a.g = c->grid;
    a.cell_lb = c->cell_lb;
    a.tables = c->tables;
    a.boxes = c->boxes;
    a.spheres = c->spheres;
    a.planes = c->planes;
    a.mk.mask = c->mk_mask;
    a.mk.ok = c->mk_ok;
    a.mk.stride = c->cap_n;
    a.mk.cq = c->mk_cq;
    a.c = k;
    a.quad = quad_lanes(c) ? 1u : 0u;
    return a;
}

// Grid build of one step: keys, sort, reorder (state flips to the other
// buffer), cell table.
void enqueue_grid(lps_ctx* c) {
    const hipStream_t s = c->stream;
    const uint32_t n = c->n;
    {
        Timed t(c, "grid_setup");
        lps_dev::launch_grid_setup(s, c->grid, scenes(c), c->bbox_part, c->der.h, c->cell_cap, c->h_energy);
    }
    {
        Timed t(c, "keys");
        lps_dev::launch_keys(s, c->pos[c->cur], n, c->grid, scenes(c), c->keysA, c->valsA, c->block_max, c->hist);
    }
    for (uint32_t p = 0; p < c->passes; p++) {
        Timed t(c, "radix_pass");
        const bool even = (p & 1u) == 0u;
        lps_dev::launch_radix_pass(s, even ? c->keysA : c->keysB, even ? c->valsA : c->valsB,
                                   even ? c->keysB : c->keysA, even ? c->valsB : c->valsA, n, 8u * p, c->grid,
                                   c->hist, c->digit_total);
    }
    {
        Timed t(c, "reorder");
        lps_dev::ReorderArgs r;
        const int a = c->cur, b = 1 - c->cur;
        r.keysA = c->keysA; r.valsA = c->valsA; r.keysB = c->keysB; r.valsB = c->valsB;
        r.n = n;
        r.g = c->grid;
        r.pos_in = c->pos[a]; r.vel_in = c->vel[a]; r.vh_in = c->vh[a]; r.first_in = c->first[a];
        r.id_in = c->id[a]; r.pprev_in = c->pprev[a];
        r.pos_out = c->pos[b]; r.vel_out = c->vel[b]; r.vh_out = c->vh[b]; r.first_out = c->first[b];
        r.id_out = c->id[b]; r.pprev_out = c->pprev[b];
        r.run_end = c->run_end;
        r.tile_max = c->block_max;
        r.skeys = c->skeys;
        r.perm = c->perm;
        r.sc = scenes(c);
        r.cq = c->mk_cq;
        lps_dev::launch_reorder(s, r);
        c->cur = b;
    }
    {
        Timed t(c, "cell_table");
        lps_dev::launch_cell_table(s, c->run_end, c->grid, c->block_max, c->cell_lb, c->cell_cap);
    }
}

struct StepResult {
    int32_t iterations = 0;
    uint32_t err_bits = 0;
    int32_t converged = 1;
    bool ok = true;  // the loop stopped without an error
};

// ---- learned warm start: when to seed, fallback bookkeeping

bool warm_ready(const lps_ctx* c) {
    return c->params.method == LPS_METHOD_PCISPH &&
           (c->params.warm_start == LPS_WARM_LEARNED || c->params.warm_start == LPS_WARM_LEARNED_ALWAYS) &&
           c->mlp_loaded;
}

// The mode in use (WarmFallback::on), except on probe steps.
bool seed_this_step(const lps_ctx* c) {
    if (!warm_ready(c)) return false;
    if (c->params.warm_start == LPS_WARM_LEARNED_ALWAYS) return true;
    return c->wf.probe() ? !c->wf.on : c->wf.on;
}

void ws_record(lps_ctx* c, bool seeded, int32_t iterations, bool converged) {
    c->wf.record(seeded, converged ? iterations : (int32_t)c->params.max_iterations);
}

// A seeded step that failed counts as max_iterations corrections worse.
void ws_fail(lps_ctx* c) { c->wf.push((double)c->params.max_iterations); }

// Energy judgement of the previous step, once the grid set-up of the current
// one has written the energy entering it (h_energy).  With negative pressures
// disabled an over-predicted seed is never pulled back by the loop and shows
// up as a step that adds mechanical energy; a seeded step whose relative
// energy change exceeds that of the zero start (running mean, at least 0) by
// LPS_WARM_ENERGY_TOL in any scene counts as a failed seed, like a non-finite
// one.
void ws_energy(lps_ctx* c, bool seeded_now) {
    const size_t S = c->h_sinfo.size();
    if (c->e_prev.size() != 2 * S) c->e_prev.assign(2 * S, 0.0);
    if (c->e_prev_valid && warm_ready(c)) {
        double worst = -INFINITY;
        for (size_t k = 0; k < S; k++) {
            const double e0 = c->e_prev[2 * k] + c->e_prev[2 * k + 1];
            const double e1 = c->h_energy[2 * k] + c->h_energy[2 * k + 1];
            const double scale = std::fabs(c->e_prev[2 * k]) + std::fabs(c->e_prev[2 * k + 1]);
            if (scale > 0.0) worst = std::max(worst, (e1 - e0) / scale);
        }
        if (std::isfinite(worst)) {
            if (c->e_prev_seeded) {
                if (worst > std::max(c->wf.ema_de_plain, 0.0) + (double)LPS_WARM_ENERGY_TOL) ws_fail(c);
            } else {
                c->wf.ema_de_plain = 0.7 * c->wf.ema_de_plain + 0.3 * worst;
            }
        }
    }
    for (size_t k = 0; k < 2 * S; k++) c->e_prev[k] = c->h_energy[k];
    c->e_prev_valid = true;
    c->e_prev_seeded = seeded_now;
}

bool capture_features(const lps_ctx* c, bool seeded) {
    return c->params.method == LPS_METHOD_PCISPH && (seeded || (c->coll_cap > 0 && c->coll_stride > 0) || c->capture);
}

int ensure_feat(lps_ctx* c) {
    int st;
    if (!c->feat || !c->seed) c->epoch++;
    if (!c->feat && (st = dalloc(&c->feat, (size_t)c->cap_n * LPS_MLP_IN))) return st;
    if (!c->seed && (st = dalloc(&c->seed, (size_t)c->cap_n))) return st;
    return LPS_OK;
}

lps_dev::MlpArgs mlp_args(const lps_ctx* c) {
    return lps_dev::MlpArgs{c->d_mlp, c->d_mlp + LPS_MLP_PARAMS, c->mlp_tscale};
}

LoopRule loop_rule(const lps_ctx* c) {
    const uint32_t kmin = (uint32_t)std::max(1, c->params.min_iterations);
    const uint32_t kmax = (uint32_t)std::max(c->params.max_iterations, (int32_t)kmin);
    return LoopRule{kmin, kmax, (uint32_t)std::max(0, c->params.stall_iterations),
                    c->params.max_density_error_ratio};
}

// Host flag slot of correction kk: {state, corrections, error bits,
// converged}, written by the device (word 0 last).  The host marks it
// pending before the correction is enqueued and polls word 0.
constexpr uint32_t FLAG_PENDING = 0xFFFFFFFFu;
volatile uint32_t* flag_slot(lps_ctx* c, uint32_t kk) {
    return c->flags + LPS_FLAG_WORDS * (kk % LPS_MAX_ITER_SLOTS);
}

// Waits until the device has written slot kk; false after ~10 s without it.
bool wait_flag(lps_ctx* c, uint32_t kk) {
    volatile uint32_t* f = flag_slot(c, kk);
    for (uint64_t spin = 0; f[0] == FLAG_PENDING; spin++) {
        if ((spin & 1023u) == 1023u) {
            // the correction may still be queued behind a long kernel
            if (hipStreamQuery(c->stream) == hipSuccess && f[0] == FLAG_PENDING) return false;
            if (spin > (1ull << 32)) return false;
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return true;
}

// One PCISPH correction kk: predicted density (+ loop rule), the seed after
// the first one, pressure force and next prediction.  `mark` sets the host
// flag slot pending (outside captured graphs; the caller marks those).
void enqueue_iter(lps_ctx* c, const lps_dev::PhysArgs& a, uint32_t kk, const LoopRule& rule, float* feat,
                  bool seeded, float delta, bool mark) {
    const hipStream_t s = c->stream;
    if (mark) flag_slot(c, kk)[0] = FLAG_PENDING;
    {
        Timed t(c, "pci_density");
        lps_dev::launch_pci_density(s, a, c->xpred[(kk - 1u) & 1u], c->rho_pred, c->press, c->xq, kk, feat,
                                    c->corr_threads);
    }
    if (kk == 1u && seeded) {
        // the first correction's pressure update is replaced by the seed
        Timed t(c, "mlp_seed");
        lps_dev::launch_mlp_seed(s, c->d_mlp_packed, c->mlp_tscale, c->feat, c->cap_n, c->n, c->press, c->xq,
                                 c->seed, c->rho_pred, c->pprev[c->cur], LPS_WARM_KAPPA,
                                 delta * c->params.rest_density, c->params.negative_pressure_factor == 0.f ? 1 : 0,
                                 c->grid, scenes(c), rule, c->params.rest_density);
    }
    {
        Timed t(c, "pci_pressure");
        lps_dev::launch_pci_pressure(s, a, c->xq, c->rho_pred, c->fp, c->xpred[kk & 1u], kk, rule,
                                     c->flags + LPS_FLAG_WORDS * (kk % LPS_MAX_ITER_SLOTS), c->corr_threads);
    }
}

// The fixed part of a PCISPH step: grid build (the state flips to the other
// buffer), density, forces + first prediction, corrections 1..min.
void enqueue_prefix(lps_ctx* c, float dt, float delta, bool seeded, float* feat, bool events) {
    const hipStream_t s = c->stream;
    enqueue_grid(c);
    const SimConst k = make_const(c, dt, delta);
    const lps_dev::PhysArgs a = phys_args(c, k);
    {
        Timed t(c, "density");
        lps_dev::launch_density(s, a, c->press, c->q);
    }
    {
        Timed t(c, "forces");
        lps_dev::launch_forces(s, a, c->xpred[0], c->pprev[c->cur], feat);
    }
    const LoopRule rule = loop_rule(c);
    for (uint32_t kk = 1; kk <= rule.min_iter; kk++) enqueue_iter(c, a, kk, rule, feat, seeded, delta, events);
}

// The prefix as a captured HIP graph, replayed when nothing it depends on
// changed.  Returns false when graphs are unavailable (nothing enqueued).
bool launch_prefix_graph(lps_ctx* c, float dt, float delta, bool seeded, bool feat_on, float* feat) {
    const int cur0 = c->cur;
    lps_ctx::PrefixGraph& pg = c->graphs[cur0][seeded ? 1 : 0][feat_on ? 1 : 0];
    if (!pg.exec || pg.epoch != c->epoch || pg.dt != dt) {
        if (pg.exec) (void)hipGraphExecDestroy(pg.exec);
        pg.exec = nullptr;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        enqueue_prefix(c, dt, delta, seeded, feat, false);
        c->cur = cur0;  // the capture enqueued nothing; the launch below flips it
        hipGraph_t graph = nullptr;
        const bool ok = hipStreamEndCapture(c->stream, &graph) == hipSuccess && graph &&
                        hipGraphInstantiate(&pg.exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
            pg.exec = nullptr;
            (void)hipGetLastError();
            return false;
        }
        pg.epoch = c->epoch;
        pg.dt = dt;
    }
    if (hipGraphLaunch(pg.exec, c->stream) != hipSuccess) return false;
    c->cur = 1 - cur0;
    return true;
}

StepResult enqueue_step(lps_ctx* c, float dt, bool seeded, bool feat_on) {
    StepResult res;
    const hipStream_t s = c->stream;
    const bool pci = c->params.method == LPS_METHOD_PCISPH;
    const float delta = pci ? lps_compute_delta(&c->params, &c->der, dt) : 0.f;
    if (!pci) {
        enqueue_grid(c);
        const SimConst k = make_const(c, dt, delta);
        lps_dev::PhysArgs a = phys_args(c, k);
        {
            Timed t(c, "density");
            lps_dev::launch_density(s, a, c->press, c->q);
        }
        {
            Timed t(c, "forces");
            lps_dev::launch_forces(s, a, nullptr, nullptr, nullptr);
        }
        {
            Timed t(c, "pressure_sph");
            lps_dev::launch_pressure_sph(s, a, c->q);
        }
        {
            Timed t(c, "integrate");
            lps_dev::launch_integrate(s, a, nullptr, c->pprev[c->cur], c->press, c->bbox_part);
        }
        c->last_fp_valid = 0;
        c->feat_valid = c->seed_valid = false;
        return res;
    }
    float* feat = feat_on ? c->feat : nullptr;
    const LoopRule rule = loop_rule(c);
    const uint32_t kmin = rule.min_iter, kmax = rule.max_iter;
    for (uint32_t kk = 1; kk <= kmin; kk++) flag_slot(c, kk)[0] = FLAG_PENDING;
    bool graphed = c->graphs_on && !c->prof && launch_prefix_graph(c, dt, delta, seeded, feat_on, feat);
    if (!graphed) {
        if (c->graphs_on && !c->prof) c->graphs_on = false;  // capture or launch failed: stay eager
        enqueue_prefix(c, dt, delta, seeded, feat, true);
    }
    const SimConst k = make_const(c, dt, delta);
    lps_dev::PhysArgs a = phys_args(c, k);
    // Corrections beyond min: `lookahead` corrections stay queued ahead of
    // the result being waited for; queued corrections past the stop return
    // at once on the device.  With automatic lanes, a correction enqueued
    // after the host learned that fewer than 512 particle blocks still run
    // (the late corrections of an ensemble) gives each particle a quad.
    uint32_t k_enq = kmin, k_known = 0;
    lps_dev::PhysArgs ac = a;
    while (true) {
        const uint32_t kw = std::max(k_known + 1u, kmin);
        while (k_enq < kmax && k_enq < kw + c->lookahead) enqueue_iter(c, ac, ++k_enq, rule, feat, seeded, delta, true);
        if (!wait_flag(c, kw)) {
            res.ok = false;  // the device never answered: reported as a HIP error
            res.iterations = -1;
            break;
        }
        volatile uint32_t* f = flag_slot(c, kw);
        const uint32_t state = f[0];
        res.iterations = (int32_t)f[1];
        res.err_bits = f[2];
        res.converged = (int32_t)f[3];
        res.ok = state == 1u;
        if (state != 0u) break;
        k_known = kw;
        if (kw >= kmax) break;
        if (c->lanes == 0u && f[4] < 512u) ac.quad = 1u;
    }
    {
        Timed t(c, "integrate");
        lps_dev::launch_integrate(s, a, c->fp, c->pprev[c->cur], c->press, c->bbox_part);
    }
    if (feat_on && res.ok && c->coll_stride > 0 && c->coll_rows < c->coll_cap) {
        // training rows: every stride-th particle's features and the
        // pressure the step applied (copied to pprev by the integration)
        const uint32_t off = c->coll_offset % c->coll_stride;
        for (const SceneInfo& si : c->h_sinfo) {
            const uint32_t n = si.count;
            uint64_t count = n > off ? ((uint64_t)(n - off) + c->coll_stride - 1u) / c->coll_stride : 0u;
            count = std::min<uint64_t>(count, c->coll_cap - c->coll_rows);
            lps_dev::launch_collect(s, c->feat + si.slot0, c->cap_n, c->pprev[c->cur] + si.slot0,
                                    delta * c->params.rest_density, n, c->coll_stride, off, (uint32_t)count,
                                    c->coll_x, c->coll_t, c->coll_rows);
            c->coll_rows += count;
        }
        c->coll_offset = (off + 1u) % c->coll_stride;
    }
    c->last_fp_valid = 1;
    c->feat_valid = feat_on;
    c->seed_valid = seeded;
    return res;
}

// Concatenate the scenes' colliders into the device arrays and upload the
// scene constants (collider ranges, gravity).
int upload_scenes(lps_ctx* c) {
    const size_t S = c->h_sinfo.size();
    std::vector<lps_box> boxes;
    std::vector<lps_sphere> spheres;
    std::vector<lps_plane> planes;
    uint32_t real0 = 0;
    for (size_t k = 0; k < S; k++) {
        SceneInfo& si = c->h_sinfo[k];
        si.real0 = real0;
        real0 += si.count;
        si.box0 = (uint32_t)boxes.size();
        si.nbox = (uint32_t)c->s_boxes[k].size();
        si.sphere0 = (uint32_t)spheres.size();
        si.nsphere = (uint32_t)c->s_spheres[k].size();
        si.plane0 = (uint32_t)planes.size();
        si.nplane = (uint32_t)c->s_planes[k].size();
        boxes.insert(boxes.end(), c->s_boxes[k].begin(), c->s_boxes[k].end());
        spheres.insert(spheres.end(), c->s_spheres[k].begin(), c->s_spheres[k].end());
        planes.insert(planes.end(), c->s_planes[k].begin(), c->s_planes[k].end());
        for (int a = 0; a < 3; a++)
            si.gravity[a] = c->s_gravity.empty() ? c->params.gravity[a] : c->s_gravity[3 * k + a];
    }
    if (boxes.size() > LPS_MAX_COLLIDERS || spheres.size() > LPS_MAX_COLLIDERS || planes.size() > LPS_MAX_COLLIDERS)
        return LPS_ERR_INVALID;
    if (!boxes.empty())
        HIPC(hipMemcpyAsync(c->boxes, boxes.data(), sizeof(lps_box) * boxes.size(), hipMemcpyHostToDevice, c->stream));
    if (!spheres.empty())
        HIPC(hipMemcpyAsync(c->spheres, spheres.data(), sizeof(lps_sphere) * spheres.size(), hipMemcpyHostToDevice,
                            c->stream));
    if (!planes.empty())
        HIPC(hipMemcpyAsync(c->planes, planes.data(), sizeof(lps_plane) * planes.size(), hipMemcpyHostToDevice,
                            c->stream));
    if (c->d_sinfo)
        HIPC(hipMemcpyAsync(c->d_sinfo, c->h_sinfo.data(), sizeof(SceneInfo) * S, hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    return LPS_OK;
}

// Real particles of every scene, in slot order, from a full slot array.
template <class T>
void compact_slots(const lps_ctx* c, const std::vector<T>& full, T* out, size_t width) {
    size_t r = 0;
    for (const SceneInfo& si : c->h_sinfo) {
        std::memcpy(out + r * width, full.data() + (size_t)si.slot0 * width, sizeof(T) * si.count * width);
        r += si.count;
    }
}

// Copies the slots of every scene from the buffer (of three, `slots` apart)
// that the scene's loop applied in the last step into `full` (a whole slot
// array, for compact_slots).  PCISPH only; the caller has synchronised.
template <class T>
int gather_applied(lps_ctx* c, const T* dev3, std::vector<T>& full) {
    const size_t S = c->h_sinfo.size();
    HIPC(hipMemcpy(c->h_sgrid, c->d_sgrid, sizeof(SceneGrid) * S, hipMemcpyDeviceToHost));
    full.resize(c->n);
    for (size_t k = 0; k < S; k++) {
        const SceneInfo& si = c->h_sinfo[k];
        const uint32_t b = std::min<uint32_t>(c->h_sgrid[k].final_buf, 2u);
        HIPC(hipMemcpy(full.data() + si.slot0, dev3 + (size_t)b * c->cap_n + si.slot0, sizeof(T) * si.count,
                       hipMemcpyDeviceToHost));
    }
    return LPS_OK;
}

}  // namespace

// ================================================================= C ABI ==
extern "C" {

const char* lps_status_string(int status) {
    switch (status) {
        case LPS_OK: return "LPS_OK";
        case LPS_ERR_INVALID: return "LPS_ERR_INVALID";
        case LPS_ERR_NO_DEVICE: return "LPS_ERR_NO_DEVICE";
        case LPS_ERR_OOM: return "LPS_ERR_OOM";
        case LPS_ERR_NONFINITE: return "LPS_ERR_NONFINITE";
        case LPS_ERR_GRID_OVERFLOW: return "LPS_ERR_GRID_OVERFLOW";
        case LPS_ERR_HIP: return "LPS_ERR_HIP";
        case LPS_ERR_STATE: return "LPS_ERR_STATE";
        default: return "LPS_ERR_UNKNOWN";
    }
}

int lps_device_info(char* name, size_t cap, int32_t* cu_count) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return LPS_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return LPS_ERR_NO_DEVICE;
    if (name && cap) {
        std::snprintf(name, cap, "%s", prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    return LPS_OK;
}

int lps_create(const lps_params* params, lps_ctx** out) {
    if (!out) return LPS_ERR_INVALID;
    *out = nullptr;
    int st = check_params(params);
    if (st) return st;
    char arch[256] = {0};
    if (lps_device_info(arch, sizeof(arch), nullptr) != LPS_OK) return LPS_ERR_NO_DEVICE;
    if (std::strncmp(arch, "gfx950", 6) != 0) return LPS_ERR_NO_DEVICE;
    lps_ctx* c = new (std::nothrow) lps_ctx();
    if (!c) return LPS_ERR_OOM;
    c->params = *params;
    lps_compute_derived(&c->params, &c->der);
    (void)hipGetDevice(&c->device);
    resolve_scratch(c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LPS_ERR_HIP;
    }
    bool ok = hipMalloc((void**)&c->grid, sizeof(DevGrid)) == hipSuccess &&
              hipHostMalloc((void**)&c->flags, sizeof(uint32_t) * LPS_FLAG_WORDS * LPS_MAX_ITER_SLOTS,
                            hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess &&
              hipHostMalloc((void**)&c->h_grid, sizeof(DevGrid), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&c->h_energy, sizeof(double) * 2 * LPS_MAX_SCENES,
                            hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess &&
              hipMemsetAsync(c->grid, 0, sizeof(DevGrid), c->stream) == hipSuccess &&
              upload_tables(c) == LPS_OK;
    if (ok) {
        ok = hipMalloc((void**)&c->boxes, sizeof(lps_box) * LPS_MAX_COLLIDERS) == hipSuccess &&
             hipMalloc((void**)&c->spheres, sizeof(lps_sphere) * LPS_MAX_COLLIDERS) == hipSuccess &&
             hipMalloc((void**)&c->planes, sizeof(lps_plane) * LPS_MAX_COLLIDERS) == hipSuccess;
    }
    if (!ok) {
        lps_destroy(c);
        return LPS_ERR_OOM;
    }
    std::memset(c->flags, 0, sizeof(uint32_t) * LPS_FLAG_WORDS * LPS_MAX_ITER_SLOTS);
    *out = c;
    return LPS_OK;
}

int lps_destroy(lps_ctx* c) {
    if (!c) return LPS_ERR_INVALID;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto& a : c->graphs)
        for (auto& b : a)
            for (auto& g : b)
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LPS_OK;
}

int lps_set_params(lps_ctx* c, const lps_params* p) {
    if (!c) return LPS_ERR_INVALID;
    int st = check_params(p);
    if (st) return st;
    c->params = *p;
    c->epoch++;
    lps_compute_derived(&c->params, &c->der);
    st = upload_tables(c);
    return st ? st : upload_scenes(c);
}

int lps_get_derived(lps_ctx* c, float* h, float* div, float* mass, float* tables) {
    if (!c) return LPS_ERR_INVALID;
    if (h) *h = c->der.h;
    if (div) *div = c->der.div;
    if (mass) *mass = c->der.mass;
    if (tables) std::memcpy(tables, c->der.tables, sizeof(c->der.tables));
    return LPS_OK;
}

int lps_index_thresholds(const lps_params* p, float* thr) {
    if (check_params(p) != LPS_OK || !thr) return LPS_ERR_INVALID;
    Derived* d = new (std::nothrow) Derived();
    if (!d) return LPS_ERR_OOM;
    lps_compute_derived(p, d);
    std::memcpy(thr, d->thr, sizeof(float) * LPS_THR_SIZE);
    delete d;
    return LPS_OK;
}

int lps_get_delta(lps_ctx* c, float dt, float* delta) {
    if (!c || !delta) return LPS_ERR_INVALID;
    *delta = lps_compute_delta(&c->params, &c->der, dt);
    return LPS_OK;
}

int lps_set_colliders(lps_ctx* c, const lps_box* boxes, uint32_t nbox, const lps_sphere* spheres, uint32_t nsphere,
                      const lps_plane* planes, uint32_t nplane) {
    if (!c || nbox > LPS_MAX_COLLIDERS || nsphere > LPS_MAX_COLLIDERS || nplane > LPS_MAX_COLLIDERS)
        return LPS_ERR_INVALID;
    if ((nbox && !boxes) || (nsphere && !spheres) || (nplane && !planes)) return LPS_ERR_INVALID;
    const size_t S = c->h_sinfo.size();
    if (S * (nbox + nsphere + nplane) > 3 * (size_t)LPS_MAX_COLLIDERS || S * nbox > LPS_MAX_COLLIDERS ||
        S * nsphere > LPS_MAX_COLLIDERS || S * nplane > LPS_MAX_COLLIDERS)
        return LPS_ERR_INVALID;
    for (size_t k = 0; k < S; k++) {
        c->s_boxes[k].assign(boxes, boxes + nbox);
        c->s_spheres[k].assign(spheres, spheres + nsphere);
        c->s_planes[k].assign(planes, planes + nplane);
    }
    return upload_scenes(c);
}

int lps_set_scene_colliders(lps_ctx* c, uint32_t scene, const lps_box* boxes, uint32_t nbox,
                            const lps_sphere* spheres, uint32_t nsphere, const lps_plane* planes, uint32_t nplane) {
    if (!c || scene >= c->h_sinfo.size()) return LPS_ERR_INVALID;
    if ((nbox && !boxes) || (nsphere && !spheres) || (nplane && !planes)) return LPS_ERR_INVALID;
    std::vector<lps_box> ob = c->s_boxes[scene];
    std::vector<lps_sphere> os = c->s_spheres[scene];
    std::vector<lps_plane> op = c->s_planes[scene];
    c->s_boxes[scene].assign(boxes, boxes + nbox);
    c->s_spheres[scene].assign(spheres, spheres + nsphere);
    c->s_planes[scene].assign(planes, planes + nplane);
    const int st = upload_scenes(c);
    if (st == LPS_ERR_INVALID) {  // too many colliders in total: keep the previous set
        c->s_boxes[scene] = ob;
        c->s_spheres[scene] = os;
        c->s_planes[scene] = op;
    }
    return st;
}

int lps_set_scenes(lps_ctx* c, uint32_t nscenes, const uint32_t* counts, const float* gravity, const float* x,
                   const float* v, const float* vhalf, const uint8_t* first) {
    if (!c || !counts || !x || nscenes == 0 || nscenes > LPS_MAX_SCENES) return LPS_ERR_INVALID;
    uint64_t total = 0, slots = 0;
    for (uint32_t k = 0; k < nscenes; k++) {
        if (counts[k] == 0) return LPS_ERR_INVALID;
        total += counts[k];
        slots += ((uint64_t)counts[k] + LPS_BLOCK - 1u) / LPS_BLOCK * LPS_BLOCK;
    }
    if (slots > LPS_MAX_PARTICLES) return LPS_ERR_INVALID;
    if (gravity)
        for (uint32_t k = 0; k < 3 * nscenes; k++)
            if (!std::isfinite(gravity[k])) return LPS_ERR_INVALID;
    const uint32_t n = (uint32_t)slots;
    int st = ensure_particles(c, n);
    if (st) return st;
    // layout; colliders of scenes that existed before are kept
    c->h_sinfo.assign(nscenes, SceneInfo{});
    c->s_boxes.resize(nscenes);
    c->s_spheres.resize(nscenes);
    c->s_planes.resize(nscenes);
    if (gravity) c->s_gravity.assign(gravity, gravity + 3 * nscenes);
    else c->s_gravity.clear();
    uint32_t slot = 0;
    uint64_t cells = 0;
    size_t off = 0;
    for (uint32_t k = 0; k < nscenes; k++) {
        SceneInfo& si = c->h_sinfo[k];
        si.slot0 = slot;
        si.count = counts[k];
        si.padded = (counts[k] + LPS_BLOCK - 1u) / LPS_BLOCK * LPS_BLOCK;
        si.block0 = slot / LPS_BLOCK;
        slot += si.padded;
        cells += scene_cell_bound(c, x + 3 * off, counts[k], c->s_boxes[k]) + 1u;
        off += counts[k];
    }
    if ((st = ensure_cells(c, (uint32_t)std::min<uint64_t>(cells, 0x7FFFFFF0ull)))) return st;
    if ((st = dalloc(&c->d_sinfo, nscenes))) return st;
    if ((st = dalloc(&c->d_sgrid, nscenes))) return st;
    if ((st = dalloc(&c->d_bscene, n / LPS_BLOCK))) return st;
    if ((st = dalloc(&c->d_actlist, 2 * (size_t)(n / LPS_BLOCK)))) return st;
    if (c->h_sgrid) (void)hipHostFree(c->h_sgrid);
    c->h_sgrid = nullptr;
    HIPC(hipHostMalloc((void**)&c->h_sgrid, sizeof(SceneGrid) * nscenes, hipHostMallocDefault));
    // particle arrays in slot order; empty slots at rest at the origin
    std::vector<float4> buf(n);
    auto up4 = [&](float4* dst, const float* src) -> int {
        size_t r0 = 0;  // index of the scene's first particle in the caller's arrays
        for (uint32_t k = 0; k < nscenes; k++) {
            const SceneInfo& si = c->h_sinfo[k];
            for (uint32_t i = 0; i < si.padded; i++) {
                const size_t r = r0 + i;
                buf[si.slot0 + i] = (i < si.count && src)
                                        ? make_float4(src[3 * r], src[3 * r + 1], src[3 * r + 2], 0.f)
                                        : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            r0 += si.count;
        }
        HIPC(hipMemcpy(dst, buf.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        return LPS_OK;
    };
    c->cur = 0;
    if ((st = up4(c->pos[0], x))) return st;
    if ((st = up4(c->vel[0], v))) return st;
    if ((st = up4(c->vh[0], vhalf))) return st;
    std::vector<uint8_t> f(n, 1);
    std::vector<uint32_t> ids(n, 0xFFFFFFFFu);
    std::vector<uint8_t> bscene(n / LPS_BLOCK);
    {
        uint32_t r = 0;
        for (uint32_t k = 0; k < nscenes; k++) {
            const SceneInfo& si = c->h_sinfo[k];
            for (uint32_t i = 0; i < si.count; i++, r++) {
                ids[si.slot0 + i] = r;
                if (first) f[si.slot0 + i] = first[r] ? 1 : 0;
            }
            for (uint32_t b = 0; b < si.padded / LPS_BLOCK; b++) bscene[si.block0 + b] = (uint8_t)k;
        }
    }
    HIPC(hipMemcpy(c->first[0], f.data(), n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->id[0], ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_bscene, bscene.data(), n / LPS_BLOCK, hipMemcpyHostToDevice));
    HIPC(hipMemset(c->pprev[0], 0, sizeof(float) * n));
    HIPC(hipMemset(c->d_sgrid, 0, sizeof(SceneGrid) * nscenes));
    c->n = n;
    c->n_real = (uint32_t)total;
    c->next_id = (uint32_t)total;
    c->have_step = false;
    c->epoch++;
    if ((st = upload_scenes(c))) return st;
    lps_dev::launch_bbox(c->stream, c->pos[0], c->vel[0], c->der.mass, n, scenes(c), c->bbox_part);
    HIPC(hipStreamSynchronize(c->stream));
    c->e_prev_valid = false;
    c->wf.prev_seeded = -1;  // a new state: its first step pairs with no earlier one
    return LPS_OK;
}

int lps_set_particles(lps_ctx* c, const float* x, const float* v, const float* vhalf, const uint8_t* first,
                      uint32_t n) {
    if (!c || !x || n == 0) return LPS_ERR_INVALID;
    // the slot bound is lps_set_scenes's
    return lps_set_scenes(c, 1u, &n, nullptr, x, v, vhalf, first);
}

int lps_add_particles(lps_ctx* c, const float* x, const float* v, const float* vhalf, const uint8_t* first,
                      uint32_t n) {
    if (!c || !x || n == 0) return LPS_ERR_INVALID;
    if (c->h_sinfo.size() != 1u || c->n_real == 0) return LPS_ERR_STATE;
    const uint32_t old = c->n_real;
    const uint64_t total = (uint64_t)old + n;
    const uint64_t slots = (total + LPS_BLOCK - 1u) / LPS_BLOCK * LPS_BLOCK;
    if (slots > LPS_MAX_PARTICLES) return LPS_ERR_INVALID;
    for (uint64_t k = 0; k < 3ull * n; k++)
        if (!std::isfinite(x[k]) || (v && !std::isfinite(v[k])) || (vhalf && !std::isfinite(vhalf[k])))
            return LPS_ERR_INVALID;
    const hipStream_t s = c->stream;
    HIPC(hipStreamSynchronize(s));
    int st;
    if (slots > c->cap_n) {
        // grow by half again, keeping the current state (slots 0 .. old)
        uint64_t cap = std::max<uint64_t>(slots, (uint64_t)c->cap_n + c->cap_n / 2u);
        cap = std::min<uint64_t>((cap + LPS_BLOCK - 1u) / LPS_BLOCK * LPS_BLOCK,
                                 (uint64_t)LPS_MAX_PARTICLES / LPS_BLOCK * LPS_BLOCK);
        const int b = c->cur;
        float4 *sp = nullptr, *sv = nullptr, *sh = nullptr;
        uint8_t* sf = nullptr;
        uint32_t* si = nullptr;
        float* spp = nullptr;
        auto release = [&]() { dfree(sp); dfree(sv); dfree(sh); dfree(sf); dfree(si); dfree(spp); };
        if ((st = dalloc(&sp, old)) || (st = dalloc(&sv, old)) || (st = dalloc(&sh, old)) || (st = dalloc(&sf, old)) ||
            (st = dalloc(&si, old)) || (st = dalloc(&spp, old))) {
            release();
            return st;
        }
        hipError_t e = hipSuccess;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
        };
        cp(sp, c->pos[b], sizeof(float4) * old); cp(sv, c->vel[b], sizeof(float4) * old);
        cp(sh, c->vh[b], sizeof(float4) * old); cp(sf, c->first[b], old);
        cp(si, c->id[b], sizeof(uint32_t) * old); cp(spp, c->pprev[b], sizeof(float) * old);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess || (st = ensure_particles(c, (uint32_t)cap))) {
            release();
            return e != hipSuccess ? hip_status(e) : st;
        }
        c->cur = 0;
        cp(c->pos[0], sp, sizeof(float4) * old); cp(c->vel[0], sv, sizeof(float4) * old);
        cp(c->vh[0], sh, sizeof(float4) * old); cp(c->first[0], sf, old);
        cp(c->id[0], si, sizeof(uint32_t) * old); cp(c->pprev[0], spp, sizeof(float) * old);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        release();
        if (e != hipSuccess) return hip_status(e);
    }
    // the new particles take slots old .. total - 1 of the current state
    const int b = c->cur;
    std::vector<float4> buf(n);
    auto up4 = [&](float4* dst, const float* src) -> int {
        for (uint32_t i = 0; i < n; i++)
            buf[i] = src ? make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
        HIPC(hipMemcpy(dst + old, buf.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        return LPS_OK;
    };
    if ((st = up4(c->pos[b], x)) || (st = up4(c->vel[b], v)) || (st = up4(c->vh[b], vhalf))) return st;
    std::vector<uint8_t> f(n, 1);
    if (first)
        for (uint32_t i = 0; i < n; i++) f[i] = first[i] ? 1 : 0;
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; i++) ids[i] = c->next_id + i;
    HIPC(hipMemcpy(c->first[b] + old, f.data(), n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->id[b] + old, ids.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPC(hipMemset(c->pprev[b] + old, 0, sizeof(float) * n));
    c->next_id += n;
    // layout: one scene over the grown slot range
    const uint32_t nslots = (uint32_t)slots;
    SceneInfo& si = c->h_sinfo[0];
    si.count = (uint32_t)total;
    si.padded = nslots;
    if (nslots != c->n) {
        if ((st = dalloc(&c->d_bscene, nslots / LPS_BLOCK))) return st;
        if ((st = dalloc(&c->d_actlist, 2 * (size_t)(nslots / LPS_BLOCK)))) return st;
        HIPC(hipMemset(c->d_bscene, 0, nslots / LPS_BLOCK));
    }
    c->n = nslots;
    c->n_real = (uint32_t)total;
    c->have_step = false;
    c->last_fp_valid = 0;
    c->feat_valid = c->seed_valid = false;
    c->epoch++;
    if ((st = upload_scenes(c))) return st;
    lps_dev::launch_bbox(s, c->pos[b], c->vel[b], c->der.mass, nslots, scenes(c), c->bbox_part);
    HIPC(hipStreamSynchronize(s));
    c->e_prev_valid = false;
    return LPS_OK;
}

int lps_copy_positions_device(lps_ctx* c, float* dst, uint32_t stride, uint32_t n, uint64_t capacity_bytes) {
    if (!c || !dst || stride < 3u || n != c->n_real || n == 0) return LPS_ERR_INVALID;
    const uint64_t need = ((uint64_t)(n - 1u) * stride + 3u) * sizeof(float);
    if (capacity_bytes < need) return LPS_ERR_INVALID;
    // the destination must be device memory of at least `need` bytes
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, dst) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return LPS_ERR_INVALID;
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)dst) == hipSuccess) {
        if ((const char*)dst + need > (const char*)base + size) return LPS_ERR_INVALID;
    } else {
        (void)hipGetLastError();
    }
    lps_dev::launch_export_positions(c->stream, c->pos[c->cur], c->n, scenes(c), dst, stride);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    return LPS_OK;
}

int lps_get_scene_count(lps_ctx* c, uint32_t* nscenes) {
    if (!c || !nscenes) return LPS_ERR_INVALID;
    *nscenes = (uint32_t)c->h_sinfo.size();
    return LPS_OK;
}

int lps_get_scene_stats(lps_ctx* c, lps_step_stats* stats, uint32_t nscenes) {
    if (!c || !stats || nscenes != c->h_sinfo.size()) return LPS_ERR_INVALID;
    if (!c->have_step) return LPS_ERR_STATE;
    HIPC(hipMemcpyAsync(c->h_sgrid, c->d_sgrid, sizeof(SceneGrid) * nscenes, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipMemcpyAsync(c->h_grid, c->grid, sizeof(DevGrid), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < nscenes; k++) {
        const SceneGrid& sg = c->h_sgrid[k];
        lps_step_stats& o = stats[k];
        const uint32_t it = sg.stop_iter;
        o.iterations = (int32_t)it;
        o.max_density_error_ratio =
            it ? __builtin_bit_cast(float, sg.final_bits) / c->params.rest_density : 0.f;
        o.warm_start_used = c->seed_valid ? 1 : 0;
        o.status = (int32_t)c->h_grid->status;
        o.converged = (int32_t)sg.converged;
    }
    return LPS_OK;
}

int lps_get_count(lps_ctx* c, uint32_t* n) {
    if (!c || !n) return LPS_ERR_INVALID;
    *n = c->n_real;
    return LPS_OK;
}

int lps_get_particles(lps_ctx* c, float* x, float* v, float* vhalf, uint8_t* first, uint32_t* id, uint32_t n) {
    if (!c || n != c->n_real || n == 0) return LPS_ERR_INVALID;
    HIPC(hipStreamSynchronize(c->stream));
    const uint32_t ns = c->n;
    std::vector<float4> buf(ns), real(n);
    auto down4 = [&](float* dst, const float4* src) -> int {
        if (!dst) return LPS_OK;
        HIPC(hipMemcpy(buf.data(), src, sizeof(float4) * ns, hipMemcpyDeviceToHost));
        compact_slots(c, buf, real.data(), 1);
        for (uint32_t i = 0; i < n; i++) {
            dst[3 * i] = real[i].x;
            dst[3 * i + 1] = real[i].y;
            dst[3 * i + 2] = real[i].z;
        }
        return LPS_OK;
    };
    int st;
    if ((st = down4(x, c->pos[c->cur]))) return st;
    if ((st = down4(v, c->vel[c->cur]))) return st;
    if ((st = down4(vhalf, c->vh[c->cur]))) return st;
    if (first) {
        std::vector<uint8_t> f(ns);
        HIPC(hipMemcpy(f.data(), c->first[c->cur], ns, hipMemcpyDeviceToHost));
        compact_slots(c, f, first, 1);
    }
    if (id) {
        std::vector<uint32_t> f(ns);
        HIPC(hipMemcpy(f.data(), c->id[c->cur], sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
        compact_slots(c, f, id, 1);
    }
    return LPS_OK;
}

int lps_step(lps_ctx* c, float dt, uint32_t nsteps, lps_step_stats* stats) {
    if (!c) return LPS_ERR_INVALID;
    if (c->n_real == 0) return LPS_ERR_STATE;
    if (c->n == 0) return LPS_ERR_STATE;
    if (!(dt > 0.f) || !std::isfinite(dt)) return LPS_ERR_INVALID;
    const bool pci = c->params.method == LPS_METHOD_PCISPH;
    // A failed step leaves the state as it was (its kernels no-op and the
    // reorder copies the identity), so after a grid overflow the cell table is
    // grown and the remaining steps are issued again.
    HIPC(hipMemcpyAsync(c->h_grid, c->grid, sizeof(DevGrid), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    const uint32_t base = c->h_grid->steps_ok;
    uint32_t completed = 0;
    int retries = 0;
    bool retry_plain = false;
    while (completed < nsteps) {
        const uint32_t batch = pci ? 1u : nsteps - completed;
        const bool seeded = pci && !retry_plain && seed_this_step(c);
        const bool feat_on = capture_features(c, seeded);
        if (feat_on) {
            int st = ensure_feat(c);
            if (st) return st;
        }
        StepResult last;
        for (uint32_t b = 0; b < batch; b++) {
            StepResult r = enqueue_step(c, dt, seeded, feat_on);
            last = r;
            if (stats) {
                lps_step_stats& o = stats[completed + b];
                o.iterations = r.iterations;
                o.max_density_error_ratio = __builtin_bit_cast(float, r.err_bits) / c->params.rest_density;
                o.warm_start_used = seeded ? 1 : 0;
                o.status = LPS_OK;
                o.converged = r.converged;
            }
        }
        c->steps = c->steps + batch;
        c->have_step = true;
        if (pci && last.ok) {
            // the correction loop ended normally, which the host has already
            // seen; no synchronisation before the next step.  The grid set-up
            // of this step has written the energy entering it.
            completed += 1;
            retry_plain = false;
            ws_energy(c, seeded);
            if (warm_ready(c)) ws_record(c, seeded, last.iterations, last.converged != 0);
            continue;
        }
        c->e_prev_valid = false;
        HIPC(hipMemcpyAsync(c->h_grid, c->grid, sizeof(DevGrid), hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        completed = c->h_grid->steps_ok - base;
        const uint32_t status = c->h_grid->status;
        if (status == LPS_ERR_GRID_OVERFLOW) {
            const uint32_t need = c->h_grid->cells_needed;
            if (need >= 0x7FFFFFF0u || ++retries > 8) return LPS_ERR_GRID_OVERFLOW;
            int st = ensure_cells(c, need + need / 4 + 64);
            if (st) return st;
            continue;
        }
        if (status == LPS_ERR_NONFINITE && seeded && retries < 8) {
            // the seeded step failed and left the state as it was: run it
            // again without the seed and count the failure against the seed
            ++retries;
            ws_fail(c);
            retry_plain = true;
            continue;
        }
        retry_plain = false;
        if (status != 0u) {
            if (stats && completed < nsteps) stats[completed].status = (int32_t)status;
            return (int)status;
        }
        if (pci && last.iterations < 0) return LPS_ERR_HIP;
        if (pci && warm_ready(c)) ws_record(c, seeded, last.iterations, last.converged != 0);
    }
    if (pci) {
        // one check per call: every step must have completed on the device
        HIPC(hipMemcpyAsync(c->h_grid, c->grid, sizeof(DevGrid), hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        if (c->h_grid->status != 0u) return (int)c->h_grid->status;
        if (c->h_grid->steps_ok - base != nsteps) return LPS_ERR_HIP;
    }
    if (c->prof) prof_collect(c);
    return LPS_OK;
}

int lps_get_energy(lps_ctx* c, double* kinetic, double* potential, uint32_t nscenes) {
    if (!c || nscenes != c->h_sinfo.size() || (!kinetic && !potential)) return LPS_ERR_INVALID;
    if (c->n_real == 0) return LPS_ERR_STATE;
    HIPC(hipStreamSynchronize(c->stream));
    // the current state's per-block records were written by the last
    // integration or particle upload
    std::vector<double> keep(c->h_energy, c->h_energy + 2 * (size_t)nscenes);
    lps_dev::launch_scene_energy(c->stream, scenes(c), c->bbox_part, c->h_energy);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < nscenes; k++) {
        if (kinetic) kinetic[k] = c->h_energy[2 * k];
        if (potential) potential[k] = c->h_energy[2 * k + 1];
    }
    // the fallback reads the energy entering the last step from the same array
    std::copy(keep.begin(), keep.end(), c->h_energy);
    return LPS_OK;
}

bool valid_corr_threads(uint32_t t) { return t == 64u || t == 128u || t == (uint32_t)LPS_BLOCK; }
bool valid_lanes(uint32_t l) { return l == 0u || l == 1u || l == 4u; }

void set_corr_threads(lps_ctx* c, uint32_t t) {
    if (c->corr_threads != t) c->epoch++;  // captured prefixes hold the old launch shape
    c->corr_threads = t;
}

void set_lanes(lps_ctx* c, uint32_t l) {
    if (c->lanes != l) c->epoch++;
    c->lanes = l;
}

int lps_set_tuning(lps_ctx* c, const lps_tuning* t) {
    if (!c || !t || (t->use_graphs != 0 && t->use_graphs != 1) || t->lookahead < 1 || t->lookahead > 4 ||
        !valid_corr_threads(t->correction_threads) || !valid_lanes(t->lanes_per_particle))
        return LPS_ERR_INVALID;
    c->graphs_on = t->use_graphs != 0;
    c->lookahead = t->lookahead;
    set_corr_threads(c, t->correction_threads);
    set_lanes(c, t->lanes_per_particle);
    return LPS_OK;
}

int lps_get_tuning(lps_ctx* c, lps_tuning* t) {
    if (!c || !t) return LPS_ERR_INVALID;
    t->use_graphs = c->graphs_on ? 1 : 0;
    t->lookahead = c->lookahead;
    t->ms_per_step = 0.f;
    t->correction_threads = c->corr_threads;
    t->lanes_per_particle = quad_lanes(c) ? 4u : 1u;
    return LPS_OK;
}

int lps_autotune(lps_ctx* c, float dt, uint32_t steps, lps_tuning* chosen) {
    if (!c || !(dt > 0.f) || !std::isfinite(dt) || steps == 0 || steps > 1000) return LPS_ERR_INVALID;
    if (c->n_real == 0) return LPS_ERR_STATE;
    if (c->prof) return LPS_ERR_STATE;
    const hipStream_t s = c->stream;
    const uint32_t n = c->n, cur = (uint32_t)c->cur;
    const size_t nparts = (size_t)LPS_BBOX_PART * lps_dev::bbox_parts(n);
    // snapshot of everything the next step reads
    float4 *sp = nullptr, *sv = nullptr, *sh = nullptr;
    uint8_t* sf = nullptr;
    uint32_t *si = nullptr, *sb = nullptr;
    float* spp = nullptr;
    int st = LPS_OK;
    if ((st = dalloc(&sp, n)) || (st = dalloc(&sv, n)) || (st = dalloc(&sh, n)) || (st = dalloc(&sf, n)) ||
        (st = dalloc(&si, n)) || (st = dalloc(&spp, n)) || (st = dalloc(&sb, nparts))) {
        dfree(sp); dfree(sv); dfree(sh); dfree(sf); dfree(si); dfree(spp); dfree(sb);
        return st;
    }
    auto copy_all = [&](bool save) -> hipError_t {
        const int b = c->cur;
        hipError_t e = hipSuccess;
        auto cp = [&](void* dst, void* src, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
        };
        if (save) {
            cp(sp, c->pos[b], sizeof(float4) * n); cp(sv, c->vel[b], sizeof(float4) * n);
            cp(sh, c->vh[b], sizeof(float4) * n); cp(sf, c->first[b], n);
            cp(si, c->id[b], sizeof(uint32_t) * n); cp(spp, c->pprev[b], sizeof(float) * n);
            cp(sb, c->bbox_part, sizeof(uint32_t) * nparts);
        } else {
            cp(c->pos[b], sp, sizeof(float4) * n); cp(c->vel[b], sv, sizeof(float4) * n);
            cp(c->vh[b], sh, sizeof(float4) * n); cp(c->first[b], sf, n);
            cp(c->id[b], si, sizeof(uint32_t) * n); cp(c->pprev[b], spp, sizeof(float) * n);
            cp(c->bbox_part, sb, sizeof(uint32_t) * nparts);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e;
    };
    const bool g0 = c->graphs_on;
    const uint32_t l0 = c->lookahead, t0 = c->corr_threads, q0 = c->lanes;
    const uint32_t coll0 = c->coll_stride;
    c->coll_stride = 0;  // no training rows from trial steps
    const WarmFallback wf0 = c->wf;
    const uint64_t steps0 = c->steps;
    if (copy_all(true) != hipSuccess) st = LPS_ERR_HIP;
    // graphs x lookahead x correction-kernel workgroup size (the particle
    // block of 256 slots run as 1, 2 or 4 workgroups) with one lane per
    // particle; below 2048 particle blocks also graphs x lookahead with a
    // quad per particle (where the workgroup size does not apply); with
    // several scenes and at least 512 blocks also the automatic lanes (one
    // lane, and quads in the corrections that run fewer than 512 blocks)
    struct Cand { bool g; uint32_t l, t, q; float ms; };
    std::vector<uint32_t> lane_opts = {1u};
    if (n / LPS_BLOCK < 2048u) lane_opts.push_back(4u);
    if (c->h_sinfo.size() > 1 && n / LPS_BLOCK >= 512u) lane_opts.push_back(0u);
    std::vector<Cand> cands;
    for (int g = 0; g < 2; g++)
        for (uint32_t l = 1; l <= 3; l++)
            for (uint32_t q : lane_opts) {
                if (q == 4u) {
                    cands.push_back({g == 1, l, t0, 4u, 0.f});
                    continue;
                }
                for (uint32_t t : {(uint32_t)LPS_BLOCK, 128u, 64u}) cands.push_back({g == 1, l, t, q, 0.f});
            }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = LPS_ERR_HIP;
    for (Cand& cd : cands) {
        if (st) break;
        c->graphs_on = cd.g;
        c->lookahead = cd.l;
        set_corr_threads(c, cd.t);
        set_lanes(c, cd.q);
        // one untimed step (graph capture), then the timed ones, each from the snapshot
        for (int pass = 0; pass < 2 && !st; pass++) {
            c->cur = (int)cur;
            c->wf = wf0;
            c->e_prev_valid = false;
            if (copy_all(false) != hipSuccess) { st = LPS_ERR_HIP; break; }
            if (pass == 1) (void)hipEventRecord(e0, s);
            st = lps_step(c, dt, pass == 0 ? 1u : steps, nullptr);
            if (pass == 1) {
                (void)hipEventRecord(e1, s);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                cd.ms = ms / (float)steps;
            }
        }
        if (cd.g && !c->graphs_on) cd.ms = INFINITY;  // graphs unavailable on this device
    }
    // restore the state and the scheduling bookkeeping; keep the fastest
    c->cur = (int)cur;
    c->wf = wf0;
    c->steps = steps0;
    c->e_prev_valid = false;
    c->coll_stride = coll0;
    if (copy_all(false) != hipSuccess && !st) st = LPS_ERR_HIP;
    c->have_step = false;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(sp); dfree(sv); dfree(sh); dfree(sf); dfree(si); dfree(spp); dfree(sb);
    if (st) {
        c->graphs_on = g0;
        c->lookahead = l0;
        set_corr_threads(c, t0);
        set_lanes(c, q0);
        return st;
    }
    const Cand* best = &cands[0];
    for (const Cand& cd : cands)
        if (cd.ms < best->ms) best = &cd;
    c->graphs_on = best->g;
    c->lookahead = best->l;
    set_corr_threads(c, best->t);
    set_lanes(c, best->q);
    if (chosen) {
        chosen->use_graphs = best->g ? 1 : 0;
        chosen->lookahead = best->l;
        chosen->ms_per_step = best->ms;
        chosen->correction_threads = best->t;
        chosen->lanes_per_particle = best->q;
    }
    return LPS_OK;
}

int lps_get_grid(lps_ctx* c, lps_grid_info* info, uint32_t* keys, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end, uint32_t cap) {
    if (!c || !c->have_step) return LPS_ERR_STATE;
    const uint32_t S = (uint32_t)c->h_sinfo.size();
    HIPC(hipMemcpyAsync(c->h_grid, c->grid, sizeof(DevGrid), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipMemcpyAsync(c->h_sgrid, c->d_sgrid, sizeof(SceneGrid) * S, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    const DevGrid& G = *c->h_grid;
    const SceneGrid& S0 = c->h_sgrid[0];
    // single scene: the grid of the reference; ensembles: scene 0's grid, the
    // global (composite) keys and the global cell table
    const uint32_t cells = S == 1 ? S0.cells : G.total_cells;
    if (info) {
        for (int a = 0; a < 3; a++) {
            info->bbox_min[a] = S0.bbox_min[a];
            info->bbox_max[a] = S0.bbox_max[a];
            info->grid_offset[a] = S0.offset[a];
            info->grid_size[a] = S0.gsize[a];
        }
        info->grid_spacing = G.spacing;
        info->cell_count = cells;
    }
    const uint32_t ns = c->n;
    if (keys || perm) {
        std::vector<uint32_t> f(ns);
        if (keys) {
            HIPC(hipMemcpy(f.data(), c->skeys, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
            compact_slots(c, f, keys, 1);
        }
        if (perm) {
            HIPC(hipMemcpy(f.data(), c->perm, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
            compact_slots(c, f, perm, 1);
        }
    }
    if (cell_start || cell_end) {
        const uint32_t m = std::min(cap, cells);
        std::vector<uint32_t> lb((size_t)m + 1);
        HIPC(hipMemcpy(lb.data(), c->cell_lb, sizeof(uint32_t) * ((size_t)m + 1), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < m; k++) {
            if (cell_start) cell_start[k] = lb[k];
            if (cell_end) cell_end[k] = lb[k + 1];
        }
    }
    return LPS_OK;
}

int lps_get_fields(lps_ctx* c, float* rho, float* pressure, float* force, float* rho_pred) {
    if (!c || !c->have_step) return LPS_ERR_STATE;
    HIPC(hipStreamSynchronize(c->stream));
    const uint32_t ns = c->n, n = c->n_real;
    std::vector<float4> a(ns), b, ra(n), rb;
    std::vector<float> f(ns);
    if (rho || force) {
        HIPC(hipMemcpy(a.data(), c->posrho, sizeof(float4) * ns, hipMemcpyDeviceToHost));
        compact_slots(c, a, ra.data(), 1);
    }
    if (rho)
        for (uint32_t i = 0; i < n; i++) rho[i] = ra[i].w;
    if (pressure) {
        // the applied pressure, copied by the integration
        HIPC(hipMemcpy(f.data(), c->pprev[c->cur], sizeof(float) * ns, hipMemcpyDeviceToHost));
        compact_slots(c, f, pressure, 1);
    }
    if (force) {
        HIPC(hipMemcpy(a.data(), c->force, sizeof(float4) * ns, hipMemcpyDeviceToHost));
        compact_slots(c, a, ra.data(), 1);
        if (c->last_fp_valid) {
            rb.resize(n);
            if (int st = gather_applied(c, c->fp, b)) return st;
            compact_slots(c, b, rb.data(), 1);
        }
        for (uint32_t i = 0; i < n; i++) {
            float fx = ra[i].x, fy = ra[i].y, fz = ra[i].z;
            if (c->last_fp_valid) {
                fx += rb[i].x;
                fy += rb[i].y;
                fz += rb[i].z;
            }
            force[3 * i] = fx;
            force[3 * i + 1] = fy;
            force[3 * i + 2] = fz;
        }
    }
    if (rho_pred) {
        if (c->last_fp_valid) {
            if (int st = gather_applied(c, c->rho_pred, f)) return st;
        } else {
            HIPC(hipMemcpy(f.data(), c->rho_pred, sizeof(float) * ns, hipMemcpyDeviceToHost));
        }
        compact_slots(c, f, rho_pred, 1);
    }
    return LPS_OK;
}

int lps_mlp_set_weights(lps_ctx* c, const float* params, const float* feature_scale, float target_scale) {
    if (!c || !params || !feature_scale || !(target_scale > 0.f) || !std::isfinite(target_scale))
        return LPS_ERR_INVALID;
    for (uint32_t k = 0; k < LPS_MLP_PARAMS; k++)
        if (!std::isfinite(params[k])) return LPS_ERR_INVALID;
    for (uint32_t k = 0; k < 2 * LPS_MLP_IN; k++)
        if (!std::isfinite(feature_scale[k])) return LPS_ERR_INVALID;
    if (!c->d_mlp) HIPC(hipMalloc((void**)&c->d_mlp, sizeof(float) * (LPS_MLP_PARAMS + 2 * LPS_MLP_IN)));
    if (!c->d_mlp_packed) HIPC(hipMalloc((void**)&c->d_mlp_packed, lps_dev::mlp_packed_bytes()));
    c->mlp_w.assign(params, params + LPS_MLP_PARAMS);
    c->mlp_fscale.assign(feature_scale, feature_scale + 2 * LPS_MLP_IN);
    c->mlp_tscale = target_scale;
    HIPC(hipMemcpyAsync(c->d_mlp, c->mlp_w.data(), sizeof(float) * LPS_MLP_PARAMS, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->d_mlp + LPS_MLP_PARAMS, c->mlp_fscale.data(), sizeof(float) * 2 * LPS_MLP_IN,
                        hipMemcpyHostToDevice, c->stream));
    lps_dev::launch_mlp_pack(c->stream, mlp_args(c), c->d_mlp_packed);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    c->mlp_loaded = true;
    c->epoch++;
    c->wf.restart();
    return LPS_OK;
}

int lps_mlp_get_weights(lps_ctx* c, float* params, float* feature_scale, float* target_scale) {
    if (!c) return LPS_ERR_INVALID;
    if (!c->mlp_loaded) return LPS_ERR_STATE;
    if (params) std::memcpy(params, c->mlp_w.data(), sizeof(float) * LPS_MLP_PARAMS);
    if (feature_scale) std::memcpy(feature_scale, c->mlp_fscale.data(), sizeof(float) * 2 * LPS_MLP_IN);
    if (target_scale) *target_scale = c->mlp_tscale;
    return LPS_OK;
}

int lps_mlp_capture(lps_ctx* c, int32_t on) {
    if (!c) return LPS_ERR_INVALID;
    c->capture = on != 0;
    return LPS_OK;
}

int lps_dump_pairs(lps_ctx* c, float* features, float* target, float* seed, uint32_t n) {
    if (!c || (!features && !target && !seed)) return LPS_ERR_INVALID;
    if (n != c->n_real) return LPS_ERR_INVALID;
    if (!c->have_step || !c->feat_valid || (seed && !c->seed_valid)) return LPS_ERR_STATE;
    HIPC(hipStreamSynchronize(c->stream));
    const uint32_t ns = c->n;
    std::vector<float> f((size_t)ns * LPS_MLP_IN);
    if (features) {
        // feature-major on the device: one row of ns slots per feature
        std::vector<float> col(ns);
        for (uint32_t k = 0; k < LPS_MLP_IN; k++) {
            HIPC(hipMemcpy(col.data(), c->feat + (size_t)k * c->cap_n, sizeof(float) * ns, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < ns; i++) f[(size_t)i * LPS_MLP_IN + k] = col[i];
        }
        compact_slots(c, f, features, LPS_MLP_IN);
    }
    if (target) {
        // the applied pressure, copied by the integration
        HIPC(hipMemcpy(f.data(), c->pprev[c->cur], sizeof(float) * (size_t)ns, hipMemcpyDeviceToHost));
        compact_slots(c, f, target, 1);
    }
    if (seed) {
        HIPC(hipMemcpy(f.data(), c->seed, sizeof(float) * (size_t)ns, hipMemcpyDeviceToHost));
        compact_slots(c, f, seed, 1);
    }
    return LPS_OK;
}

int lps_mlp_collect(lps_ctx* c, uint64_t capacity, uint32_t stride) {
    if (!c || capacity > (1ull << 32)) return LPS_ERR_INVALID;
    if (capacity && capacity == c->coll_cap) {
        c->coll_stride = stride;  // rows kept; 0 pauses
        return LPS_OK;
    }
    HIPC(hipStreamSynchronize(c->stream));
    dfree(c->coll_x);
    dfree(c->coll_t);
    c->coll_cap = c->coll_rows = 0;
    c->coll_offset = 0;
    if (!capacity) return LPS_OK;
    int st;
    if ((st = dalloc(&c->coll_x, (size_t)capacity * LPS_MLP_IN))) return st;
    if ((st = dalloc(&c->coll_t, (size_t)capacity))) return st;
    c->coll_cap = capacity;
    c->coll_stride = stride;
    return LPS_OK;
}

int lps_mlp_collected(lps_ctx* c, uint64_t* rows) {
    if (!c || !rows) return LPS_ERR_INVALID;
    *rows = c->coll_rows;
    return LPS_OK;
}

}  // extern "C"

namespace {

int ensure_train_buffers(lps_ctx* c) {
    int st;
    if (!c->d_mlp && (st = dalloc(&c->d_mlp, LPS_MLP_PARAMS + 2 * LPS_MLP_IN))) return st;
    if (!c->d_mlp_packed && (st = dalloc(&c->d_mlp_packed, lps_dev::mlp_packed_bytes()))) return st;
    if (!c->d_grad && (st = dalloc(&c->d_grad, LPS_MLP_PARAMS))) return st;
    if (!c->d_m1 && (st = dalloc(&c->d_m1, LPS_MLP_PARAMS))) return st;
    if (!c->d_m2 && (st = dalloc(&c->d_m2, LPS_MLP_PARAMS))) return st;
    if (!c->d_loss && (st = dalloc(&c->d_loss, 1))) return st;
    if (!c->d_stats && (st = dalloc(&c->d_stats, 4 * LPS_MLP_IN))) return st;
    return LPS_OK;
}

// He-normal weights (ReLU layers), zero biases.
void init_weights(std::vector<float>& w, uint64_t seed) {
    std::mt19937_64 rng(seed);
    w.assign(LPS_MLP_PARAMS, 0.f);
    const int H = LPS_MLP_HIDDEN, I = LPS_MLP_IN;
    std::normal_distribution<float> n1(0.f, std::sqrt(2.f / I)), n2(0.f, std::sqrt(2.f / H)),
        n3(0.f, std::sqrt(1.f / H));
    float* p = w.data();
    for (int k = 0; k < H * I; k++) *p++ = n1(rng);
    p += H;
    for (int k = 0; k < H * H; k++) *p++ = n2(rng);
    p += H;
    for (int k = 0; k < H; k++) *p++ = n3(rng);
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

extern "C" {

int lps_mlp_train(lps_ctx* c, const lps_mlp_train_cfg* cfg, float* epoch_loss) {
    if (!c || !cfg || cfg->epochs == 0 || cfg->batch == 0 || !(cfg->lr > 0.f) || !(cfg->lr_final >= 0.f) ||
        !(cfg->over_weight > 0.f) || !std::isfinite(cfg->over_weight) ||
        !(cfg->beta1 >= 0.f && cfg->beta1 < 1.f) || !(cfg->beta2 >= 0.f && cfg->beta2 < 1.f) || !(cfg->eps > 0.f))
        return LPS_ERR_INVALID;
    if (c->coll_rows == 0) return LPS_ERR_STATE;
    int st = ensure_train_buffers(c);
    if (st) return st;
    const hipStream_t s = c->stream;
    const uint64_t N = c->coll_rows;
    if (cfg->refit_scale || !c->mlp_loaded) {
        // feature normalisation (mean, 1/std; constant features get weight 0)
        // and target scale (root mean square of the targets)
        HIPC(hipMemsetAsync(c->d_stats, 0, sizeof(double) * 4 * LPS_MLP_IN, s));
        lps_dev::launch_feature_stats(s, c->coll_x, N, LPS_MLP_IN, c->d_stats);
        lps_dev::launch_feature_stats(s, c->coll_t, N, 1, c->d_stats + 2 * LPS_MLP_IN);
        double h[4 * LPS_MLP_IN];
        HIPC(hipMemcpyAsync(h, c->d_stats, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        c->mlp_fscale.assign(2 * LPS_MLP_IN, 0.f);
        for (int f = 0; f < LPS_MLP_IN; f++) {
            const double mean = h[f] / (double)N;
            const double var = std::max(h[LPS_MLP_IN + f] / (double)N - mean * mean, 0.0);
            c->mlp_fscale[f] = (float)mean;
            c->mlp_fscale[LPS_MLP_IN + f] = var > 1e-12 * std::max(1.0, mean * mean) ? (float)(1.0 / std::sqrt(var)) : 0.f;
        }
        const double rms = std::sqrt(h[2 * LPS_MLP_IN + 1] / (double)N);
        c->mlp_tscale = rms > 0.0 && std::isfinite(rms) ? (float)rms : 1.f;
    }
    if (cfg->reinit || !c->mlp_loaded) init_weights(c->mlp_w, cfg->seed);
    HIPC(hipMemcpyAsync(c->d_mlp, c->mlp_w.data(), sizeof(float) * LPS_MLP_PARAMS, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(c->d_mlp + LPS_MLP_PARAMS, c->mlp_fscale.data(), sizeof(float) * 2 * LPS_MLP_IN,
                        hipMemcpyHostToDevice, s));
    HIPC(hipMemsetAsync(c->d_grad, 0, sizeof(float) * LPS_MLP_PARAMS, s));
    HIPC(hipMemsetAsync(c->d_m1, 0, sizeof(float) * LPS_MLP_PARAMS, s));
    HIPC(hipMemsetAsync(c->d_m2, 0, sizeof(float) * LPS_MLP_PARAMS, s));
    const lps_dev::MlpArgs m = mlp_args(c);
    const uint64_t per_epoch = (N + cfg->batch - 1u) / cfg->batch;
    const double total = (double)per_epoch * cfg->epochs;
    std::mt19937_64 rng(cfg->seed ^ 0x9E3779B97F4A7C15ull);
    uint32_t t = 0;
    for (uint32_t e = 0; e < cfg->epochs; e++) {
        // a fresh affine permutation of the rows each epoch
        uint64_t mul = N > 1 ? 1u + rng() % (N - 1u) : 1u;
        while (gcd64(mul, N) != 1u) mul = mul + 1u < N ? mul + 1u : 1u;
        const uint64_t add = rng() % N;
        HIPC(hipMemsetAsync(c->d_loss, 0, sizeof(float), s));
        for (uint64_t b = 0; b < per_epoch; b++) {
            const uint64_t start = b * cfg->batch;
            const uint32_t rows = (uint32_t)std::min<uint64_t>(cfg->batch, N - start);
            const lps_dev::RowMap rm{nullptr, start, mul, add, N};
            lps_dev::launch_mlp_grad(s, m, c->coll_x, c->coll_t, rm, rows, cfg->over_weight, c->d_grad, c->d_loss);
            const double u = (double)t / std::max(total - 1.0, 1.0);
            const float lr = (float)(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(3.14159265358979323846 * u)));
            ++t;
            lps_dev::launch_adam(s, c->d_mlp, c->d_grad, c->d_m1, c->d_m2, LPS_MLP_PARAMS, lr, cfg->beta1, cfg->beta2,
                                 cfg->eps, t);
        }
        float loss = 0.f;
        HIPC(hipMemcpyAsync(&loss, c->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        if (epoch_loss) epoch_loss[e] = loss / (float)per_epoch;
    }
    lps_dev::launch_mlp_pack(s, m, c->d_mlp_packed);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(c->mlp_w.data(), c->d_mlp, sizeof(float) * LPS_MLP_PARAMS, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (float w : c->mlp_w)
        if (!std::isfinite(w)) {
            c->mlp_loaded = false;
            return LPS_ERR_NONFINITE;
        }
    c->mlp_loaded = true;
    c->epoch++;
    c->wf.restart();
    return LPS_OK;
}

int lps_mlp_forward(lps_ctx* c, const float* x, uint32_t n, float* y) {
    if (!c || !x || !y || n == 0) return LPS_ERR_INVALID;
    if (!c->mlp_loaded) return LPS_ERR_STATE;
    float *dx = nullptr, *dy = nullptr;
    int st = dalloc(&dx, (size_t)n * LPS_MLP_IN);
    if (!st) st = dalloc(&dy, n);
    if (!st && hipMemcpy(dx, x, sizeof(float) * LPS_MLP_IN * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        st = LPS_ERR_HIP;
    if (!st) {
        lps_dev::launch_mlp_forward(c->stream, c->d_mlp_packed, dx, n, dy);
        if (hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = LPS_ERR_HIP;
    }
    dfree(dx);
    dfree(dy);
    return st;
}

int lps_mlp_grad(lps_ctx* c, const float* x, const float* t, uint32_t n, float over_weight, float* grad,
                 float* loss) {
    if (!c || !x || !t || !grad || n == 0 || !(over_weight > 0.f) || !std::isfinite(over_weight))
        return LPS_ERR_INVALID;
    if (!c->mlp_loaded) return LPS_ERR_STATE;
    int st = ensure_train_buffers(c);
    if (st) return st;
    float *dx = nullptr, *dt = nullptr;
    if ((st = dalloc(&dx, (size_t)n * LPS_MLP_IN)) || (st = dalloc(&dt, n))) {
        dfree(dx);
        return st;
    }
    const hipStream_t s = c->stream;
    if (hipMemcpy(dx, x, sizeof(float) * LPS_MLP_IN * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, t, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(c->d_grad, 0, sizeof(float) * LPS_MLP_PARAMS, s) != hipSuccess ||
        hipMemsetAsync(c->d_loss, 0, sizeof(float), s) != hipSuccess)
        st = LPS_ERR_HIP;
    if (!st) {
        const lps_dev::RowMap rm{nullptr, 0u, 1u, 0u, n};
        lps_dev::launch_mlp_grad(s, mlp_args(c), dx, dt, rm, n, over_weight, c->d_grad, c->d_loss);
        float l = 0.f;
        if (hipMemcpyAsync(grad, c->d_grad, sizeof(float) * LPS_MLP_PARAMS, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(&l, c->d_loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemsetAsync(c->d_grad, 0, sizeof(float) * LPS_MLP_PARAMS, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = LPS_ERR_HIP;
        if (loss) *loss = l;
    }
    dfree(dx);
    dfree(dt);
    return st;
}

int lps_mlp_grad_noalloc(lps_ctx* c, const float* x, const float* t, uint32_t n, float over_weight, float* grad,
                         float* loss) {
    if (!c || !x || !t || !grad || n == 0 || !(over_weight > 0.f) || !std::isfinite(over_weight))
        return LPS_ERR_INVALID;
    if (n > LPS_NOALLOC_MAX_ROWS) return LPS_ERR_INVALID;
    if (!c->mlp_loaded) return LPS_ERR_STATE;
    unsigned char* base = scratch_ptr(c->device);
    if (!base) return LPS_ERR_HIP;
    float* dx = reinterpret_cast<float*>(base);
    float* dt = reinterpret_cast<float*>(base + LPS_NOALLOC_X_BYTES);
    float* dg = reinterpret_cast<float*>(base + LPS_NOALLOC_X_BYTES + LPS_NOALLOC_T_BYTES);
    float* dl = reinterpret_cast<float*>(base + LPS_NOALLOC_X_BYTES + LPS_NOALLOC_T_BYTES + LPS_NOALLOC_G_BYTES);
    const hipStream_t s = c->stream;
    // the accumulators keep the previous call's sums: clear them (a fill
    // kernel, so the call stays free of memset nodes)
    lps_dev::launch_fill_zero(s, dg, LPS_MLP_PARAMS);
    lps_dev::launch_fill_zero(s, dl, 1u);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(dx, x, sizeof(float) * LPS_MLP_IN * (size_t)n, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(dt, t, sizeof(float) * n, hipMemcpyHostToDevice, s));
    const lps_dev::RowMap rm{nullptr, 0u, 1u, 0u, n};
    lps_dev::launch_mlp_grad(s, mlp_args(c), dx, dt, rm, n, over_weight, dg, dl);
    HIPC(hipGetLastError());
    float l = 0.f;
    HIPC(hipMemcpyAsync(grad, dg, sizeof(float) * LPS_MLP_PARAMS, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return LPS_OK;
}

int lps_profile_enable(lps_ctx* c, int32_t on) {
    if (!c) return LPS_ERR_INVALID;
    if (!on) prof_collect(c);
    c->prof = on != 0;
    return LPS_OK;
}

int lps_profile_query(lps_ctx* c, const char* kernel, uint32_t* launches, double* total_ms) {
    if (!c || !kernel) return LPS_ERR_INVALID;
    prof_collect(c);
    auto it = c->prof_acc.find(kernel);
    if (launches) *launches = it == c->prof_acc.end() ? 0u : it->second.first;
    if (total_ms) *total_ms = it == c->prof_acc.end() ? 0.0 : it->second.second;
    return LPS_OK;
}

int lps_profile_reset(lps_ctx* c) {
    if (!c) return LPS_ERR_INVALID;
    prof_collect(c);
    c->prof_acc.clear();
    return LPS_OK;
}

int lps_get_stream(lps_ctx* c, void** stream) {
    if (!c || !stream) return LPS_ERR_INVALID;
    *stream = (void*)c->stream;
    return LPS_OK;
}

}  // extern "C"
