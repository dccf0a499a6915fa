// lps_kernels.h -- host launchers of the gfx950 kernels (lps_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "lps_common.h"

// Loop rule of the PCISPH corrections, evaluated on the device.
struct LoopRule {
    uint32_t min_iter, max_iter, stall_limit;
    float tol;
};

namespace lps_dev {

// The scenes of a context: constants, device grid/loop state, and the scene
// of every particle block (blocks of LPS_BLOCK slots never straddle scenes).
struct Scenes {
    const SceneInfo* info;
    SceneGrid* grid;
    const uint8_t* block_scene;
    uint32_t count;
    uint32_t nblocks;        // particle blocks (slots / LPS_BLOCK)
    uint32_t* act_list;      // [2][nblocks] active-block lists of the correction kernels
};

// Candidate masks recorded by one neighbour pass for the next pass over the
// same positions and cells (k_density -> k_forces / k_pressure_sph,
// k_pci_density -> k_pci_pressure of the same correction): per slot, the
// accept mask of the first 32 candidates of each of the 9 rows of its
// 27-cell stencil (bit t: candidate row_start + t), row-major (row r of slot
// i at mask[r * stride + i], so a wave's lanes touch consecutive words), and
// a flag that the slot was recorded.
struct NbMasks {
    uint32_t* mask;    // [LPS_MC_ROWS][stride]
    uint32_t* ok;      // [stride] 1: recorded by the last recording pass
    uint32_t stride;   // slots (LPS_MC_ROWS * stride < 2^32)
    uint2* cq;         // [stride + LPS_STENCIL_PAD] positions in 16-bit fixed point for the recording
                       // passes' filter: the sorted positions (written by the reorder, read by the
                       // density pass), then each prediction (read by the predicted-density pass)
};

struct ReorderArgs {
    const uint32_t *keysA, *valsA, *keysB, *valsB;
    uint32_t n;
    const DevGrid* g;
    const float4 *pos_in, *vel_in, *vh_in;
    const uint8_t* first_in;
    const uint32_t* id_in;
    const float* pprev_in;
    float4 *pos_out, *vel_out, *vh_out;
    uint8_t* first_out;
    uint32_t* id_out;
    float* pprev_out;
    uint32_t* run_end;
    uint32_t* tile_max;  // per-tile maxima of the run ends (cell-table scan tiles)
    uint32_t* skeys;
    uint32_t* perm;
    Scenes sc;            // slot -> scene (for the compact copy's grid)
    uint2* cq;            // compact copy of pos_out for the density pass's filter
};

// Buffers shared by the physics kernels of one step (sorted order).
struct PhysArgs {
    Scenes sc;
    float4* pos;          // x (w unused)
    float4* vel;          // v
    float4* vh;           // v_1/2
    uint8_t* first;       // first-step flag
    float4* posrho;       // (x, rho) written by the density pass
    float4* force;        // non-pressure (+ SPH pressure) force
    DevGrid* g;
    const uint32_t* cell_lb;
    const float* tables;  // 5 x LPS_TABLE_SIZE
    const lps_box* boxes;
    const lps_sphere* spheres;
    const lps_plane* planes;
    NbMasks mk;
    SimConst c;
    uint32_t quad;        // 1: density, forces and SPH pressure give each particle a quad of lanes
};

// per-block partial records (bounding box and energies) of a state
void launch_bbox(hipStream_t s, const float4* pos, const float4* vel, float mass, uint32_t n, const Scenes& sc,
                 uint32_t* part);
uint32_t bbox_parts(uint32_t n);
// per-scene kinetic and potential energy from the partial records to
// out[2 * scene], out[2 * scene + 1] (coherent host memory)
void launch_scene_energy(hipStream_t s, const Scenes& sc, const uint32_t* part, double* out);
// grid of the step; also writes each scene's energy of the entering state
// to energy[2 * scene .. + 1] when energy is given
void launch_grid_setup(hipStream_t s, DevGrid* g, const Scenes& sc, const uint32_t* part, float h, uint32_t cell_cap,
                       double* energy);
// keys, and the first radix pass's histogram into hist (that pass skips its own)
void launch_keys(hipStream_t s, const float4* pos, uint32_t n, const DevGrid* g, const Scenes& sc, uint32_t* keys,
                 uint32_t* vals, uint32_t* tile_max, uint32_t* hist);
uint32_t radix_tiles(uint32_t n);
void launch_radix_pass(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       uint32_t n, uint32_t shift, DevGrid* g, uint32_t* hist, uint32_t* digit_total);
void launch_reorder(hipStream_t s, const ReorderArgs& a);
uint32_t cell_blocks(uint32_t cell_cap);
void launch_cell_table(hipStream_t s, uint32_t* run_end, const DevGrid* g, uint32_t* block_max, uint32_t* cell_lb,
                       uint32_t cell_cap);
void launch_density(hipStream_t s, const PhysArgs& a, float* press, float* q);
void launch_forces(hipStream_t s, const PhysArgs& a, float4* xpred, const float* pprev, float* feat);
void launch_pressure_sph(hipStream_t s, const PhysArgs& a, const float* q);
// rho_pred and press: three buffers of a.mk.stride slots each; correction k
// writes the scene's buffer buf_of[k % 64] (and reads the previous pressure
// from buf_of[(k - 1) % 64]); xq receives (prediction xyz, p / rho*^2), the
// pressure pass's single gather per candidate (a.n + LPS_STENCIL_PAD slots)
// threads: workgroup size of the correction kernels (64, 128 or LPS_BLOCK)
void launch_pci_density(hipStream_t s, const PhysArgs& a, const float4* xpred, float* rho_pred, float* press,
                        float4* xq, uint32_t k, float* feat, uint32_t threads);

// ---- learned warm start (lps_mlp.hip)
// Network parameters, fp32 master copy in this order: W1 [64][16], b1 [64],
// W2 [64][64], b2 [64], W3 [64], b3 [1]; features are normalised as
// (f - mean) * inv_std before the first layer, the output scaled by tscale.
struct MlpArgs {
    const float* params;
    const float* fscale;  // mean[16], inv_std[16]
    float tscale;
};
// Row of batch sample s: idx[s] when idx is given, else the affine
// permutation ((start + s) * mul + add) mod rows (mul coprime to rows).
struct RowMap {
    const uint32_t* idx;
    uint64_t start, mul, add, rows;
    __host__ __device__ uint32_t row(uint32_t s) const {
        return idx ? idx[s] : (uint32_t)(((start + s) * mul + add) % rows);
    }
};
// y_i = pscale * tscale * MLP(features_i) (0 when not finite) and seed_i =
// y_i; the first correction's pressure (buffer buf_of[1] of press, stride
// fs) becomes y_i, or with negative pressures disabled (clamp_nonneg)
// max(p1_i, min(y_i, kappa * pprev_i)), p1_i the pressure that correction
// computed; the w slot of xq_i becomes p_i / rho*_i^2.  No-op once the step
// has stopped or failed.
// feat: feature-major, feature f of slot i at f * fs + i
// skips the scenes whose loop stops at the first correction (rule, rest_density)
void launch_mlp_seed(hipStream_t s, const void* packed, float tscale, const float* feat, uint32_t fs, uint32_t n,
                     float* press, float4* xq, float* seed, const float* rho_pred, const float* pprev,
                     float kappa, float pscale, int clamp_nonneg, const DevGrid* g, const Scenes& sc,
                     const LoopRule& rule, float rest_density);
// BF16 inference weights (k_mlp_seed's LDS layout) from the FP32 master copy
size_t mlp_packed_bytes();
void launch_mlp_pack(hipStream_t s, const MlpArgs& m, void* packed);
// batch inference (raw output, before tscale) / gradients, for training and tests
void launch_mlp_forward(hipStream_t s, const void* packed, const float* x, uint32_t n, float* y);
// p[0..n) = 0 with a kernel (no memset node)
void launch_fill_zero(hipStream_t s, float* p, uint32_t n);
void launch_mlp_grad(hipStream_t s, const MlpArgs& m, const float* x, const float* t, const RowMap& rm, uint32_t n,
                     float over_weight, float* grad, float* loss);
void launch_adam(hipStream_t s, float* params, float* grad, float* m1, float* m2, uint32_t count, float lr,
                 float b1, float b2, float eps, uint32_t step);
void launch_feature_stats(hipStream_t s, const float* x, uint64_t n, uint32_t cols, double* acc);
void launch_collect(hipStream_t s, const float* feat, uint32_t fs, const float* press, float pscale, uint32_t n,
                    uint32_t stride, uint32_t offset, uint32_t count, float* X, float* T, uint64_t base);
// also applies the loop rule of correction k (block 0) and answers host_flag;
// xq is the predicted-density pass's (prediction, p / rho*^2)
void launch_pci_pressure(hipStream_t s, const PhysArgs& a, const float4* xq, const float* rho_pred, float4* fp,
                         float4* xpred_next, uint32_t k, const LoopRule& rule, uint32_t* host_flag, uint32_t threads);
// PCISPH (fp given): fp and press hold three buffers of a.mk.stride slots;
// each scene reads the one its loop applied (SceneGrid::final_buf).  SPH:
// press is the density pass's pressure.  pprev receives the applied pressure.
void launch_integrate(hipStream_t s, const PhysArgs& a, const float4* fp, float* pprev, const float* press,
                      uint32_t* bbox_part);
// positions of the real particles, scenes back to back, to dst[r * stride + 0..2]
void launch_export_positions(hipStream_t s, const float4* pos, uint32_t n, const Scenes& sc, float* dst,
                             uint32_t stride);

}  // namespace lps_dev
