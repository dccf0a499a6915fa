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
    uint32_t* skeys;
    uint32_t* perm;
};

// Buffers shared by the physics kernels of one step (sorted order).
struct PhysArgs {
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
    SimConst c;
};

void launch_bbox(hipStream_t s, const float4* pos, uint32_t n, uint32_t* part);
uint32_t bbox_parts(uint32_t n);
void launch_grid_setup(hipStream_t s, DevGrid* g, const uint32_t* part, uint32_t n, float h, uint32_t cell_cap);
void launch_keys(hipStream_t s, const float4* pos, uint32_t n, const DevGrid* g, uint32_t* keys, uint32_t* vals);
uint32_t radix_tiles(uint32_t n);
void launch_radix_pass(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       uint32_t n, uint32_t shift, DevGrid* g, uint32_t* hist, uint32_t* digit_total);
void launch_reorder(hipStream_t s, const ReorderArgs& a);
uint32_t cell_blocks(uint32_t cell_cap);
void launch_cell_table(hipStream_t s, uint32_t* run_end, const DevGrid* g, uint32_t* block_max, uint32_t* cell_lb,
                       uint32_t cell_cap);
void launch_density(hipStream_t s, const float4* pos, const DevGrid* g, const uint32_t* cell_lb, const SimConst& c,
                    const float* tables, float4* posrho, float* press, float* q);
void launch_forces(hipStream_t s, const PhysArgs& a, float4* xpred);
void launch_pressure_sph(hipStream_t s, const PhysArgs& a, const float* q);
void launch_pci_density(hipStream_t s, const PhysArgs& a, const float4* xpred, float* rho_pred, float* press,
                        float* q, uint32_t k, const LoopRule& rule, uint32_t* host_flag);
void launch_pci_pressure(hipStream_t s, const PhysArgs& a, const float4* xpred, const float* rho_pred,
                         const float* q, float4* fp, float4* xpred_next, uint32_t k);
void launch_integrate(hipStream_t s, const PhysArgs& a, const float4* fp, float* pprev, const float* press,
                      uint32_t* bbox_part);

}  // namespace lps_dev
