// Device check of the hardware float -> bfloat16 conversion (v_cvt_pk_bf16_f32,
// what a (__bf16) conversion compiles to on gfx950) against round-to-nearest-
// even on the bit pattern, the rounding of the warm-start network's inputs and
// activations (liquidsimulation_amd/csrc/lps_mlp.hip, oracle/mlp.py):
//   bf16_check            every one of the 2^32 float bit patterns; a NaN must
//                         give a NaN, every other value the rounded pattern
// Prints the number of patterns and of mismatches and the first mismatch, and
// exits 0 when none.  Each launch covers 2^26 patterns, one per thread.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                               \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                                   \
        }                                                                                   \
    } while (0)

constexpr uint32_t kBlock = 256;
constexpr uint32_t kChunk = 1u << 26;

__device__ __forceinline__ uint32_t rne_bits(uint32_t u) { return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16; }

__global__ __launch_bounds__(kBlock) void k_bf16(uint32_t u0, uint32_t* bad, uint32_t* first) {
    const uint32_t u = u0 + blockIdx.x * kBlock + threadIdx.x;
    const float f = __uint_as_float(u);
    const uint32_t hw = __builtin_bit_cast(unsigned short, (__bf16)f);
    bool ok;
    if ((u & 0x7FFFFFFFu) > 0x7F800000u)
        ok = (hw & 0x7FFFu) > 0x7F80u;  // NaN in, NaN out
    else
        ok = hw == rne_bits(u);
    uint32_t nbad = ok ? 0u : 1u;
    for (int o = 32; o > 0; o >>= 1) nbad += __shfl_xor(nbad, o, 64);
    __shared__ uint32_t s[kBlock / 64];
    if ((threadIdx.x & 63u) == 0u) s[threadIdx.x >> 6] = nbad;
    if (!ok) first[blockIdx.x] = u;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kBlock / 64; w++) t += s[w];
        bad[blockIdx.x] += t;
    }
}

int main() {
    const uint32_t nblk = kChunk / kBlock;
    uint32_t *dbad, *dfirst;
    CK(hipMalloc(&dbad, nblk * sizeof(uint32_t)));
    CK(hipMalloc(&dfirst, nblk * sizeof(uint32_t)));
    CK(hipMemset(dbad, 0, nblk * sizeof(uint32_t)));
    CK(hipMemset(dfirst, 0, nblk * sizeof(uint32_t)));
    for (uint64_t u0 = 0; u0 < (1ull << 32); u0 += kChunk) {
        k_bf16<<<nblk, kBlock>>>((uint32_t)u0, dbad, dfirst);
        CK(hipGetLastError());
    }
    CK(hipDeviceSynchronize());
    std::vector<uint32_t> bad(nblk), first(nblk);
    CK(hipMemcpy(bad.data(), dbad, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost));
    CK(hipMemcpy(first.data(), dfirst, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost));
    uint64_t nbad = 0;
    uint32_t f = 0;
    bool have = false;
    for (uint32_t b = 0; b < nblk; b++) {
        if (bad[b] && !have) {
            f = first[b];
            have = true;
        }
        nbad += bad[b];
    }
    std::printf("bf16: %llu patterns checked, %llu wrong\n", 1ull << 32, (unsigned long long)nbad);
    if (nbad) std::printf("a mismatch: pattern 0x%08x (block-local first)\n", f);
    return nbad != 0;
}
