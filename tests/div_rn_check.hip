// Device check of the shared-reciprocal division (liquidsimulation_amd/csrc/lps_div.h)
// against the IEEE quotient, with the hardware reciprocal estimate:
//   div_rn_check rcp                  every d in [2^-78, 2^40):
//                                     rcp_rn(d) == 1.0f / d
//   div_rn_check pairs A0 A1          every significand pair a = 1.m_a, d = 1.m_d
//                                     with m_a in [A0, A1):
//                                     div_rn(a, d, rcp_rn(d)) == a / d
// (significands as 23-bit integers; pairs 0 8388608 is the exhaustive run).
// Prints the number of cases and of mismatches, the first mismatch if any, and
// exits 0 when none.  Each launch covers a bounded slice; every thread's loop
// has a fixed trip count.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lps_div.h"

#define CK(x)                                                                       \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                           \
        }                                                                           \
    } while (0)

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t block_sum(uint32_t v) {
    __shared__ uint32_t s[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0u) s[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (uint32_t w = 0; w < kBlock / 64; w++) t += s[w];
    return t;
}

// Per-block mismatch counts accumulate in bad[blockIdx.x]; first[blockIdx.x]
// keeps the first failing case's bits (a in the high word, d in the low word).
__global__ __launch_bounds__(kBlock) void k_rcp(uint32_t u0, uint32_t count, uint32_t* bad, uint64_t* first) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    uint32_t nbad = 0;
    uint64_t f = 0;
    if (t < count) {
        const uint32_t u = u0 + t;
        const float d = __uint_as_float(u);
        if (lps_div::rcp_rn(d) != 1.f / d) {
            nbad = 1;
            f = u;
        }
    }
    const uint32_t n = block_sum(nbad);
    if (nbad) first[blockIdx.x] = f;
    if (threadIdx.x == 0) bad[blockIdx.x] += n;
}

__global__ __launch_bounds__(kBlock) void k_pairs(uint32_t a0, uint32_t na, uint32_t* bad, uint64_t* first) {
    const uint32_t md = blockIdx.x * kBlock + threadIdx.x;  // grid covers exactly 2^23 significands
    const float d = __uint_as_float(0x3f800000u | md);
    const float y = lps_div::rcp_rn(d);
    uint32_t nbad = 0;
    uint64_t f = 0;
    for (uint32_t k = 0; k < na; k++) {
        const uint32_t ua = 0x3f800000u | (a0 + k);
        const float a = __uint_as_float(ua);
        if (lps_div::div_rn(a, d, y) != a / d) {
            if (nbad == 0) f = ((uint64_t)ua << 32) | __float_as_uint(d);
            nbad++;
        }
    }
    const uint32_t n = block_sum(nbad);
    if (nbad) first[blockIdx.x] = f;
    if (threadIdx.x == 0) bad[blockIdx.x] += n;
}

static int report(const char* what, uint64_t cases, const std::vector<uint32_t>& bad, const std::vector<uint64_t>& first) {
    uint64_t nbad = 0, f = 0;
    for (size_t b = 0; b < bad.size(); b++) {
        if (bad[b] && !f) f = first[b];
        nbad += bad[b];
    }
    std::printf("%s: %llu cases checked, %llu wrong\n", what, (unsigned long long)cases, (unsigned long long)nbad);
    if (nbad) {
        float a, d;
        const uint32_t ua = (uint32_t)(f >> 32), ud = (uint32_t)f;
        std::memcpy(&a, &ua, 4);
        std::memcpy(&d, &ud, 4);
        std::printf("first mismatch: a=%a d=%a\n", a, d);
    }
    std::fflush(stdout);
    return nbad != 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "rcp") == 0) {
        float lo = 0x1p-78f, hi = 0x1p40f;
        uint32_t u0, u1;
        std::memcpy(&u0, &lo, 4);
        std::memcpy(&u1, &hi, 4);
        const uint32_t total = u1 - u0, chunk = 1u << 26, nblk = chunk / kBlock;
        uint32_t* dbad;
        uint64_t* dfirst;
        CK(hipMalloc(&dbad, nblk * sizeof(uint32_t)));
        CK(hipMalloc(&dfirst, nblk * sizeof(uint64_t)));
        CK(hipMemset(dbad, 0, nblk * sizeof(uint32_t)));
        CK(hipMemset(dfirst, 0, nblk * sizeof(uint64_t)));
        for (uint32_t s = 0; s < total; s += chunk) {
            const uint32_t cnt = total - s < chunk ? total - s : chunk;
            k_rcp<<<(cnt + kBlock - 1) / kBlock, kBlock>>>(u0 + s, cnt, dbad, dfirst);
            CK(hipGetLastError());
        }
        CK(hipDeviceSynchronize());
        std::vector<uint32_t> bad(nblk);
        std::vector<uint64_t> first(nblk);
        CK(hipMemcpy(bad.data(), dbad, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost));
        CK(hipMemcpy(first.data(), dfirst, nblk * sizeof(uint64_t), hipMemcpyDeviceToHost));
        return report("reciprocal", total, bad, first);
    }
    if (argc >= 4 && std::strcmp(argv[1], "pairs") == 0) {
        const uint32_t a0 = (uint32_t)std::strtoul(argv[2], nullptr, 0), a1 = (uint32_t)std::strtoul(argv[3], nullptr, 0);
        if (a1 > (1u << 23) || a0 >= a1) {
            std::fprintf(stderr, "pairs: need 0 <= A0 < A1 <= 8388608\n");
            return 2;
        }
        const uint32_t nblk = (1u << 23) / kBlock, step = 256;
        uint32_t* dbad;
        uint64_t* dfirst;
        CK(hipMalloc(&dbad, nblk * sizeof(uint32_t)));
        CK(hipMalloc(&dfirst, nblk * sizeof(uint64_t)));
        CK(hipMemset(dbad, 0, nblk * sizeof(uint32_t)));
        CK(hipMemset(dfirst, 0, nblk * sizeof(uint64_t)));
        uint32_t launches = 0;
        for (uint32_t a = a0; a < a1; a += step) {
            const uint32_t na = a1 - a < step ? a1 - a : step;
            k_pairs<<<nblk, kBlock>>>(a, na, dbad, dfirst);
            CK(hipGetLastError());
            if (++launches % 2048 == 0) {
                CK(hipDeviceSynchronize());
                std::printf("pairs: significands %u..%u of a done\n", a0, a + na);
                std::fflush(stdout);
            }
        }
        CK(hipDeviceSynchronize());
        std::vector<uint32_t> bad(nblk);
        std::vector<uint64_t> first(nblk);
        CK(hipMemcpy(bad.data(), dbad, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost));
        CK(hipMemcpy(first.data(), dfirst, nblk * sizeof(uint64_t), hipMemcpyDeviceToHost));
        return report("quotients", (uint64_t)(a1 - a0) << 23, bad, first);
    }
    std::fprintf(stderr, "usage: div_rn_check rcp | pairs A0 A1\n");
    return 2;
}
