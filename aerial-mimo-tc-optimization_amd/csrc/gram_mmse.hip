// gram_mmse.hip -- single-launch MMSE coefficient kernels: the FP32 tile pass and the FP64 pass over flagged
// subcarriers; the per-subcarrier arithmetic is in mmse_device.hpp.
#include "mmse_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mmse;

// One workgroup of W waves per tile of T subcarriers (the last tile may be short); its LDS holds T slots.
// C16: coefficients written as binary16 pairs to c16 and exponents to cexp (C unused); NA antennas per lane
// (1: A <= 64, 2: A <= 128).
template <int L, int KR, bool C16, int NA = 2>
__global__ __launch_bounds__(512) void gram_mmse_f32_kernel(const float2 *__restrict__ H,
                                                            const float *__restrict__ sigma2,
                                                            float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                            uint8_t *__restrict__ flags, int nSC, int A, int P,
                                                            int T, int unbiased, float thr,
                                                            uint32_t *__restrict__ c16, int *__restrict__ cexp) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const long sc0 = (long)blockIdx.x * T;
    const int n = (int)min((long)T, (long)nSC - sc0);
    tile_fp32<L, KR, C16, NA>(H, sigma2, C, noise_inv, flags, sc0, n, A, P, unbiased != 0, thr, lds,
                              threadIdx.x, blockDim.x >> 6, nullptr, c16, cexp);
}

// One wave per 64 subcarriers; leaves after one flag load unless one of them was flagged.
template <int L, bool C16>
__global__ __launch_bounds__(64) void gram_mmse_f64_kernel(const float2 *__restrict__ H,
                                                           const float *__restrict__ sigma2,
                                                           float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                           uint8_t *__restrict__ flags, int nSC, int A, int P,
                                                           int unbiased, uint32_t *__restrict__ c16,
                                                           int *__restrict__ cexp) {
    const long sc0 = (long)blockIdx.x * 64;
    const int n = (int)min(64L, (long)nSC - sc0);
    tile_fp64<L, C16>(H, sigma2, C, noise_inv, flags, sc0, n, A, P, unbiased != 0, threadIdx.x, c16, cexp);
}

template <int L, bool C16>
hipError_t launch_mmse_L(const float2 *H, const float *sigma2, float2 *C, float *ni, uint8_t *flags, int nSC,
                         int A, bool unbiased, float thr, uint32_t *c16, int *cexp, hipStream_t st) {
    const int P = mmse_lanes_for(L, A);
    const MmseTile mt = mmse_tile_for(L, A, nSC);
    const int T = mt.tile;
    const int grid = (nSC + T - 1) / T;
    // rows kept in registers when asked for, every wave has one Gram round and a lane has at most 4 rows
    const bool keep = !C16 && mt.keep && T <= mt.waves * (64 / P) && A <= 4 * P;
    const dim3 g(grid), b(64 * mt.waves);
    const size_t lds = (size_t)tile_lds_bytes(L, T);
    if (keep)
        hipLaunchKernelGGL((gram_mmse_f32_kernel<L, 4, false>), g, b, lds, st, H, sigma2, C, ni, flags, nSC, A, P,
                           T, unbiased ? 1 : 0, thr, c16, cexp);
    else if (C16 && A <= 64)
        hipLaunchKernelGGL((gram_mmse_f32_kernel<L, 0, C16, 1>), g, b, lds, st, H, sigma2, C, ni, flags, nSC, A, P,
                           T, unbiased ? 1 : 0, thr, c16, cexp);
    else
        hipLaunchKernelGGL((gram_mmse_f32_kernel<L, 0, C16>), g, b, lds, st, H, sigma2, C, ni, flags, nSC, A, P,
                           T, unbiased ? 1 : 0, thr, c16, cexp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((gram_mmse_f64_kernel<L, C16>), dim3((nSC + 63) / 64), dim3(64), 0, st, H, sigma2, C, ni,
                       flags, nSC, A, P, unbiased ? 1 : 0, c16, cexp);
    return hipGetLastError();
}

template <bool C16>
hipError_t launch_mmse(const float2 *H, const float *sigma2, float2 *C, float *ni, uint8_t *flags, int nSC, int L,
                       int A, bool unbiased, float thr, uint32_t *c16, int *cexp, hipStream_t st) {
    switch (L) {
    case 1: return launch_mmse_L<1, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 2: return launch_mmse_L<2, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 3: return launch_mmse_L<3, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 4: return launch_mmse_L<4, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 5: return launch_mmse_L<5, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 6: return launch_mmse_L<6, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 7: return launch_mmse_L<7, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    case 8: return launch_mmse_L<8, C16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, c16, cexp, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_gram_mmse(const float2 *H, const float *sigma2, float2 *C, float *noise_inv, uint8_t *flags,
                            int nSC, int L, int A, bool unbiased, float thr, hipStream_t st) {
    return launch_mmse<false>(H, sigma2, C, noise_inv, flags, nSC, L, A, unbiased, thr, nullptr, nullptr, st);
}

hipError_t launch_gram_mmse_fp64_pass(const float2 *H, const float *sigma2, float2 *C, float *noise_inv,
                                      uint8_t *flags, int nSC, int L, int A, bool unbiased, hipStream_t st) {
    const int P = mmse_lanes_for(L, A);
    const dim3 g((unsigned)((nSC + 63) / 64)), b(64);
    const int u = unbiased ? 1 : 0;
    switch (L) {
#define PE_F64_L(LL)                                                                                              \
    case LL:                                                                                                      \
        hipLaunchKernelGGL((gram_mmse_f64_kernel<LL, false>), g, b, 0, st, H, sigma2, C, noise_inv, flags, nSC, A, P, \
                           u, nullptr, nullptr);                                                                  \
        break;
        PE_F64_L(1)
        PE_F64_L(2)
        PE_F64_L(3)
        PE_F64_L(4)
        PE_F64_L(5)
        PE_F64_L(6)
        PE_F64_L(7)
        PE_F64_L(8)
#undef PE_F64_L
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gram_mmse_c16(const float2 *H, const float *sigma2, uint32_t *C16, int *cexp, float *noise_inv,
                                uint8_t *flags, int nSC, int L, int A, bool unbiased, float thr, hipStream_t st) {
    return launch_mmse<true>(H, sigma2, nullptr, noise_inv, flags, nSC, L, A, unbiased, thr, C16, cexp, st);
}

}  // namespace pe
