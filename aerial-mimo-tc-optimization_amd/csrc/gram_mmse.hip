// gram_mmse.hip -- the single-cell MMSE coefficient kernel: one launch, the FP32 tile pass with the FP64 solves of
// the subcarriers it flags done by the same workgroup; the per-subcarrier arithmetic is in mmse_device.hpp.
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
                              threadIdx.x, blockDim.x >> 6, c16, cexp);
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

hipError_t launch_gram_mmse_c16(const float2 *H, const float *sigma2, uint32_t *C16, int *cexp, float *noise_inv,
                                uint8_t *flags, int nSC, int L, int A, bool unbiased, float thr, hipStream_t st) {
    return launch_mmse<true>(H, sigma2, nullptr, noise_inv, flags, nSC, L, A, unbiased, thr, C16, cexp, st);
}

}  // namespace pe
