// gram_mmse.hip -- single-launch MMSE coefficient kernels: the FP32 tile pass and the FP64 pass over flagged
// subcarriers; the per-subcarrier arithmetic is in mmse_device.hpp.
#include "mmse_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mmse;

// One workgroup of W waves per tile of T subcarriers (the last tile may be short); its LDS holds T slots.
template <int L, int KR>
__global__ __launch_bounds__(512) void gram_mmse_f32_kernel(const float2 *__restrict__ H,
                                                            const float *__restrict__ sigma2,
                                                            float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                            uint8_t *__restrict__ flags, int nSC, int A, int P,
                                                            int T, int unbiased, float thr) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const long sc0 = (long)blockIdx.x * T;
    const int n = (int)min((long)T, (long)nSC - sc0);
    tile_fp32<L, KR>(H, sigma2, C, noise_inv, flags, sc0, n, A, P, unbiased != 0, thr, lds, threadIdx.x,
                     blockDim.x >> 6, nullptr);
}

// One wave per 64 subcarriers; leaves after one flag load unless one of them was flagged.
template <int L>
__global__ __launch_bounds__(64) void gram_mmse_f64_kernel(const float2 *__restrict__ H,
                                                           const float *__restrict__ sigma2,
                                                           float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                           uint8_t *__restrict__ flags, int nSC, int A, int P,
                                                           int unbiased) {
    const long sc0 = (long)blockIdx.x * 64;
    const int n = (int)min(64L, (long)nSC - sc0);
    tile_fp64<L>(H, sigma2, C, noise_inv, flags, sc0, n, A, P, unbiased != 0, threadIdx.x);
}

template <int L>
hipError_t launch_mmse_L(const float2 *H, const float *sigma2, float2 *C, float *ni, uint8_t *flags, int nSC,
                         int A, bool unbiased, float thr, hipStream_t st) {
    const int P = mmse_lanes_for(L, A);
    const MmseTile mt = mmse_tile_for(L, A, nSC);
    const int T = mt.tile;
    const int grid = (nSC + T - 1) / T;
    // rows kept in registers when asked for, every wave has one Gram round and a lane has at most 4 rows
    const bool keep = mt.keep && T <= mt.waves * (64 / P) && A <= 4 * P;
    const dim3 g(grid), b(64 * mt.waves);
    const size_t lds = (size_t)tile_lds_bytes(L, T);
    if (keep)
        hipLaunchKernelGGL((gram_mmse_f32_kernel<L, 4>), g, b, lds, st, H, sigma2, C, ni, flags, nSC, A, P, T,
                           unbiased ? 1 : 0, thr);
    else
        hipLaunchKernelGGL((gram_mmse_f32_kernel<L, 0>), g, b, lds, st, H, sigma2, C, ni, flags, nSC, A, P, T,
                           unbiased ? 1 : 0, thr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((gram_mmse_f64_kernel<L>), dim3((nSC + 63) / 64), dim3(64), 0, st, H, sigma2, C, ni, flags,
                       nSC, A, P, unbiased ? 1 : 0);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gram_mmse(const float2 *H, const float *sigma2, float2 *C, float *noise_inv, uint8_t *flags,
                            int nSC, int L, int A, bool unbiased, float thr, hipStream_t st) {
    switch (L) {
    case 1: return launch_mmse_L<1>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 2: return launch_mmse_L<2>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 3: return launch_mmse_L<3>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 4: return launch_mmse_L<4>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 5: return launch_mmse_L<5>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 6: return launch_mmse_L<6>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 7: return launch_mmse_L<7>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    case 8: return launch_mmse_L<8>(H, sigma2, C, noise_inv, flags, nSC, A, unbiased, thr, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pe
