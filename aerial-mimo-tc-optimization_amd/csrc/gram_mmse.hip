// gram_mmse.hip -- single-launch MMSE coefficient kernels (FP32 pass, FP64 pass); the per-subcarrier solve
// is in mmse_device.hpp.
#include "mmse_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mmse;

template <int L, int P>
__global__ __launch_bounds__(256) void gram_mmse_f32_kernel(const float2 *__restrict__ H,
                                                            const float *__restrict__ sigma2,
                                                            float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                            uint8_t *__restrict__ flags, int nSC, int A,
                                                            int unbiased, float thr) {
    const int lane = threadIdx.x & 63;
    const int p = lane & (P - 1);
    const long sc = (long)blockIdx.x * (blockDim.x / P) + threadIdx.x / P;
    if (sc >= nSC) return;  // whole group leaves together; shuffles stay inside groups
    const bool ok = mmse_one<float, L, P>(H + (size_t)sc * A * L, sigma2[sc], A, p, unbiased != 0, thr, true,
                                          C + (size_t)sc * L * A, noise_inv + (size_t)sc * L);
    if (p == 0) flags[sc] = ok ? 0 : kPendingFp64;
}

template <int L, int P>
__global__ __launch_bounds__(256) void gram_mmse_f64_kernel(const float2 *__restrict__ H,
                                                            const float *__restrict__ sigma2,
                                                            float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                            uint8_t *__restrict__ flags, int nSC, int A,
                                                            int unbiased) {
    const int lane = threadIdx.x & 63;
    const int p = lane & (P - 1);
    const long sc = (long)blockIdx.x * (blockDim.x / P) + threadIdx.x / P;
    if (sc >= nSC) return;
    if (!(flags[sc] & kPendingFp64)) return;
    float2 *c = C + (size_t)sc * L * A;
    float *ni = noise_inv + (size_t)sc * L;
    const bool ok = mmse_one<double, L, P>(H + (size_t)sc * A * L, (double)sigma2[sc], A, p, unbiased != 0,
                                           0.0, false, c, ni);
    if (!ok) {
        for (int a = p; a < A; a += P)
            for (int l = 0; l < L; ++l) c[(size_t)l * A + a] = make_float2(0.f, 0.f);
        for (int l = 0; l < L; ++l)
            if (p == l % P) ni[l] = 0.f;
    }
    // every lane of the group has read the flag before lane 0 rewrites it (one wave, lockstep)
    if (p == 0) flags[sc] = ok ? PE_SCFLAG_FP64 : (PE_SCFLAG_FP64 | PE_SCFLAG_SINGULAR);
}

template <int L, int P>
hipError_t launch_mmse_LP(const float2 *H, const float *sigma2, float2 *C, float *ni, uint8_t *flags, int nSC,
                          int A, bool unbiased, float thr, hipStream_t st) {
    const int per_block = 256 / P;
    const int grid = (nSC + per_block - 1) / per_block;
    hipLaunchKernelGGL((gram_mmse_f32_kernel<L, P>), dim3(grid), dim3(256), 0, st, H, sigma2, C, ni, flags, nSC,
                       A, unbiased ? 1 : 0, thr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((gram_mmse_f64_kernel<L, P>), dim3(grid), dim3(256), 0, st, H, sigma2, C, ni, flags, nSC,
                       A, unbiased ? 1 : 0);
    return hipGetLastError();
}

template <int L>
hipError_t launch_mmse_L(const float2 *H, const float *sigma2, float2 *C, float *ni, uint8_t *flags, int nSC,
                         int A, bool unbiased, float thr, hipStream_t st) {
    if (mmse_lanes_for(L, A, nSC) == 16)
        return launch_mmse_LP<L, 16>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, st);
    return launch_mmse_LP<L, 8>(H, sigma2, C, ni, flags, nSC, A, unbiased, thr, st);
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
