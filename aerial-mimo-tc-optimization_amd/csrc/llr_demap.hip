// llr_demap.hip -- standalone max-log LLR soft demappers (computePamLlr).
//
// pe_llr_demap_f16: the reference's operator (llr_v1_baseline, aerial_llr_bench_v2.cu:114-156): binary16
//   complex in, noiseInv per symbol, binary16 LLRs out [N][2q]. One lane per symbol; the PAM order is a
//   template parameter so the LUT entries are immediates (the approach of llr_v3_reg_template,
//   aerial_llr_v3.cu:119-229); the operation sequence is the reference's, so results are bit-identical.
// pe_llr_demap: FP32 on the equalizer output layout Y_eq[sc][l][s] with noise_inv[sc][l].
#include "pe_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

__device__ __forceinline__ float half_bits_to_float(uint32_t b) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xffffu));
}

template <int NPB>
__global__ __launch_bounds__(256) void llr_f16_kernel(const uint32_t *__restrict__ soft,
                                                      const float *__restrict__ noise_inv,
                                                      uint16_t *__restrict__ out, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t v = soft[i];
    float l[2 * NPB];
    qam_llr<NPB>(half_bits_to_float(v), half_bits_to_float(v >> 16), noise_inv[i], l);
    uint32_t w[NPB];
#pragma unroll
    for (int k = 0; k < NPB; ++k) w[k] = float_to_half_bits(l[2 * k]) | (float_to_half_bits(l[2 * k + 1]) << 16);
    uint32_t *dst = reinterpret_cast<uint32_t *>(out + (size_t)i * (2 * NPB));
    if constexpr (NPB == 4) {
        *reinterpret_cast<uint4 *>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (NPB == 2) {
        *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0], w[1]);
    } else {
#pragma unroll
        for (int k = 0; k < NPB; ++k) dst[k] = w[k];
    }
}

template <int NPB>
__global__ __launch_bounds__(256) void llr_f32_kernel(const float2 *__restrict__ Yeq,
                                                      const float *__restrict__ noise_inv,
                                                      float *__restrict__ llr, long total, int S) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const float2 y = Yeq[e];
    float v[2 * NPB];
    qam_llr<NPB>(y.x, y.y, noise_inv[e / S], v);
    store_llr<NPB>(llr + e * (2 * NPB), v);
}

}  // namespace

hipError_t launch_llr_f16(const uint32_t *soft, const float *noise_inv, uint16_t *out, int N, int nPamBits,
                          hipStream_t st) {
    const int grid = (N + 255) / 256;
    switch (nPamBits) {
    case 1: hipLaunchKernelGGL(llr_f16_kernel<1>, dim3(grid), dim3(256), 0, st, soft, noise_inv, out, N); break;
    case 2: hipLaunchKernelGGL(llr_f16_kernel<2>, dim3(grid), dim3(256), 0, st, soft, noise_inv, out, N); break;
    case 3: hipLaunchKernelGGL(llr_f16_kernel<3>, dim3(grid), dim3(256), 0, st, soft, noise_inv, out, N); break;
    case 4: hipLaunchKernelGGL(llr_f16_kernel<4>, dim3(grid), dim3(256), 0, st, soft, noise_inv, out, N); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_llr_f32(const float2 *Yeq, const float *noise_inv, float *llr, int nSC, int L, int S,
                          int nPamBits, hipStream_t st) {
    const long total = (long)nSC * L * S;
    const int grid = (int)((total + 255) / 256);
    switch (nPamBits) {
    case 1: hipLaunchKernelGGL(llr_f32_kernel<1>, dim3(grid), dim3(256), 0, st, Yeq, noise_inv, llr, total, S); break;
    case 2: hipLaunchKernelGGL(llr_f32_kernel<2>, dim3(grid), dim3(256), 0, st, Yeq, noise_inv, llr, total, S); break;
    case 3: hipLaunchKernelGGL(llr_f32_kernel<3>, dim3(grid), dim3(256), 0, st, Yeq, noise_inv, llr, total, S); break;
    case 4: hipLaunchKernelGGL(llr_f32_kernel<4>, dim3(grid), dim3(256), 0, st, Yeq, noise_inv, llr, total, S); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pe
