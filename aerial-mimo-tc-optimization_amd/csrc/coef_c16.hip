// coef_c16.hip -- coefficients kept in HBM in the binary16 operand format of the MFMA product: C16[nSC][L][A]
// binary16 pairs {re, im} and one exponent per subcarrier (C = C16 * 2^e, rule in c16_exp). The fused apply
// reads 4 bytes per coefficient instead of 8 and copies them into LDS unconverted; the products, the
// accumulation and the demapper are those of the binary16 path on complex64 coefficients.
//   coef_to_c16_kernel: complex64 coefficients -> C16 + exponents (one wave per subcarrier).
//   coef_apply_c16_kernel: fused apply + demap (or Y_eq) from C16, one subcarrier per wave, WPB waves per
//   workgroup as tuned for the binary16 path of the same shape.
#include "apply_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mfma;

template <int KS, int WPB>
__global__ __launch_bounds__(64 * WPB) void coef_to_c16_kernel(const float2 *__restrict__ C,
                                                                uint32_t *__restrict__ C16, int *__restrict__ cexp,
                                                                int nSC, int L, int A) {
    const long sc = (long)blockIdx.x * WPB + (threadIdx.x >> 6);
    if (sc >= nSC) return;  // uniform per wave; no barrier follows
    const int lane = threadIdx.x & 63;
    const int nC = L * (A >> 1);  // 16-byte chunks (2 complex values) of the block
    const float4 *cb = reinterpret_cast<const float4 *>(C + (size_t)sc * L * A);
    float4 v[KS];
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int c = lane + 64 * j;
        v[j] = c < nC ? cb[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        m = max4abs(m, v[j]);
    }
    const int e = c16_exp(wave_max_xor(m, 1, 64));
    const float s = ldexpf(1.f, -e);
    uint2 *ob = reinterpret_cast<uint2 *>(C16 + (size_t)sc * L * A);
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int c = lane + 64 * j;
        if (c < nC) {
            float4 w = v[j];
            if (e != 0) { w.x *= s; w.y *= s; w.z *= s; w.w *= s; }
            ob[c] = make_uint2(pack2<false>(w.x, w.y), pack2<false>(w.z, w.w));
        }
    }
    if (lane == 0) cexp[sc] = e;
}

template <int KS, int NPB, int WPB>
__global__ __launch_bounds__(64 * WPB) void coef_apply_c16_kernel(
    const uint32_t *__restrict__ C16, const int *__restrict__ cexp, const float2 *__restrict__ Y,
    float2 *__restrict__ O, const float *__restrict__ noise_inv, float *__restrict__ llr, int nSC, int L, int S,
    int A, int f16) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_all[WPB * LdsLayout<KS>::bytes(1)];
    const int wave = threadIdx.x >> 6;
    unsigned char *lds = lds_all + wave * LdsLayout<KS>::bytes(1);
    const long sc = (long)blockIdx.x * WPB + wave;
    if (sc >= nSC) return;  // uniform per wave; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    if (A < 16 * KS) zero_operands<KS, 1>(lds, lane);
    const int nC = L * (A >> 2), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    const int ec = cexp[sc];
    RawSC<KS> ra;
    load_raw_c16<KS>(ra, C16, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
    process_sc<KS, 1, NPB, true>(ra, (size_t)sc, lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, 0, f16,
                                 ec);
}

template <int KS, int NPB>
hipError_t launch_apply(const uint32_t *C16, const int *cexp, const float2 *Y, float2 *O, const float *ni,
                        float *llr, int nSC, int L, int S, int A, int f16, hipStream_t st) {
    const int wpb = coef_tile_for(L, S, A, 1).mfma_waves;
    if (wpb == 4)
        hipLaunchKernelGGL((coef_apply_c16_kernel<KS, NPB, 4>), dim3((unsigned)((nSC + 3) / 4)), dim3(256), 0, st,
                           C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16);
    else if (wpb == 2)
        hipLaunchKernelGGL((coef_apply_c16_kernel<KS, NPB, 2>), dim3((unsigned)((nSC + 1) / 2)), dim3(128), 0, st,
                           C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16);
    else
        hipLaunchKernelGGL((coef_apply_c16_kernel<KS, NPB, 1>), dim3((unsigned)nSC), dim3(64), 0, st, C16, cexp, Y,
                           O, ni, llr, nSC, L, S, A, f16);
    return hipGetLastError();
}

template <int KS>
hipError_t launch_apply_ks(const uint32_t *C16, const int *cexp, const float2 *Y, float2 *O, const float *ni,
                           float *llr, int nSC, int L, int S, int A, int npb, int f16, hipStream_t st) {
    switch (npb) {
    case 0: return launch_apply<KS, 0>(C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 1: return launch_apply<KS, 1>(C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 2: return launch_apply<KS, 2>(C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 3: return launch_apply<KS, 3>(C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 4: return launch_apply<KS, 4>(C16, cexp, Y, O, ni, llr, nSC, L, S, A, f16, st);
    default: return hipErrorInvalidValue;
    }
}

template <int KS>
hipError_t launch_conv(const float2 *C, uint32_t *C16, int *cexp, int nSC, int L, int A, hipStream_t st) {
    hipLaunchKernelGGL((coef_to_c16_kernel<KS, 4>), dim3((unsigned)((nSC + 3) / 4)), dim3(256), 0, st, C, C16, cexp,
                       nSC, L, A);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_coef_to_c16(const float2 *C, uint32_t *C16, int *cexp, int nSC, int L, int A, hipStream_t st) {
    if (A <= 16) return launch_conv<1>(C, C16, cexp, nSC, L, A, st);
    if (A <= 32) return launch_conv<2>(C, C16, cexp, nSC, L, A, st);
    if (A <= 64) return launch_conv<4>(C, C16, cexp, nSC, L, A, st);
    if (A <= 128) return launch_conv<8>(C, C16, cexp, nSC, L, A, st);
    return hipErrorInvalidValue;
}

hipError_t launch_coef_apply_c16(const uint32_t *C16, const int *cexp, const float2 *Y, float2 *O,
                                 const float *noise_inv, void *llr, int nSC, int L, int S, int A, int nPamBits,
                                 int llr_f16, hipStream_t st) {
    const int npb = llr ? nPamBits : 0;
    float *l = reinterpret_cast<float *>(llr);
    if (A <= 16) return launch_apply_ks<1>(C16, cexp, Y, O, noise_inv, l, nSC, L, S, A, npb, llr_f16, st);
    if (A <= 32) return launch_apply_ks<2>(C16, cexp, Y, O, noise_inv, l, nSC, L, S, A, npb, llr_f16, st);
    if (A <= 64) return launch_apply_ks<4>(C16, cexp, Y, O, noise_inv, l, nSC, L, S, A, npb, llr_f16, st);
    if (A <= 128) return launch_apply_ks<8>(C16, cexp, Y, O, noise_inv, l, nSC, L, S, A, npb, llr_f16, st);
    return hipErrorInvalidValue;
}

}  // namespace pe
