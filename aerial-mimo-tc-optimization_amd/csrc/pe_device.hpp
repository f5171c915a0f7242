// pe_device.hpp -- device helpers shared by the gfx950 kernels of the PUSCH equalization library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pe {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Complex multiply-accumulate acc + a*b (no conjugation), as the contracted FP32 sequence
//   re = fma(-a.im, b.im, fma(a.re, b.re, acc.re)),  im = fma(a.im, b.re, fma(a.re, b.im, acc.im))
// i.e. the fused form of cma() in aerial_coef_wmma.cu:60-64. Written with explicit fmaf so the rounding
// sequence does not depend on compiler contraction.
__device__ __forceinline__ float2 cma(float2 a, float2 b, float2 acc) {
    float re = __builtin_fmaf(a.x, b.x, acc.x);
    re = __builtin_fmaf(-a.y, b.y, re);
    float im = __builtin_fmaf(a.x, b.y, acc.y);
    im = __builtin_fmaf(a.y, b.x, im);
    return make_float2(re, im);
}

// Orders this wave's LDS accesses across the point of call (each wave works in its own LDS region):
// LDS instructions of a wave execute in order, so only compiler reordering has to be stopped -- unlike
// __syncthreads(), this does not wait for the wave's outstanding global loads (prefetches stay in flight).
__device__ __forceinline__ void lds_order_wave() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- max-log PAM LLR (computePamLlr) ---------------------------------------------------------------
// Distances of the PAM2/4/8/16 axes at offsets 0, 1, 3, 6, as the decimal literals of
// aerial_llr_bench_v2.cu:47-54 rounded to float (they equal the float roundings of 1/sqrt(2),
// {2,1}/sqrt(10), {4,2,1}/sqrt(42), {8,4,2,1}/sqrt(170)).
template <int NPB> struct PamLut;
template <> struct PamLut<1> {
    static constexpr float D[1] = {0.707106781186548f};
};
template <> struct PamLut<2> {
    static constexpr float D[2] = {0.632455532033676f, 0.316227766016838f};
};
template <> struct PamLut<3> {
    static constexpr float D[3] = {0.617213399848368f, 0.308606699924184f, 0.154303349962092f};
};
template <> struct PamLut<4> {
    static constexpr float D[4] = {0.613571991077897f, 0.306785995538948f, 0.153392997769474f,
                                   0.076696498884737f};
};

// One PAM axis: folding chain sb_{i+1} = D_i - |sb_i|, md_i = (dist + |sb_i|)^2, m1 = (|sb_last| - dist)^2,
// llr_i = +-(md_i - m1) * noise (negative when sb_i < 0), clamped to +-65504. Same operation sequence as
// llr_v1_baseline (aerial_llr_bench_v2.cu:134-155), so FP32 results are bit-identical to it.
// Every operation is rounded on its own (no fused multiply-add contraction, e.g. of md_i - m1 with
// md_i = m*m), as the source lines of the reference read; the host restatement does the same.
template <int NPB>
__device__ __forceinline__ void pam_llr_axis(float x, float noise, float (&out)[NPB]) {
#pragma clang fp contract(off)
    const float dist = PamLut<NPB>::D[NPB - 1];
    float sb = x;
    unsigned sign = 0;
    float md[NPB];
#pragma unroll
    for (int i = 0; i < NPB; ++i) {
        const float a = fabsf(sb);
        float m = dist + a;
        md[i] = m * m;
        if (sb < 0.f) sign |= 1u << i;
        if (i < NPB - 1) sb = PamLut<NPB>::D[i] - a;
    }
    float m1 = fabsf(sb) - dist;
    m1 = m1 * m1;
#pragma unroll
    for (int i = 0; i < NPB; ++i) {
        float v = (md[i] - m1) * noise;
        if (sign & (1u << i)) v = -v;
        out[i] = fmaxf(-65504.f, fminf(65504.f, v));
    }
}

// Both axes of one complex symbol, interleaved [b0(I), b1(Q), b2(I), ...] (aerial_llr_bench_v2.cu:221-227).
template <int NPB>
__device__ __forceinline__ void qam_llr(float re, float im, float noise, float (&out)[2 * NPB]) {
    float li[NPB], lq[NPB];
    pam_llr_axis<NPB>(re, noise, li);
    pam_llr_axis<NPB>(im, noise, lq);
#pragma unroll
    for (int i = 0; i < NPB; ++i) {
        out[2 * i] = li[i];
        out[2 * i + 1] = lq[i];
    }
}

template <int NPB>
__device__ __forceinline__ void store_llr(float *dst, const float (&v)[2 * NPB]) {
    if constexpr (NPB == 1) {
        *reinterpret_cast<float2 *>(dst) = make_float2(v[0], v[1]);
    } else if constexpr (NPB == 2) {
        *reinterpret_cast<float4 *>(dst) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (NPB == 3) {
        reinterpret_cast<float2 *>(dst)[0] = make_float2(v[0], v[1]);
        reinterpret_cast<float2 *>(dst)[1] = make_float2(v[2], v[3]);
        reinterpret_cast<float2 *>(dst)[2] = make_float2(v[4], v[5]);
    } else {
        reinterpret_cast<float4 *>(dst)[0] = make_float4(v[0], v[1], v[2], v[3]);
        reinterpret_cast<float4 *>(dst)[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
}

// binary16 bits of a float, round to nearest even (the conversion of __float2half)
__device__ __forceinline__ uint32_t float_to_half_bits(float f) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
}

// two floats -> binary16 pair bits (a in the low half), round to nearest even
__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    return float_to_half_bits(a) | (float_to_half_bits(b) << 16);
}

// The 2*NPB LLRs of one symbol as binary16 (the reference's LLR wire format, aerial_llr_bench_v2.cu:221-227;
// the values are already clamped to +-65504 by pam_llr_axis).
template <int NPB>
__device__ __forceinline__ void store_llr_f16(uint16_t *dst, const float (&v)[2 * NPB]) {
    uint32_t w[NPB];
#pragma unroll
    for (int k = 0; k < NPB; ++k) w[k] = float_to_half_bits(v[2 * k]) | (float_to_half_bits(v[2 * k + 1]) << 16);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    if constexpr (NPB == 4) {
        *reinterpret_cast<uint4 *>(d) = make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (NPB == 2) {
        *reinterpret_cast<uint2 *>(d) = make_uint2(w[0], w[1]);
    } else {
#pragma unroll
        for (int k = 0; k < NPB; ++k) d[k] = w[k];
    }
}

// ---- binary16 coefficient blocks ------------------------------------------------------------------------
// A subcarrier's coefficient block C[L][A] stored as binary16 pairs {re, im} with one power-of-two exponent:
// C = C16 * 2^e. e = 0 when the block's largest |Re|, |Im| lies in [2^-9, 2^14) (binary16 holds it at full
// precision as is, the range the MFMA path converts without scaling); otherwise the frexp exponent of that
// maximum, so that the scaled block peaks in [0.5, 1). 0 for an all-zero block.
__device__ __forceinline__ int c16_exp(float m) {
    if (m >= 0x1p-9f && m < 0x1p14f) return 0;
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}

// ---- wave helpers (wave64) -------------------------------------------------------------------------
__device__ __forceinline__ float wave_max_xor(float v, int from, int to) {
    // butterfly over lane-index bits [from, to): every lane of the group ends with the same value
    for (int off = from; off < to; off <<= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace pe
