// coef_apply_mfma.hip -- coefficient apply on gfx950 MFMA (fp16 / bf16 inputs, fp32 accumulation), with an
// optional fused max-log demapper in the epilogue (Y_eq then never leaves registers).
//
// Complex product as one real GEMM per subcarrier ("4-real" decomposition in real-embedding form):
//   D[16 x 16] = A'[16 x 2A] * B'[2A x 16]
//   B'[(a,0)][n] = Re Y[n][a],  B'[(a,1)][n] = Im Y[n][a]           (= the raw interleaved Y row of symbol n)
//   A'[l][(a,*)]   = ( Re C[l][a], -Im C[l][a])  -> D[l][n]   = Re Y_eq[l][n]   (rows 0..7)
//   A'[8+l][(a,*)] = ( Im C[l][a],  Re C[l][a])  -> D[8+l][n] = Im Y_eq[l][n]   (rows 8..15)
// so M = 16 rows are filled by the 8 layers (re and im), N = 16 columns hold up to 16 symbols (11 used),
// K = 2A runs in steps of 32 through v_mfma_f32_16x16x32_{f16,bf16}. One wave owns one subcarrier.
//
// Operand loading: the reduction index K is assigned so that each lane reads contiguous memory. Lane
// (g = lane/16, r = lane%16) covers antennas [g*4*KS, (g+1)*4*KS) -- 4 antennas (32 bytes, two 128-bit loads)
// per k-step -- of C row (r % 8) for the A operand and of Y row r for the B operand. Any assignment works as
// long as A and B agree, because a k-step sums over the lane groups and elements alike.
//
// Range: every C row (layer) and every Y row (symbol) is scaled by its own power of two so that its largest
// component lies in [0.5, 1) before rounding to 16 bits; the exponents are put back on the fp32 accumulators
// (exact). The split variants add the rounding residuals: A*B ~ Ahi*Bhi + (Ahi*Blo + Alo*Bhi).
#include "pe_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

template <bool BF>
struct Op16;
template <>
struct Op16<false> {
    typedef f16x8 vec;
    static __device__ __forceinline__ vec cvt(const float (&v)[8]) {
        vec r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (_Float16)v[i];
        return r;
    }
    static __device__ __forceinline__ vec cvt_lo(const float (&v)[8], const vec &hi) {
        vec r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (_Float16)(v[i] - (float)hi[i]);
        return r;
    }
    static __device__ __forceinline__ f32x4 mfma(const vec &a, const vec &b, const f32x4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Op16<true> {
    typedef bf16x8 vec;
    static __device__ __forceinline__ vec cvt(const float (&v)[8]) {
        vec r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (__bf16)v[i];
        return r;
    }
    static __device__ __forceinline__ vec cvt_lo(const float (&v)[8], const vec &hi) {
        vec r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (__bf16)(v[i] - (float)hi[i]);
        return r;
    }
    static __device__ __forceinline__ f32x4 mfma(const vec &a, const vec &b, const f32x4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ float absmax4(float4 v) {
    return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
}

// exponent e such that m * 2^-e lies in [0.5, 1); 0 for m == 0
__device__ __forceinline__ int scale_exp(float m) {
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}

// PREC: 1 f16, 2 bf16, 3 f16 split, 4 bf16 split.  NPB: 0 -> store Y_eq, 1..4 -> fused demap.
template <int KS, int PREC, int NPB>
__global__ __launch_bounds__(256) void coef_apply_mfma_kernel(
    const float2 *__restrict__ C, const float2 *__restrict__ Y, float2 *__restrict__ O,
    const float *__restrict__ noise_inv, float *__restrict__ llr, int nSC, int L, int S, int A) {
    constexpr bool BF = (PREC == 2 || PREC == 4);
    constexpr bool SPLIT = (PREC == 3 || PREC == 4);
    constexpr int AG = 4 * KS;  // antennas per lane group
    typedef Op16<BF> Op;

    const int lane = threadIdx.x & 63;
    const int sc = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (sc >= nSC) return;  // uniform per wave
    const int g = lane >> 4;
    const int r = lane & 15;
    const int la = r & 7;
    const bool part_im = r >= 8;
    const bool arow = la < L;
    const bool bcol = r < S;

    const float4 *cp = reinterpret_cast<const float4 *>(C + ((size_t)sc * L + (arow ? la : 0)) * A);
    const float4 *yp = reinterpret_cast<const float4 *>(Y + ((size_t)sc * S + (bcol ? r : 0)) * A);

    float4 cr[KS][2], yr[KS][2];
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        const int a0 = g * AG + 4 * t;
        const bool ok = a0 < A;
        cr[t][0] = (arow && ok) ? cp[a0 / 2] : z4;
        cr[t][1] = (arow && ok) ? cp[a0 / 2 + 1] : z4;
        yr[t][0] = (bcol && ok) ? yp[a0 / 2] : z4;
        yr[t][1] = (bcol && ok) ? yp[a0 / 2 + 1] : z4;
    }

    // per-row / per-column power-of-two range scaling
    float mc = 0.f, my = 0.f;
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        mc = fmaxf(mc, fmaxf(absmax4(cr[t][0]), absmax4(cr[t][1])));
        my = fmaxf(my, fmaxf(absmax4(yr[t][0]), absmax4(yr[t][1])));
    }
    mc = wave_max_xor(mc, 16, 64);  // reduce over the 4 lane groups (same row / column)
    my = wave_max_xor(my, 16, 64);
    const int ec = scale_exp(mc);
    const int ey = scale_exp(my);
    const float sc_c = ldexpf(1.f, -ec);
    const float sc_y = ldexpf(1.f, -ey);

    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    f32x4 acc2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        const float4 c0 = cr[t][0], c1 = cr[t][1];
        float av[8], bv[8];
        if (!part_im) {
            av[0] = c0.x; av[1] = -c0.y; av[2] = c0.z; av[3] = -c0.w;
            av[4] = c1.x; av[5] = -c1.y; av[6] = c1.z; av[7] = -c1.w;
        } else {
            av[0] = c0.y; av[1] = c0.x; av[2] = c0.w; av[3] = c0.z;
            av[4] = c1.y; av[5] = c1.x; av[6] = c1.w; av[7] = c1.z;
        }
        const float4 y0 = yr[t][0], y1 = yr[t][1];
        bv[0] = y0.x; bv[1] = y0.y; bv[2] = y0.z; bv[3] = y0.w;
        bv[4] = y1.x; bv[5] = y1.y; bv[6] = y1.z; bv[7] = y1.w;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            av[i] *= sc_c;
            bv[i] *= sc_y;
        }
        const typename Op::vec ah = Op::cvt(av);
        const typename Op::vec bh = Op::cvt(bv);
        acc = Op::mfma(ah, bh, acc);
        if constexpr (SPLIT) {
            const typename Op::vec al = Op::cvt_lo(av, ah);
            const typename Op::vec bl = Op::cvt_lo(bv, bh);
            acc2 = Op::mfma(ah, bl, acc2);
            acc2 = Op::mfma(al, bh, acc2);
        }
    }
    if constexpr (SPLIT) acc += acc2;

    // acc[j] = D[4g + j][r]; undo the row (layer) and column (symbol) scaling
    float d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ecR = __shfl(ec, (lane & 48) | (4 * g + j), 64);
        d[j] = ldexpf(acc[j], ecR + ey);
    }
    // pair real and imaginary parts: lanes g<2 hold Re of layers 4g..4g+3, lanes g>=2 hold Im of layers
    // 4(g-2)..4(g-2)+3; after one xor-32 exchange each lane holds two complete symbols
    const bool lo = g < 2;
    const float x0 = lo ? d[2] : d[0];
    const float x1 = lo ? d[3] : d[1];
    const float y0 = __shfl_xor(x0, 32, 64);
    const float y1 = __shfl_xor(x1, 32, 64);
    float2 s0, s1;
    int l0;
    if (lo) {
        l0 = 4 * g;
        s0 = make_float2(d[0], y0);
        s1 = make_float2(d[1], y1);
    } else {
        l0 = 4 * (g - 2) + 2;
        s0 = make_float2(y0, d[2]);
        s1 = make_float2(y1, d[3]);
    }
    const int n = r;
    if (n >= S) return;
    if constexpr (NPB == 0) {
        if (l0 < L) O[((size_t)sc * L + l0) * S + n] = s0;
        if (l0 + 1 < L) O[((size_t)sc * L + l0 + 1) * S + n] = s1;
    } else {
        if (l0 < L) {
            float v[2 * NPB];
            qam_llr<NPB>(s0.x, s0.y, noise_inv[(size_t)sc * L + l0], v);
            store_llr<NPB>(llr + (((size_t)sc * L + l0) * S + n) * (2 * NPB), v);
        }
        if (l0 + 1 < L) {
            float v[2 * NPB];
            qam_llr<NPB>(s1.x, s1.y, noise_inv[(size_t)sc * L + l0 + 1], v);
            store_llr<NPB>(llr + (((size_t)sc * L + l0 + 1) * S + n) * (2 * NPB), v);
        }
    }
}

template <int KS, int PREC, int NPB>
hipError_t launch3(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, hipStream_t st) {
    const CoefTile &t = coef_tile_for(L, S, A, PREC);
    const int waves = (t.mfma_waves == 1 || t.mfma_waves == 2) ? t.mfma_waves : 4;
    const int grid = (nSC + waves - 1) / waves;
    hipLaunchKernelGGL((coef_apply_mfma_kernel<KS, PREC, NPB>), dim3(grid), dim3(64 * waves), 0, st, C, Y, O,
                       ni, llr, nSC, L, S, A);
    return hipGetLastError();
}

template <int KS, int PREC>
hipError_t launch2(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, int npb, hipStream_t st) {
    switch (npb) {
    case 0: return launch3<KS, PREC, 0>(C, Y, O, ni, llr, nSC, L, S, A, st);
    case 1: return launch3<KS, PREC, 1>(C, Y, O, ni, llr, nSC, L, S, A, st);
    case 2: return launch3<KS, PREC, 2>(C, Y, O, ni, llr, nSC, L, S, A, st);
    case 3: return launch3<KS, PREC, 3>(C, Y, O, ni, llr, nSC, L, S, A, st);
    case 4: return launch3<KS, PREC, 4>(C, Y, O, ni, llr, nSC, L, S, A, st);
    default: return hipErrorInvalidValue;
    }
}

template <int KS>
hipError_t launch1(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, int npb, int prec, hipStream_t st) {
    switch (prec) {
    case 1: return launch2<KS, 1>(C, Y, O, ni, llr, nSC, L, S, A, npb, st);
    case 2: return launch2<KS, 2>(C, Y, O, ni, llr, nSC, L, S, A, npb, st);
    case 3: return launch2<KS, 3>(C, Y, O, ni, llr, nSC, L, S, A, npb, st);
    case 4: return launch2<KS, 4>(C, Y, O, ni, llr, nSC, L, S, A, npb, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_coef_apply_mfma(const float2 *C, const float2 *Y, float2 *O, const float *noise_inv,
                                  float *llr, int nSC, int L, int S, int A, int nPamBits, int prec,
                                  hipStream_t st) {
    const int npb = llr ? nPamBits : 0;
    if (A <= 16) return launch1<1>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, st);
    if (A <= 32) return launch1<2>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, st);
    if (A <= 64) return launch1<4>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, st);
    if (A <= 128) return launch1<8>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, st);
    return hipErrorInvalidValue;
}

}  // namespace pe
