// apply_device.hpp -- per-subcarrier MFMA coefficient apply (+ fused demap) as device functions, shared by the
// single-cell kernel (coef_apply_mfma.hip) and the multi-cell tile kernels (multicell.hip).
//
// Complex product as one real GEMM per subcarrier ("4-real" decomposition in real-embedding form):
//   D[16 x 16] = A'[16 x 2A] * B'[2A x 16]
//   B'[(a,0)][n] = Re Y[n][a],  B'[(a,1)][n] = Im Y[n][a]
//   A'[l][(a,*)]   = ( Re C[l][a], -Im C[l][a])  -> D[l][n]   = Re Y_eq[l][n]   (rows 0..7)
//   A'[8+l][(a,*)] = ( Im C[l][a],  Re C[l][a])  -> D[8+l][n] = Im Y_eq[l][n]   (rows 8..15)
// so M = 16 rows are filled by the 8 layers (re and im), N = 16 columns hold up to 16 symbols (11 used),
// K = 2A runs in steps of 32 through v_mfma_f32_16x16x32_{f16,bf16}. One wave owns one subcarrier at a time.
//
// Data path per subcarrier (one wave):
//   1. the C block [L][A] and the Y block [S][A] are contiguous in HBM; lane i loads 16-byte chunks i, i+64, ...
//      so every instruction reads 1 KB contiguous and every byte is read by exactly one lane;
//   2. per-subcarrier power-of-two range scales of C and of Y (wave max-reduction, exact scaling), skipped when
//      binary16 already covers the data;
//   3. each lane rounds its chunks to 16 bits once and writes them into the LDS operand layout (the real-part
//      and imaginary-part rows of A' from the same C chunk, via a sign xor and a half swap on packed pairs);
//   4. MFMA: lane (g = lane/16, r = lane%16) reads its A' row r / B' column r, antennas [g*4KS + 4t, +4) for
//      k-step t, with one 128-bit LDS read per operand;
//   5. epilogue: undo the scale, pair real and imaginary parts with one xor-32 lane exchange, store Y_eq or run
//      the max-log demapper in registers and store LLRs.
// The split variants add the rounding residuals: A*B ~ Ahi*Bhi + (Ahi*Blo + Alo*Bhi).
#pragma once
#include "pe_device.hpp"
#include "pe_internal.hpp"

namespace pe {
namespace mfma {

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

// two floats -> packed 16-bit pair (element 0 in the low half), round to nearest even
template <bool BF>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    if constexpr (BF) {
        typedef __bf16 v2 __attribute__((ext_vector_type(2)));
        v2 v = {(__bf16)a, (__bf16)b};
        return __builtin_bit_cast(uint32_t, v);
    } else {
        typedef _Float16 v2 __attribute__((ext_vector_type(2)));
        v2 v = {(_Float16)a, (_Float16)b};
        return __builtin_bit_cast(uint32_t, v);
    }
}
template <bool BF>
__device__ __forceinline__ float2 unpack2(uint32_t p) {
    if constexpr (BF) {
        typedef __bf16 v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_bit_cast(v2, p);
        return make_float2((float)v[0], (float)v[1]);
    } else {
        typedef _Float16 v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_bit_cast(v2, p);
        return make_float2((float)v[0], (float)v[1]);
    }
}

__device__ __forceinline__ float max3abs(float m, float a, float b) {
    return fmaxf(fmaxf(m, fabsf(a)), fabsf(b));
}

// exponent e such that m * 2^-e lies in [0.5, 1); 0 for m == 0
__device__ __forceinline__ int scale_exp(float m) {
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}

// Raw complex64 data of one subcarrier, loaded cooperatively: the C block [L][A] and the Y block [S][A] of a
// subcarrier are contiguous, so lane i loads 16-byte chunks i, i+64, ... of each (RC / RY rounds at most).
template <int KS>
struct RawSC {
    static constexpr int RC = KS;      // ceil(8 * 16KS / 128): C rounds for L <= 8, A <= 16KS
    static constexpr int RY = 2 * KS;  // Y rounds for S <= 16
    float4 c[RC];
    float4 y[RY];
};

// 16-byte load of raw operand data. Each byte is read once per launch, so the load is marked nontemporal
// (measured on MI355X: +7% HBM throughput for the fused kernel over ordinary loads).
typedef float f32x4_nt __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream(const float4 *p) {
    const f32x4_nt v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_nt *>(p));
    return make_float4(v[0], v[1], v[2], v[3]);
}

template <int KS>
__device__ __forceinline__ void load_raw(RawSC<KS> &R, const float2 *__restrict__ C, const float2 *__restrict__ Y,
                                         size_t sc, int L, int S, int A, int lane, int nC, int nY, int rc,
                                         int ry) {
    const float4 *cb = reinterpret_cast<const float4 *>(C + sc * L * A);
    const float4 *yb = reinterpret_cast<const float4 *>(Y + sc * S * A);
    // chunks past the block end re-read chunk 0 of the same block: the per-subcarrier maximum is unchanged
    // and they are not written to LDS
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RC; ++j)
        if (j < rc) {
            const int c = lane + 64 * j;
            R.c[j] = ld_stream(cb + (c < nC ? c : 0));
        }
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RY; ++j)
        if (j < ry) {
            const int c = lane + 64 * j;
            R.y[j] = ld_stream(yb + (c < nY ? c : 0));
        }
}

// The Y block alone (coefficients supplied another way).
template <int KS>
__device__ __forceinline__ void load_raw_y(RawSC<KS> &R, const float2 *__restrict__ Y, size_t sc, int S, int A,
                                           int lane, int nY, int ry) {
    const float4 *yb = reinterpret_cast<const float4 *>(Y + sc * S * A);
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RY; ++j)
        if (j < ry) {
            const int c = lane + 64 * j;
            R.y[j] = ld_stream(yb + (c < nY ? c : 0));
        }
}

// Coefficients of one subcarrier as columns in registers (computed in the same kernel instead of loaded):
// lane q holds C[l][a] for antennas a = q + 64 u (u < 2, A <= 128) and every layer l < 8; entries for
// l >= L or a >= A are zero.
struct CCols {
    float r[2][8];
    float i[2][8];
};

// Binary16 coefficient blocks (C16 [nSC][L][A] pairs, c16_exp scaling): the C chunks hold 4 pairs each,
// L*A/4 of them per subcarrier.
template <int KS>
__device__ __forceinline__ void load_raw_c16(RawSC<KS> &R, const uint32_t *__restrict__ C16,
                                             const float2 *__restrict__ Y, size_t sc, int L, int S, int A,
                                             int lane, int nC, int nY, int rc, int ry) {
    const float4 *cb = reinterpret_cast<const float4 *>(C16 + sc * L * A);
    const float4 *yb = reinterpret_cast<const float4 *>(Y + sc * S * A);
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RC; ++j)
        if (j < rc) {
            const int c = lane + 64 * j;
            R.c[j] = ld_stream(cb + (c < nC ? c : 0));
        }
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RY; ++j)
        if (j < ry) {
            const int c = lane + 64 * j;
            R.y[j] = ld_stream(yb + (c < nY ? c : 0));
        }
}

__device__ __forceinline__ float max4abs(float m, float4 v) {
    return max3abs(max3abs(m, v.x, v.y), v.z, v.w);
}

// LDS operand layout of one wave (rows padded by 16 bytes):
//   C region [8 rows][2*16KS halfs]:  row l holds packed pairs (Re C[l][a], Im C[l][a]), a = 0..A-1
//   Y region [16 rows][2*16KS halfs]: row s holds packed pairs (Re Y[s][a], Im Y[s][a])
// The A' operand rows are built from the C rows at read time: row l (real part of Y_eq) negates the high
// half of every pair -> (Re, -Im); row 8 + l (imaginary part) swaps the halves -> (Im, Re). Rows past L and
// S are never written; they only feed outputs that are never stored.
template <int KS>
struct LdsLayout {
    static constexpr int RB = 64 * KS + 16;     // bytes per row
    static constexpr int CREG = 8 * RB;         // C region
    static constexpr int YREG = 16 * RB;        // Y region
    static constexpr int HI = CREG + YREG;      // one precision level (hi); the split variants add a lo copy
    static constexpr int bytes(int prec) { return (prec == 3 || prec == 4 ? 2 : 1) * HI; }
};


// Antennas past A are never written by process_sc; zero every operand row once before a subcarrier range
// whose A is below the 16*KS the layout holds.
template <int KS, int PREC>
__device__ __forceinline__ void zero_operands(unsigned char *lds, int lane) {
    for (int i = lane; i < LdsLayout<KS>::bytes(PREC) / 16; i += 64)
        reinterpret_cast<uint4 *>(lds)[i] = make_uint4(0u, 0u, 0u, 0u);
    lds_order_wave();
}

// Max-log LLRs of the two symbols (layers l0, l0 + 1, symbol n) a lane holds after the re/im exchange.
template <int NPB>
// f16: llr points to binary16 storage of the same layout (the LLR wire format) instead of FP32.
__device__ __forceinline__ void demap_store(float2 s0, float2 s1, size_t sc, int l0, int n, int L, int S,
                                            const float *__restrict__ noise_inv, float *__restrict__ llr, int f16) {
    if (l0 < L) {
        float v[2 * NPB];
        qam_llr<NPB>(s0.x, s0.y, noise_inv[sc * L + l0], v);
        const size_t e = ((sc * L + l0) * S + n) * (2 * NPB);
        if (f16)
            store_llr_f16<NPB>(reinterpret_cast<uint16_t *>(llr) + e, v);
        else
            store_llr<NPB>(llr + e, v);
    }
    if (l0 + 1 < L) {
        float v[2 * NPB];
        qam_llr<NPB>(s1.x, s1.y, noise_inv[sc * L + l0 + 1], v);
        const size_t e = ((sc * L + l0 + 1) * S + n) * (2 * NPB);
        if (f16)
            store_llr_f16<NPB>(reinterpret_cast<uint16_t *>(llr) + e, v);
        else
            store_llr<NPB>(llr + e, v);
    }
}

// ---- codeword output: descrambling and the rate-recovery layout (TS 38.211 6.3.1, TS 38.212 5.4.2) ----------
// Codeword records (CwLayer, pe_internal.hpp): LLR i of layer l's symbol at (subcarrier sc, data symbol n) is
// codeword bit g = ((n nSC + sc) nu + v) Qm + i; it is negated where c(g) = 1 and stored at g (codeword order) or at
// its place in the de-interleaved code block.
template <int NPB>
__device__ __forceinline__ void cw_store(float (&v)[2 * NPB], const CwLayer &cl, size_t sc, int n, int nSC,
                                         int f16) {
    constexpr uint32_t QM = 2 * NPB;
    const uint32_t nu = cl.nu_v & 255u, vr = cl.nu_v >> 8;
    const uint32_t g = (((uint32_t)n * (uint32_t)nSC + (uint32_t)sc) * nu + vr) * QM;
    if (cl.cseq) {
        const uint32_t *cs = reinterpret_cast<const uint32_t *>(cl.cseq);
        const uint32_t w = g >> 5, sh = g & 31u;
        uint32_t bits = cs[w] >> sh;
        if (sh + QM > 32u) bits |= cs[w + 1] << (32u - sh);
#pragma unroll
        for (uint32_t i = 0; i < QM; ++i)
            if ((bits >> i) & 1u) v[i] = -v[i];
    }
    if (cl.e_lo == 0u) {  // codeword order: the symbol's Qm LLRs are consecutive
        if (f16)
            store_llr_f16<NPB>(reinterpret_cast<uint16_t *>(cl.out) + g, v);
        else
            store_llr<NPB>(reinterpret_cast<float *>(cl.out) + g, v);
        return;
    }
    const uint32_t lo_len = cl.r0 * cl.e_lo;
    uint32_t start, E;
    if (g < lo_len) {
        E = cl.e_lo;
        start = g / E * E;
    } else {
        E = cl.e_lo + nu * QM;
        start = lo_len + (g - lo_len) / E * E;
    }
    const uint32_t j = (g - start) / QM, stride = E / QM;  // symbol j of its block; bit i goes to plane i
#pragma unroll
    for (uint32_t i = 0; i < QM; ++i) {
        const uint32_t pos = start + i * stride + j;
        if (f16)
            reinterpret_cast<uint16_t *>(cl.out)[pos] = (uint16_t)float_to_half_bits(v[i]);
        else
            reinterpret_cast<float *>(cl.out)[pos] = v[i];
    }
}

// The demapped LLRs of the two symbols a lane holds (layers l0, l0 + 1, data symbol n) in codeword order.
template <int NPB>
__device__ __forceinline__ void cw_demap_store(float2 s0, float2 s1, size_t sc, int l0, int n, int L, int nSC,
                                               const float *__restrict__ noise_inv, const CwLayer *cwl, int f16) {
    if (l0 < L) {
        float v[2 * NPB];
        qam_llr<NPB>(s0.x, s0.y, noise_inv[sc * L + l0], v);
        cw_store<NPB>(v, cwl[l0], sc, n, nSC, f16);
    }
    if (l0 + 1 < L) {
        float v[2 * NPB];
        qam_llr<NPB>(s1.x, s1.y, noise_inv[sc * L + l0 + 1], v);
        cw_store<NPB>(v, cwl[l0 + 1], sc, n, nSC, f16);
    }
}

// One subcarrier: range scaling, operand conversion into LDS, MFMA, epilogue to Y_eq or LLRs.
// NPB: 0 -> store Y_eq; 1..4 -> fused demap of that QAM order; -1 -> fused demap of order npb_rt.
// C16IN: R.c holds binary16 coefficient chunks (load_raw_c16; nC, rc count those) scaled by 2^ec_in; only the
// samples are range-checked here, and only PREC 1 applies.
// CCOLS: the coefficients come as columns in *cc instead of R.c (nC, rc unused). The values, the scale
// decision and every operand written to LDS are those of the chunk path on the same complex64 coefficients.
// CW: the LLRs go to their codewords (cwl: the eight layer records, nSC: subcarriers of the allocation) instead
// of llr; the QAM order is npb_rt.
template <int KS, int PREC, int NPB, bool C16IN = false, bool CCOLS = false, bool CW = false>
__device__ __forceinline__ void process_sc(RawSC<KS> &R, size_t sc, unsigned char *lds, float2 *__restrict__ O,
                                           const float *__restrict__ noise_inv, float *__restrict__ llr, int L,
                                           int S, int A, int lane, int nC, int nY, int rc, int ry,
                                           int npb_rt = 0, int f16 = 0, int ec_in = 0,
                                           const CCols *cc = nullptr, const CwLayer *cwl = nullptr, int nSC = 0) {
    static_assert(!C16IN || PREC == 1, "binary16 coefficient input is for the binary16 product");
    static_assert(!(C16IN && CCOLS), "one coefficient source");
    constexpr int NU = KS > 4 ? 2 : 1;  // antennas per lane in the column form
    constexpr bool BF = (PREC == 2 || PREC == 4);
    constexpr bool SPLIT = (PREC == 3 || PREC == 4);
    constexpr int AG = 4 * KS;
    constexpr int RB = LdsLayout<KS>::RB;
    constexpr int CREG = LdsLayout<KS>::CREG;
    constexpr int HI = LdsLayout<KS>::HI;
    typedef Op16<BF> Op;
    typedef typename Op::vec vec;
    const int g = lane >> 4;
    const int r = lane & 15;
    const int cpr = A >> 1;  // 16-byte chunks per row
    // c / cpr as a multiply-shift (error < c / 2^20 < 1 / cpr): exact for every c < 3072, cpr <= 128
    const unsigned cpr_mag = ((1u << 20) + (unsigned)cpr - 1u) / (unsigned)cpr;

    // per-subcarrier power-of-two scales of C and Y
    float mc = 0.f, my = 0.f;
    if constexpr (CCOLS) {
#pragma unroll
        for (int u = 0; u < NU; ++u)
#pragma unroll
            for (int l = 0; l < 8; ++l) mc = max3abs(mc, cc->r[u][l], cc->i[u][l]);
    } else {
#pragma unroll
        for (int j = 0; j < RawSC<KS>::RC; ++j)
            if (!C16IN && j < rc) mc = max4abs(mc, R.c[j]);
    }
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RY; ++j)
        if (j < ry) my = max4abs(my, R.y[j]);
    // binary16 keeps full precision on [2^-14, 65504]: when both per-subcarrier maxima lie in [2^-9, 2^14)
    // the operands are already in range and scaling is skipped. The test needs only wave votes on per-lane
    // maxima; the cross-lane max reductions run only when scaling is needed. bf16 has the fp32 exponent range.
    bool scaled = true;
    if constexpr (C16IN) {
        scaled = __any(my >= 0x1p14f) || !__any(my >= 0x1p-9f);  // the samples alone
    } else if constexpr (!SPLIT) {
        if constexpr (BF) {
            scaled = false;
        } else {
            const float lo_ = 0x1p-9f, hi_ = 0x1p14f;
            scaled = __any(mc >= hi_ || my >= hi_) || !__any(mc >= lo_) || !__any(my >= lo_);
        }
    }
    int ec = 0, ey = 0;
    if constexpr (C16IN) {
        ec = ec_in;
        if (scaled) ey = scale_exp(wave_max_xor(my, 1, 64));
    } else if (scaled) {
        ec = scale_exp(wave_max_xor(mc, 1, 64));
        ey = scale_exp(wave_max_xor(my, 1, 64));
    }
    const float sc_c = ldexpf(1.f, -ec), sc_y = ldexpf(1.f, -ey);

    // convert and write the operands
    if constexpr (C16IN) {  // binary16 chunks of 4 pairs, copied as they are
        const int cpr4 = A >> 2;
        const unsigned cpr4_mag = ((1u << 20) + (unsigned)cpr4 - 1u) / (unsigned)cpr4;
#pragma unroll
        for (int j = 0; j < RawSC<KS>::RC; ++j) {
            const int c = lane + 64 * j;
            if (j < rc && c < nC) {
                const int row = (int)(((unsigned)c * cpr4_mag) >> 20), part = c - row * cpr4;
                *reinterpret_cast<float4 *>(lds + row * RB + part * 16) = R.c[j];
            }
        }
    }
    if constexpr (CCOLS) {  // pair (Re, Im) of C[l][a] at the same place as in the chunk path
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int a = lane + 64 * u;
            if (a < A) {
#pragma unroll
                for (int l = 0; l < 8; ++l)
                    if (l < L) {
                        float re = cc->r[u][l], im = cc->i[u][l];
                        if (scaled) { re *= sc_c; im *= sc_c; }
                        const uint32_t p = pack2<BF>(re, im);
                        *reinterpret_cast<uint32_t *>(lds + l * RB + a * 4) = p;
                        if constexpr (SPLIT) {
                            const float2 h = unpack2<BF>(p);
                            *reinterpret_cast<uint32_t *>(lds + HI + l * RB + a * 4) =
                                pack2<BF>(re - h.x, im - h.y);
                        }
                    }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RC; ++j) {
        const int c = lane + 64 * j;
        if (!C16IN && !CCOLS && j < rc && c < nC) {
            float4 v = R.c[j];
            if (scaled) { v.x *= sc_c; v.y *= sc_c; v.z *= sc_c; v.w *= sc_c; }
            const int row = (int)(((unsigned)c * cpr_mag) >> 20), part = c - row * cpr;
            const uint32_t p0 = pack2<BF>(v.x, v.y), p1 = pack2<BF>(v.z, v.w);
            *reinterpret_cast<uint2 *>(lds + row * RB + part * 8) = make_uint2(p0, p1);
            if constexpr (SPLIT) {
                const float2 h0 = unpack2<BF>(p0), h1 = unpack2<BF>(p1);
                *reinterpret_cast<uint2 *>(lds + HI + row * RB + part * 8) =
                    make_uint2(pack2<BF>(v.x - h0.x, v.y - h0.y), pack2<BF>(v.z - h1.x, v.w - h1.y));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < RawSC<KS>::RY; ++j) {
        const int c = lane + 64 * j;
        if (j < ry && c < nY) {
            float4 v = R.y[j];
            if (scaled) { v.x *= sc_y; v.y *= sc_y; v.z *= sc_y; v.w *= sc_y; }
            const int col = (int)(((unsigned)c * cpr_mag) >> 20), part = c - col * cpr;
            const uint32_t p0 = pack2<BF>(v.x, v.y), p1 = pack2<BF>(v.z, v.w);
            *reinterpret_cast<uint2 *>(lds + CREG + col * RB + part * 8) = make_uint2(p0, p1);
            if constexpr (SPLIT) {
                const float2 h0 = unpack2<BF>(p0), h1 = unpack2<BF>(p1);
                *reinterpret_cast<uint2 *>(lds + HI + CREG + col * RB + part * 8) =
                    make_uint2(pack2<BF>(v.x - h0.x, v.y - h0.y), pack2<BF>(v.z - h1.x, v.w - h1.y));
            }
        }
    }
    lds_order_wave();  // operand writes before the cross-lane reads

    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    f32x4 acc2 = {0.f, 0.f, 0.f, 0.f};
    const bool part_im = r >= 8;
    const uint32_t perm_sel = part_im ? 0x01000302u : 0x03020100u;  // swap halves : identity
    const uint32_t sign_mask = part_im ? 0u : 0x80000000u;          // negate Im on the real-part rows
    const int coff = (r & 7) * RB + g * AG * 4;
    const int yoff = CREG + r * RB + g * AG * 4;
    auto a_operand = [&](int base) {
        uint4 q = *reinterpret_cast<const uint4 *>(lds + base);
        q.x = __builtin_amdgcn_perm(q.x, q.x, perm_sel) ^ sign_mask;
        q.y = __builtin_amdgcn_perm(q.y, q.y, perm_sel) ^ sign_mask;
        q.z = __builtin_amdgcn_perm(q.z, q.z, perm_sel) ^ sign_mask;
        q.w = __builtin_amdgcn_perm(q.w, q.w, perm_sel) ^ sign_mask;
        return __builtin_bit_cast(vec, q);
    };
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        const vec a = a_operand(coff + 16 * t);
        const vec b = __builtin_bit_cast(vec, *reinterpret_cast<const uint4 *>(lds + yoff + 16 * t));
        acc = Op::mfma(a, b, acc);
        if constexpr (SPLIT) {
            const vec al = a_operand(HI + coff + 16 * t);
            const vec bl = __builtin_bit_cast(vec, *reinterpret_cast<const uint4 *>(lds + HI + yoff + 16 * t));
            acc2 = Op::mfma(a, bl, acc2);
            acc2 = Op::mfma(al, b, acc2);
        }
    }
    if constexpr (SPLIT) acc += acc2;
    lds_order_wave();  // operand reads before the next subcarrier rewrites LDS

    // acc[j] = D[4g + j][r]; undo the scaling (exact power of two)
    const int e = ec + ey;
    float d[4] = {acc[0], acc[1], acc[2], acc[3]};
    if (e != 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = ldexpf(d[j], e);
    }
    // pair real and imaginary parts: lanes g<2 hold Re of layers 4g..4g+3, lanes g>=2 hold Im of layers
    // 4(g-2)..4(g-2)+3; after one xor-32 exchange each lane holds two complete symbols
    const bool lo = g < 2;
    const float x0 = lo ? d[2] : d[0];
    const float x1 = lo ? d[3] : d[1];
    const float y0 = __shfl_xor(x0, 32, 64);
    const float y1 = __shfl_xor(x1, 32, 64);
    const int l0 = lo ? 4 * g : 4 * (g - 2) + 2;
    const float2 s0 = lo ? make_float2(d[0], y0) : make_float2(y0, d[2]);
    const float2 s1 = lo ? make_float2(d[1], y1) : make_float2(y1, d[3]);
    const int n = r;
    if constexpr (CW) {
        if (n < S) {
            switch (npb_rt) {
            case 1: cw_demap_store<1>(s0, s1, sc, l0, n, L, nSC, noise_inv, cwl, f16); break;
            case 2: cw_demap_store<2>(s0, s1, sc, l0, n, L, nSC, noise_inv, cwl, f16); break;
            case 3: cw_demap_store<3>(s0, s1, sc, l0, n, L, nSC, noise_inv, cwl, f16); break;
            default: cw_demap_store<4>(s0, s1, sc, l0, n, L, nSC, noise_inv, cwl, f16); break;
            }
        }
        return;
    }
    if (n < S) {
        if constexpr (NPB == 0) {
            if (l0 < L) O[(sc * L + l0) * S + n] = s0;
            if (l0 + 1 < L) O[(sc * L + l0 + 1) * S + n] = s1;
        } else if constexpr (NPB > 0) {
            demap_store<NPB>(s0, s1, sc, l0, n, L, S, noise_inv, llr, f16);
        } else {  // QAM order chosen at run time (multi-cell kernels): only the epilogue is duplicated
            switch (npb_rt) {
            case 1: demap_store<1>(s0, s1, sc, l0, n, L, S, noise_inv, llr, f16); break;
            case 2: demap_store<2>(s0, s1, sc, l0, n, L, S, noise_inv, llr, f16); break;
            case 3: demap_store<3>(s0, s1, sc, l0, n, L, S, noise_inv, llr, f16); break;
            default: demap_store<4>(s0, s1, sc, l0, n, L, S, noise_inv, llr, f16); break;
            }
        }
    }
}

// One subcarrier as a separate function: called in a loop, it keeps the per-subcarrier registers short-lived
// (inlined into a loop, the compiler hoists per-lane address arithmetic out of it and doubles the registers).
template <int KS, int PREC, int NPB>
__device__ __attribute__((noinline)) void apply_one_sc(const float2 *__restrict__ C, const float2 *__restrict__ Y,
                                                      float2 *__restrict__ O, const float *__restrict__ noise_inv,
                                                      float *__restrict__ llr, long sc, int L, int S, int A,
                                                      unsigned char *lds, int lane, int npb_rt, int f16 = 0) {
    const int nC = L * (A >> 1), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw<KS>(ra, C, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
    process_sc<KS, PREC, NPB>(ra, (size_t)sc, lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, npb_rt, f16);
}

// The same from binary16 coefficients (C16 pairs, exponents in cexp; binary16 product only).
template <int KS, int NPB>
__device__ __attribute__((noinline)) void apply_one_sc_c16(const uint32_t *__restrict__ C16,
                                                          const int *__restrict__ cexp,
                                                          const float2 *__restrict__ Y, float2 *__restrict__ O,
                                                          const float *__restrict__ noise_inv,
                                                          float *__restrict__ llr, long sc, int L, int S, int A,
                                                          unsigned char *lds, int lane, int npb_rt) {
    const int nC = L * (A >> 2), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw_c16<KS>(ra, C16, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
    process_sc<KS, 1, NPB, true>(ra, (size_t)sc, lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, npb_rt, 0,
                                 cexp[sc]);
}

// Subcarriers [first, last) of one cell by one wave (a one-wave workgroup). PIPE: the next subcarrier's raw
// data is loaded into a second register set while the current one is processed (one subcarrier of loads
// always in flight, at about 3x the registers); without it the wave relies on other resident waves to
// overlap its loads, which measured faster on MI355X.
template <int KS, int PREC, int NPB, bool PIPE>
__device__ __forceinline__ void apply_range(const float2 *__restrict__ C, const float2 *__restrict__ Y,
                                            float2 *__restrict__ O, const float *__restrict__ noise_inv,
                                            float *__restrict__ llr, long first, long last, int L, int S, int A,
                                            unsigned char *lds, int lane, int npb_rt = 0, int f16 = 0) {
    const int nC = L * (A >> 1), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    if constexpr (!PIPE) {
#pragma nounroll
        for (long sc = first; sc < last; ++sc)
            apply_one_sc<KS, PREC, NPB>(C, Y, O, noise_inv, llr, sc, L, S, A, lds, lane, npb_rt, f16);
        return;
    }
    RawSC<KS> ra, rb;
    load_raw<KS>(ra, C, Y, (size_t)first, L, S, A, lane, nC, nY, rc, ry);
    for (long sc = first; sc < last; sc += 2) {
        if (sc + 1 < last) load_raw<KS>(rb, C, Y, (size_t)(sc + 1), L, S, A, lane, nC, nY, rc, ry);
        process_sc<KS, PREC, NPB>(ra, (size_t)sc, lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, npb_rt,
                                  f16);
        if (sc + 1 >= last) break;
        if (sc + 2 < last) load_raw<KS>(ra, C, Y, (size_t)(sc + 2), L, S, A, lane, nC, nY, rc, ry);
        process_sc<KS, PREC, NPB>(rb, (size_t)(sc + 1), lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, npb_rt,
                                  f16);
    }
}

}  // namespace mfma
}  // namespace pe
