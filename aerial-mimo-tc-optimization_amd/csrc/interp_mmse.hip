// interp_mmse.hip -- the chain from DMRS least-squares estimates to LLRs with H formed on chip once.
//
// The reference names the DMRS frequency-interpolation GEMM ([64 x 96] @ [96 x 4], batched over OFDM symbols and
// UEs; README.md:199-200) as the stage upstream of the equalizer, where fusing the two removes the round trip of H
// through HBM. pe_freq_interp followed by pe_mmse_apply_demap writes H, reads it for the Gram matrices, and reads it
// again for the coefficients; here the interpolation kernel forms the Gram matrices from its MFMA results while
// they are in registers and solves them, so H is written once and read once (by the apply).
//
//   interp_solve_kernel<KS, SPLIT, P>: one workgroup per block of 64 output subcarriers (4 waves x 16), looping
//     over the block's antennas 4 at a time: one MFMA tile holds 4 items x 4 columns (item = block * A + antenna,
//     columns = layers) and is computed as pe_freq_interp's bf16 kernel computes it (same operands, fragments and
//     MFMA sequence), so H carries the same values, stored in the equalizer's layout H[sc][ant][layer] with
//     sc = 64 block + m. Lane (g, c) of wave w holds antenna 4 k + g of subcarrier m = 16 w + c after tile k and adds
//     its outer product to one of its P / 4 partial Gram matrices: partial u holds the antennas a = p + P r,
//     p = g + 4 u, in increasing r, which is lane p's sum in the MMSE kernels' P-lane groups (gram_group). The
//     partials are then combined in the order of those kernels' butterfly (group_sum: p ^ 1, p ^ 2, then the
//     half-row and row mirrors p ^ 7, p ^ 15; lanes g ^ 1, g ^ 2, g ^ 3 are 16, 32, 48 lanes away), so the Gram
//     matrices are bit-identical to those of pe_mmse_apply_demap on that H for the same P, and so are the solve,
//     the records, noise_inv and flags. The 16 lanes with g = 0 of each wave solve their subcarriers and store the
//     records as the solve kernel does; flagged subcarriers are solved in FP64 by the workgroup from H in HBM.
//   The apply is pe_mmse_apply_demap's own (launch_apply_records), hence the same LLRs.
// Bytes per subcarrier at 64 antennas, 96 pilots, 4 layers: X 3 KB in (the pilots of 64 antennas), H 2 KB and the
// record 96 B out, against X in + H out, then H in + record out for the two calls.
#include "apply_device.hpp"
#include "mmse_device.hpp"
#include "pe_internal.hpp"
#include "solve_record.hpp"

namespace pe {
namespace {

using namespace mmse;

constexpr int kIsM = 64;   // output subcarriers per block (4 waves x 16)
constexpr int kIsNC = 4;   // layers (MFMA tile columns per item)
constexpr int kIsIPT = 4;  // items (antennas) per tile

// LDS of a workgroup: region 0 holds the two X tile buffers (and a spare word) during the tile loop, then the 64
// ok-flag slots of the solve; without SPLIT, W follows as bf16 pairs [64][KS 16 + 4] (the pad spreads the rows
// over the LDS banks), read there for the MFMA B operands instead of being held in registers (48 fewer VGPRs at 96
// pilots, which lets four waves share a SIMD).
template <int KS, bool SPLIT>
__host__ __device__ constexpr int is_buf_bytes() {
    return 2 * (SPLIT ? 2 : 1) * KS * 64 * 16 + 16;  // two tile buffers and one spare word
}
template <int KS, bool SPLIT>
__host__ __device__ constexpr int is_region0_bytes() {
    return is_buf_bytes<KS, SPLIT>() > kIsM * Dim<kIsNC>::SLOT * 4 ? is_buf_bytes<KS, SPLIT>()
                                                                 : kIsM * Dim<kIsNC>::SLOT * 4;
}
template <int KS> __host__ __device__ constexpr int is_wpitch() { return KS * 16 + 4; }
template <int KS, bool SPLIT>
__host__ __device__ constexpr int is_lds_bytes() {
    return is_region0_bytes<KS, SPLIT>() + (SPLIT ? 0 : kIsM * is_wpitch<KS>() * 4);
}

// Without SPLIT and up to 112 pilots the kernel is held to 128 VGPRs (four waves per SIMD: four workgroups per
// compute unit, as many as the LDS allows; 124 at 96 pilots, no spills); the SPLIT operands need more registers.
template <int KS, bool SPLIT>
__host__ __device__ constexpr int is_min_waves() { return !SPLIT && KS <= 7 ? 4 : 1; }

template <int KS, bool SPLIT, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(is_min_waves<KS, SPLIT>()))) void interp_solve_kernel(const float2 *__restrict__ W, const float2 *__restrict__ X,
                                                           const float *__restrict__ sigma2, float2 *__restrict__ H,
                                                           float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                           uint8_t *__restrict__ flags, int K, int A, int unbiased,
                                                           float thr) {
    static_assert(P == 4 || P == 8 || P == 16, "lanes per subcarrier");
    using Op = mfma::Op16<true>;
    typedef Op::vec vec;
    constexpr int L = kIsNC;
    constexpr int NT = Dim<L>::NT;
    constexpr int SLOT = Dim<L>::SLOT;
    constexpr int PU = P / 4;                 // partial Gram matrices per lane
    constexpr int LV = SPLIT ? 2 : 1;         // precision levels of the operands
    constexpr int TILEV = LV * KS * 64;       // vec (16 B) per tile buffer
    constexpr int CH = KS;                    // X values per thread per tile: 16 rows x 16 KS pilots / 256
    extern __shared__ __attribute__((aligned(16))) unsigned char is_lds[];
    vec *bufv = reinterpret_cast<vec *>(is_lds);
    uint32_t *bufw = reinterpret_cast<uint32_t *>(is_lds);
    constexpr int sink = 2 * TILEV * 4;       // the spare word: target of values past the tile
    constexpr int KP = is_wpitch<KS>();
    float *slots = reinterpret_cast<float *>(is_lds);  // [64][SLOT]: the ok flags, after the tile loop
    uint32_t *wl = reinterpret_cast<uint32_t *>(is_lds + is_region0_bytes<KS, SPLIT>());  // [64][KP] W (no SPLIT)

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int m = wave * 16 + c;  // this lane's output subcarrier in the block
    const long blk = blockIdx.x;
    const long sc = blk * kIsM + m;

    // B fragments of this wave's 16 subcarriers (zero past K), as in the interpolation kernel: in registers with
    // SPLIT, else formed per tile from W's bf16 pairs in LDS (the same bits: bf16(-x) = -bf16(x))
    constexpr int KSR = SPLIT ? KS : 1;
    vec br[LV][KSR], bi[LV][KSR];
    if constexpr (!SPLIT) {
        for (int i = tid; i < kIsM * KP; i += 256) {
            const int mm = i / KP, k = i - mm * KP;
            const float2 wv = k < K ? W[(size_t)mm * K + k] : make_float2(0.f, 0.f);
            wl[i] = mfma::pack2<true>(wv.x, wv.y);
        }
    }
#pragma unroll
    for (int t = 0; t < KSR && SPLIT; ++t) {
        float vr[8], vi[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * t + 4 * g + j;
            const float2 wv = k < K ? W[(size_t)m * K + k] : make_float2(0.f, 0.f);
            vr[2 * j] = wv.x;
            vr[2 * j + 1] = -wv.y;
            vi[2 * j] = wv.y;
            vi[2 * j + 1] = wv.x;
        }
        br[0][t] = Op::cvt(vr);
        bi[0][t] = Op::cvt(vi);
        if constexpr (SPLIT) {
            br[1][t] = Op::cvt_lo(vr, br[0][t]);
            bi[1][t] = Op::cvt_lo(vi, bi[0][t]);
        }
    }
    // where value e = tid + 256 q of a tile (item, pilot k, column n) lands in the buffer: row = item NC + n, word
    // ((k / 16) 64 + ((k / 4) % 4) 16 + row) 4 + k % 4; past the tile, the spare word
    const int KN = K * kIsNC, TE = kIsIPT * KN;
    int off[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int e = tid + 256 * q;
        if (e < TE) {
            const int item = e / KN, rem = e - item * KN;
            const int k = rem / kIsNC, n = rem - k * kIsNC;
            const int row = item * kIsNC + n;
            off[q] = (((k >> 4) * 64 + ((k >> 2) & 3) * 16 + row) << 2) + (k & 3);
        } else {
            off[q] = sink;
        }
    }
    for (int i = tid; i < 2 * TILEV * 4; i += 256) bufw[i] = 0u;  // pilots past K read as zero

    const int ntile = A / kIsIPT;
    float2 rv[CH];
    auto load = [&](int k) {  // tile k: items blk A + 4 k .. + 3, one contiguous run of X
        typedef float f2v __attribute__((ext_vector_type(2)));
        const f2v *src = reinterpret_cast<const f2v *>(X + ((size_t)blk * A + (size_t)kIsIPT * k) * KN);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = tid + 256 * q;
            const f2v v = __builtin_nontemporal_load(src + min(e, TE - 1));
            rv[q] = e < TE ? make_float2(v.x, v.y) : make_float2(0.f, 0.f);
        }
    };
    auto stage = [&](int bsel) {
        uint32_t *dst = bufw + bsel * TILEV * 4;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const uint32_t hi = mfma::pack2<true>(rv[q].x, rv[q].y);
            const bool spare = off[q] == sink;
            dst[spare ? sink - bsel * TILEV * 4 : off[q]] = hi;
            if constexpr (SPLIT) {
                const float2 h = mfma::unpack2<true>(hi);
                const uint32_t lo = mfma::pack2<true>(rv[q].x - h.x, rv[q].y - h.y);
                dst[spare ? sink - bsel * TILEV * 4 : off[q] + KS * 256] = lo;
            }
        }
    };

    load(0);
    __syncthreads();  // zero fill before the first tile lands
    stage(0);
    __syncthreads();
    float pr[PU][NT], pim[PU][NT];
#pragma unroll
    for (int u = 0; u < PU; ++u)
#pragma unroll
        for (int e = 0; e < NT; ++e) pr[u][e] = pim[u][e] = 0.f;
    int cur = 0;
    for (int k0 = 0; k0 < ntile; k0 += PU) {
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int k = k0 + u;  // tile k: antenna 4 k + g, in partial (4 k + g) % P / 4 = u
            if (k >= ntile) break;  // uniform
            const bool more = k + 1 < ntile;
            if (more) load(k + 1);  // in flight during the MFMAs below
            const vec *a = bufv + cur * TILEV;
            f32x4 ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < KS; ++t) {
                const vec ah = a[t * 64 + lane];
                if constexpr (SPLIT) {
                    const vec al = a[(KS + t) * 64 + lane];
                    ar = Op::mfma(al, br[0][t], ar);
                    ai = Op::mfma(al, bi[0][t], ai);
                    ar = Op::mfma(ah, br[1][t], ar);
                    ai = Op::mfma(ah, bi[1][t], ai);
                    ar = Op::mfma(ah, br[0][t], ar);
                    ai = Op::mfma(ah, bi[0][t], ai);
                } else {
                    // pilots 16 t + 4 g .. + 3 of subcarrier m: words (bf16 re | bf16 im << 16); the B operands
                    // are (re, -im) and (im, re) per pilot
                    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                    const u32x4 w4 = *reinterpret_cast<const u32x4 *>(wl + m * KP + 16 * t + 4 * g);
                    const u32x4 r4 = w4 ^ 0x80000000u;
                    const u32x4 i4 = (w4 >> 16) | (w4 << 16);
                    ar = Op::mfma(ah, __builtin_bit_cast(vec, r4), ar);
                    ai = Op::mfma(ah, __builtin_bit_cast(vec, i4), ai);
                }
            }
            // D rows 4 g .. 4 g + 3 of column m: item g's layers 0..3 = antenna 4 k + g's row of H at subcarrier sc
            const float vr[L] = {ar[0], ar[1], ar[2], ar[3]};
            const float vi[L] = {ai[0], ai[1], ai[2], ai[3]};
            f32x4 *dst = reinterpret_cast<f32x4 *>(H + ((size_t)sc * A + (size_t)kIsIPT * k + g) * L);
            dst[0] = f32x4{vr[0], vi[0], vr[1], vi[1]};
            dst[1] = f32x4{vr[2], vi[2], vr[3], vi[3]};
#pragma unroll
            for (int i = 0; i < L; ++i)
#pragma unroll
                for (int j = 0; j <= i; ++j) {
                    pr[u][tri(i, j)] = fma_(vr[i], vr[j], fma_(vi[i], vi[j], pr[u][tri(i, j)]));
                    if (j < i) pim[u][tri(i, j)] = fma_(vr[i], vi[j], fma_(-vi[i], vr[j], pim[u][tri(i, j)]));
                }
            // the next tile is staged after this tile's H stores and Gram updates, which so overlap its loads
            if (more) stage(cur ^ 1);
            __syncthreads();  // next buffer written; every wave done reading this one
            cur ^= 1;
        }
    }

    // the P-lane butterfly over the partials (lane g, partial u = MMSE lane p = g + 4 u)
    float gr[NT], gi[NT];
#pragma unroll
    for (int e = 0; e < NT; ++e) {
        float s[PU], t2[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) s[u] = pr[u][e] + __shfl_xor(pr[u][e], 16);  // p ^ 1
#pragma unroll
        for (int u = 0; u < PU; ++u) s[u] = s[u] + __shfl_xor(s[u], 32);  // p ^ 2
        if constexpr (PU >= 2) {  // p ^ 7: lane g ^ 3, partial u ^ 1
#pragma unroll
            for (int u = 0; u < PU; ++u) t2[u] = s[u] + __shfl_xor(s[u ^ 1], 48);
#pragma unroll
            for (int u = 0; u < PU; ++u) s[u] = t2[u];
        }
        if constexpr (PU >= 4) {  // p ^ 15: lane g ^ 3, partial u ^ 3
#pragma unroll
            for (int u = 0; u < PU; ++u) t2[u] = s[u] + __shfl_xor(s[u ^ 3], 48);
#pragma unroll
            for (int u = 0; u < PU; ++u) s[u] = t2[u];
        }
        gr[e] = s[0];
        float si[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) si[u] = pim[u][e] + __shfl_xor(pim[u][e], 16);
#pragma unroll
        for (int u = 0; u < PU; ++u) si[u] = si[u] + __shfl_xor(si[u], 32);
        if constexpr (PU >= 2) {
#pragma unroll
            for (int u = 0; u < PU; ++u) t2[u] = si[u] + __shfl_xor(si[u ^ 1], 48);
#pragma unroll
            for (int u = 0; u < PU; ++u) si[u] = t2[u];
        }
        if constexpr (PU >= 4) {
#pragma unroll
            for (int u = 0; u < PU; ++u) t2[u] = si[u] + __shfl_xor(si[u ^ 3], 48);
#pragma unroll
            for (int u = 0; u < PU; ++u) si[u] = t2[u];
        }
        gi[e] = si[0];
    }
#pragma unroll
    for (int i = 0; i < L; ++i) gi[tri(i, i)] = 0.f;  // the diagonal is real (never accumulated)

    // the solve and its record, by the lanes with g = 0, as the solve kernel's tile_solve stores them
    if (g == 0) {
        float vr[NT], vi[NT], scl[L], x[L];
        const bool ok = solve<float, L>(gr, gi, sigma2[sc], unbiased != 0, thr, true, vr, vi, scl, x);
        float out[2 * NT + L];
#pragma unroll
        for (int e = 0; e < NT; ++e) {
            out[e] = vr[e];
            out[NT + e] = vi[e];
        }
#pragma unroll
        for (int l = 0; l < L; ++l) out[2 * NT + l] = scl[l];
        slots[m * SLOT + 2 * NT + L] = ok ? 1.f : 0.f;
        if (ok) {
#pragma unroll
            for (int l = 0; l < L; ++l) noise_inv[sc * L + l] = noise_inv_of(x[l]);
            flags[sc] = 0;
            store_floats<2 * NT + L>(reinterpret_cast<float *>(C + (size_t)sc * L * A), out, (L * A) % 2 == 0);
        }
    }
    __syncthreads();  // the block's H stores and ok flags visible to the workgroup
    tile_fp64_inline<L>(H, sigma2, C, noise_inv, flags, blk * kIsM, kIsM, A, unbiased != 0, slots, tid, 4);
}

template <int KS, bool SPLIT>
hipError_t launch_ks(const float2 *W, const float2 *X, const float *s2, float2 *H, float2 *C, float *ni,
                     uint8_t *flags, int nBlk, int K, int A, int P, bool unbiased, float thr, hipStream_t st) {
    const size_t lds = (size_t)is_lds_bytes<KS, SPLIT>();
#define PE_IS_P(PP)                                                                                                \
    case PP:                                                                                                       \
        hipLaunchKernelGGL((interp_solve_kernel<KS, SPLIT, PP>), dim3((unsigned)nBlk), dim3(256), lds, st, W, X, s2, \
                           H, C, ni, flags, K, A, unbiased ? 1 : 0, thr);                                         \
        break;
    switch (P) {
        PE_IS_P(4)
        PE_IS_P(8)
        PE_IS_P(16)
    default: return hipErrorInvalidValue;
    }
#undef PE_IS_P
    return hipGetLastError();
}

template <bool SPLIT>
hipError_t launch_split(const float2 *W, const float2 *X, const float *s2, float2 *H, float2 *C, float *ni,
                        uint8_t *flags, int nBlk, int K, int A, int P, bool unbiased, float thr, hipStream_t st) {
    switch ((K + 15) / 16) {
    case 1: return launch_ks<1, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 2: return launch_ks<2, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 3: return launch_ks<3, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 4: return launch_ks<4, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 5: return launch_ks<5, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 6: return launch_ks<6, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 7: return launch_ks<7, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    case 8: return launch_ks<8, SPLIT>(W, X, s2, H, C, ni, flags, nBlk, K, A, P, unbiased, thr, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_interp_mmse_apply_demap(const float2 *W, const float2 *X, const float *sigma2, const float2 *Y,
                                          float2 *H, float2 *C, float *noise_inv, uint8_t *flags, void *llr, int nBlk,
                                          int K, int S, int A, int nPamBits, int interp_prec, int prec, bool unbiased,
                                          float thr, int llr_f16, int write_c, hipStream_t st) {
    if (nBlk <= 0) return hipSuccess;
    if (K < 1 || K > 128 || A % kIsIPT != 0 || !rec::record_pays(kIsNC, A)) return hipErrorInvalidValue;
    const int P = mmse_lanes_for(kIsNC, A);
    hipError_t e = interp_prec == 4
                       ? launch_split<true>(W, X, sigma2, H, C, noise_inv, flags, nBlk, K, A, P, unbiased, thr, st)
                       : launch_split<false>(W, X, sigma2, H, C, noise_inv, flags, nBlk, K, A, P, unbiased, thr, st);
    if (e != hipSuccess) return e;
    return launch_apply_records(H, Y, C, noise_inv, flags, reinterpret_cast<float *>(llr), nBlk * kIsM, kIsNC, S, A,
                                nPamBits, prec, llr_f16, write_c, st);
}

}  // namespace pe
