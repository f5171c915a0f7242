// freq_interp.hip -- DMRS frequency interpolation: the channel-estimation GEMM that produces the equalizer's H.
//
//   H_f[b][m][n] = sum_{k < K} W[m][k] * X[b][k][n]      complex64, one filter W [M][K] for the whole batch
//
// The reference names this GEMM ([64 x 96] @ [96 x 4], batched over OFDM symbols and UEs; README.md:200) but has
// no code for it. X holds the least-squares estimates on the K pilot tones of batch item b (NC columns: ports or
// layers), H_f the interpolated estimates on M subcarriers. Per item at 64 x 96 x 4: 3 KB in, 2 KB out, 196,608
// FLOPs (AI 38 FLOP/B: above the FP32 ridge of 20 FLOP/B, far below the bf16 MFMA ridge).
//
// fi_f32_kernel: the FP32 chain of cma() (aerial_coef_wmma.cu:60-64, fused form) in k order, bit-identical to the
//   oracle. W stays in LDS for the workgroup's lifetime; a thread computes one row m of one item for all NC
//   columns (one W read per 4 NC fused multiply-adds), the items' X blocks are staged once per pass and read by
//   broadcast. VALU-bound (11.3 GFLOP per launch on the bench batch: 143 us at the 78.6 TFLOP/s FP32 VALU rate).
// fi_mfma_kernel: bf16 operands (PE_PREC_BF16; PE_PREC_BF16_SPLIT adds the rounding residuals, A*B ~ Ahi*Bhi +
//   Ahi*Blo + Alo*Bhi), v_mfma_f32_16x16x32_bf16 with FP32 accumulation, real embedding
//     A[(item, n)][2k + p] = part p of X[item][k][n],   B_re[2k][m] = Re W[m][k], B_re[2k+1][m] = -Im W[m][k],
//     B_im[2k][m] = Im W[m][k], B_im[2k+1][m] = Re W[m][k]
//   The filter is shared, so the batch packs into the MFMA rows: one 16-row tile holds 16 / NC items (4 at
//   NC = 4), no padding to 16 columns. Each wave owns 16 output subcarriers m and keeps their B fragments in
//   registers for the whole run; the workgroup's 4 waves cover 64 subcarriers of the same item tile, so the tile
//   is read from HBM once (coalesced 8-byte lanes), converted once into LDS in fragment order, and read there by
//   all 4 waves. Item tiles are double-buffered: the next tile's loads are in flight during the current tile's
//   MFMAs. Output: lane (g, c) holds rows 4g..4g+3 of column m, i.e. at NC = 4 the 4 columns of one item: 32
//   contiguous bytes per lane, 512 per 16 lanes. HBM-bound.
#include <algorithm>
#include <atomic>

#include "apply_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

constexpr int kFiMaxLds = 64 * 1024;

// ---- FP32 chain ------------------------------------------------------------------------------------------
// Thread (slot, m) computes row m of item b0 + slot, all NC columns at once: per pilot one W value (LDS, rows
// padded to an odd length so the distinct rows of a wave fall in distinct banks) feeds 4 NC fused multiply-adds,
// the X row is read by broadcast. Each output keeps its own cma() chain in k order.
template <int NC>
__global__ __launch_bounds__(256) void fi_f32_kernel(const float2 *__restrict__ W, const float2 *__restrict__ X,
                                                     float2 *__restrict__ Hf, int nBatch, int M, int K, int WS,
                                                     int IPW) {
    extern __shared__ __attribute__((aligned(16))) float2 fi_lds[];
    float2 *w = fi_lds;
    float2 *x = fi_lds + (size_t)M * WS;
    const int tid = threadIdx.x;
    for (int i = tid; i < M * K; i += 256) {
        const int m = i / K;
        w[m * WS + (i - m * K)] = W[i];
    }
    const int KN = K * NC;
    const int slot = tid / M, m = tid - slot * M;
    for (long b0 = (long)blockIdx.x * IPW; b0 < nBatch; b0 += (long)gridDim.x * IPW) {
        const int ni = (int)min((long)IPW, (long)nBatch - b0);
        __syncthreads();  // W staged; every thread is done with the previous items' x
        const float2 *xs = X + (size_t)b0 * KN;
        const int nx = ni * KN;
        for (int i0 = 0; i0 < nx; i0 += 8 * 256) {  // 8 loads in flight per thread before their LDS writes
            float2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = xs[min(i0 + u * 256 + tid, nx - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (i0 + u * 256 + tid < nx) x[i0 + u * 256 + tid] = v[u];
        }
        __syncthreads();
        if (slot < ni) {
            const float2 *wr = w + m * WS;
            const float2 *xr = x + slot * KN;
            float2 acc[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n) acc[n] = make_float2(0.f, 0.f);
#pragma unroll 2
            for (int k = 0; k < K; ++k) {
                const float2 wv = wr[k];
#pragma unroll
                for (int n = 0; n < NC; ++n) acc[n] = cma(wv, xr[k * NC + n], acc[n]);
            }
            float2 *out = Hf + ((size_t)(b0 + slot) * M + m) * NC;
#pragma unroll
            for (int n = 0; n < NC; ++n) out[n] = acc[n];
        }
    }
}

hipError_t launch_f32(const float2 *W, const float2 *X, float2 *Hf, int nBatch, int M, int K, int NC,
                      hipStream_t st) {
    const int WS = (K & 1) ? K : K + 1;
    const size_t wbytes = (size_t)M * WS * sizeof(float2), xbytes = (size_t)K * NC * sizeof(float2);
    if (wbytes + xbytes > (size_t)kFiMaxLds) return hipErrorInvalidValue;
    // items per workgroup pass: one row per thread, as many items as fit the 256 threads and the LDS limit
    const int IPW = std::max(1, std::min(256 / M, (int)(((size_t)kFiMaxLds - wbytes) / xbytes)));
    const size_t lds = wbytes + (size_t)IPW * xbytes;
    const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / lds)));
    const long grid = std::min(((long)nBatch + IPW - 1) / IPW, (long)device_cus() * per_cu);
    const dim3 gd((unsigned)grid), bd(256);
#define PE_FI_NC(N)                                                                                           \
    case N:                                                                                                   \
        hipLaunchKernelGGL((fi_f32_kernel<N>), gd, bd, lds, st, W, X, Hf, nBatch, M, K, WS, IPW);             \
        break;
    switch (NC) {
        PE_FI_NC(1) PE_FI_NC(2) PE_FI_NC(3) PE_FI_NC(4) PE_FI_NC(5) PE_FI_NC(6) PE_FI_NC(7) PE_FI_NC(8)
        PE_FI_NC(9) PE_FI_NC(10) PE_FI_NC(11) PE_FI_NC(12) PE_FI_NC(13) PE_FI_NC(14) PE_FI_NC(15) PE_FI_NC(16)
    default: return hipErrorInvalidValue;
    }
#undef PE_FI_NC
    return hipGetLastError();
}

// ---- bf16 MFMA -------------------------------------------------------------------------------------------
// One iteration = one group of TPI tiles of 16 rows (TPI * IPT consecutive items, one contiguous run of X).
// Loads are branch-free (the index is clamped, values past the batch replaced by zero) so that the wait before
// staging covers them alone; the group's stores are issued after the next group is staged, so they drain
// during the next iteration instead of being waited for.
template <int KS, bool SPLIT, int TPI>
__global__ __launch_bounds__(256) void fi_mfma_kernel(const float2 *__restrict__ W, const float2 *__restrict__ X,
                                                      float2 *__restrict__ Hf, int nBatch, int M, int K, int NC,
                                                      int IPT, long nGroups, int nMG, int vec4) {
    using Op = mfma::Op16<true>;
    typedef Op::vec vec;
    constexpr int LV = SPLIT ? 2 : 1;         // precision levels held in LDS and registers
    constexpr int TILEV = LV * KS * 64;       // vec (16 B) per tile
    constexpr int BUFV = TPI * TILEV;         // vec per group buffer
    constexpr int CH = KS * TPI;              // X values per thread per group: 16 rows x 16 KS pilots per tile
    extern __shared__ __attribute__((aligned(16))) unsigned char fi_ops[];
    vec *bufv = reinterpret_cast<vec *>(fi_ops);
    uint32_t *bufw = reinterpret_cast<uint32_t *>(fi_ops);
    const int sink = 2 * BUFV * 4;            // one spare word after both buffers: target of unused chunks

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int mg = (int)(blockIdx.x % nMG);
    const int mt = mg * 4 + wave;
    const int m = mt * 16 + c;
    const bool active = mt * 16 < M;  // waves past the last subcarrier tile still stage X and meet the barriers

    // B fragments of this wave's 16 subcarriers: lane (g, c) holds B[32t + 8g + i][c], i = 2j + p, pilot
    // k = 16t + 4g + j; zero past M and K (the padded K range then adds nothing, whatever A holds there)
    vec br[LV][KS], bi[LV][KS];
#pragma unroll
    for (int t = 0; t < KS; ++t) {
        float vr[8], vi[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * t + 4 * g + j;
            const float2 wv = (m < M && k < K) ? W[(size_t)m * K + k] : make_float2(0.f, 0.f);
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

    // Where the group's value e = tid + 256 q lands (the same for every group): tile tt = e / TE, then
    // item = e' / (K NC), pilot k, column n; row = item NC + n; word tt TILEV 4 + ((t 64 + g' 16 + row) 4 + j)
    // with t = k / 16, g' = (k / 4) % 4, j = k % 4. Past the group: the spare word.
    const int KN = K * NC, TE = IPT * KN, GE = TPI * TE;
    int off[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int e = tid + 256 * q;
        if (e < GE) {
            const int tt = e / TE, e1 = e - tt * TE;
            const int item = e1 / KN, rem = e1 - item * KN;
            const int k = rem / NC, n = rem - k * NC;
            const int row = item * NC + n;
            off[q] = tt * TILEV * 4 + ((((k >> 4) * 64 + ((k >> 2) & 3) * 16 + row) << 2) + (k & 3));
        } else {
            off[q] = sink;
        }
    }
    // output rows of this lane: D[4g + j][c] -> item (4g + j) / NC, column (4g + j) % NC
    int oitem[4], ocol[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        oitem[j] = (4 * g + j) / NC;
        ocol[j] = (4 * g + j) - oitem[j] * NC;
    }

    // both buffers zero: pilots past K and rows past IPT NC are never written and must read as zero
    for (int i = tid; i < 2 * BUFV * 4; i += 256) bufw[i] = 0u;

    const int GI = TPI * IPT;  // items per group
    float2 rv[CH];
    auto load = [&](long grp) {
        const long b0 = grp * GI;
        const int E = (int)min((long)GI, (long)nBatch - b0) * KN;  // >= KN: the group holds an item
        typedef float f2v __attribute__((ext_vector_type(2)));
        const f2v *src = reinterpret_cast<const f2v *>(X + (size_t)b0 * KN);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = tid + 256 * q;
            const f2v v = __builtin_nontemporal_load(src + min(e, E - 1));
            rv[q] = e < E ? make_float2(v.x, v.y) : make_float2(0.f, 0.f);
        }
    };
    auto stage = [&](int bsel) {
        uint32_t *dst = bufw + bsel * BUFV * 4;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const uint32_t hi = mfma::pack2<true>(rv[q].x, rv[q].y);
            const bool spare = off[q] == sink;
            dst[spare ? sink - bsel * BUFV * 4 : off[q]] = hi;
            if constexpr (SPLIT) {
                const float2 h = mfma::unpack2<true>(hi);
                const uint32_t lo = mfma::pack2<true>(rv[q].x - h.x, rv[q].y - h.y);
                dst[spare ? sink - bsel * BUFV * 4 : off[q] + KS * 256] = lo;
            }
        }
    };

    long grp = blockIdx.x / nMG;
    const long stride = gridDim.x / nMG;
    if (grp < nGroups) {
        load(grp);
        __syncthreads();  // zero fill before the first group lands
        stage(0);
    }
    __syncthreads();
    int cur = 0;
    for (; grp < nGroups; grp += stride) {
        const long nx = grp + stride;
        if (nx < nGroups) load(nx);  // in flight during the MFMAs below
        f32x4 ar[TPI], ai[TPI];
        if (active) {
#pragma unroll
            for (int tt = 0; tt < TPI; ++tt) {
                const vec *a = bufv + cur * BUFV + tt * TILEV;
                ar[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
                ai[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < KS; ++t) {
                    const vec ah = a[t * 64 + lane];
                    if constexpr (SPLIT) {
                        const vec al = a[(KS + t) * 64 + lane];
                        ar[tt] = Op::mfma(al, br[0][t], ar[tt]);
                        ai[tt] = Op::mfma(al, bi[0][t], ai[tt]);
                        ar[tt] = Op::mfma(ah, br[1][t], ar[tt]);
                        ai[tt] = Op::mfma(ah, bi[1][t], ai[tt]);
                    }
                    ar[tt] = Op::mfma(ah, br[0][t], ar[tt]);
                    ai[tt] = Op::mfma(ah, bi[0][t], ai[tt]);
                }
            }
        }
        if (nx < nGroups) stage(cur ^ 1);
        if (active && m < M) {
            const long b0 = grp * GI;
            if (vec4 && b0 + GI <= nBatch) {  // full group, 4 columns, H_f 16-byte aligned: one item's 32 bytes
#pragma unroll
                for (int tt = 0; tt < TPI; ++tt) {
                    f32x4 *dst = reinterpret_cast<f32x4 *>(Hf + ((size_t)(b0 + tt * IPT + g) * M + m) * 4);
                    __builtin_nontemporal_store(f32x4{ar[tt][0], ai[tt][0], ar[tt][1], ai[tt][1]}, dst);
                    __builtin_nontemporal_store(f32x4{ar[tt][2], ai[tt][2], ar[tt][3], ai[tt][3]}, dst + 1);
                }
            } else {
#pragma unroll
                for (int tt = 0; tt < TPI; ++tt) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const long b = b0 + tt * IPT + oitem[j];
                        if (oitem[j] < IPT && b < nBatch)
                            Hf[((size_t)b * M + m) * NC + ocol[j]] = make_float2(ar[tt][j], ai[tt][j]);
                    }
                }
            }
        }
        __syncthreads();  // next buffer written; every wave is done reading this one
        cur ^= 1;
    }
}

// MFMA tiles per iteration. Measured on the 57,344-item batch (64 x 96 filter, 4 columns): 1 tile, 116 VGPRs,
// 4 waves per SIMD: 63 us; 2 tiles, 184 VGPRs, 2 waves per SIMD: 68 us.
constexpr int kFiTilesPerIter = 1;

template <bool SPLIT>
hipError_t launch_mfma(const float2 *W, const float2 *X, float2 *Hf, int nBatch, int M, int K, int NC,
                       hipStream_t st) {
    constexpr int TPI = kFiTilesPerIter;
    const int KS = (K + 15) / 16;
    const int IPT = 16 / NC;
    const long nGroups = ((long)nBatch + TPI * IPT - 1) / (TPI * IPT);
    const int nMT = (M + 15) / 16, nMG = (nMT + 3) / 4;
    const int vec4 = NC == 4 && (reinterpret_cast<uintptr_t>(Hf) & 15u) == 0;
    const size_t lds = (size_t)2 * TPI * (SPLIT ? 2 : 1) * KS * 64 * 16 + 16;
    // one resident round of workgroups (groups are strided over them, so a second round would run alone), at
    // most 3 per CU: on the 57,344-item batch 3 ran at 58.8 us against 63.8 us for the 4 that fit the registers
    static std::atomic<int> occ_cache[9];
    auto resident = [&](const void *fn) -> long {
        int occ = occ_cache[KS].load(std::memory_order_relaxed);
        if (occ <= 0) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, lds) != hipSuccess || occ <= 0) occ = 1;
            occ = std::min(occ, 3);
            occ_cache[KS].store(occ, std::memory_order_relaxed);
        }
        return (long)occ * device_cus();
    };
#define PE_FI_KS(N)                                                                                           \
    case N: {                                                                                                 \
        const void *fn = (const void *)&fi_mfma_kernel<N, SPLIT, TPI>;                                        \
        const long per_mg = std::min(nGroups, std::max(1L, resident(fn) / nMG));                              \
        if (per_mg * nMG > 0x7fffffffL) return hipErrorInvalidConfiguration;                                  \
        hipLaunchKernelGGL((fi_mfma_kernel<N, SPLIT, TPI>), dim3((unsigned)(per_mg * nMG)), dim3(256), lds, st, W, X, \
                           Hf, nBatch, M, K, NC, IPT, nGroups, nMG, vec4);                                    \
        break;                                                                                                \
    }
    switch (KS) {
        PE_FI_KS(1)
        PE_FI_KS(2)
        PE_FI_KS(3)
        PE_FI_KS(4)
        PE_FI_KS(5)
        PE_FI_KS(6)
        PE_FI_KS(7)
        PE_FI_KS(8)
    default: return hipErrorInvalidValue;
    }
#undef PE_FI_KS
    return hipGetLastError();
}

}  // namespace

size_t freq_interp_f32_lds_bytes(int M, int K, int NC) {
    const int WS = (K & 1) ? K : K + 1;
    return ((size_t)M * WS + (size_t)K * NC) * sizeof(float2);
}

hipError_t launch_freq_interp(const float2 *W, const float2 *X, float2 *Hf, int nBatch, int M, int K, int NC,
                              int prec, hipStream_t st) {
    if (nBatch <= 0) return hipSuccess;
    if (M < 1 || K < 1 || K > 128 || NC < 1 || NC > 16) return hipErrorInvalidValue;
    switch (prec) {
    case 0: return launch_f32(W, X, Hf, nBatch, M, K, NC, st);
    case 2: return launch_mfma<false>(W, X, Hf, nBatch, M, K, NC, st);
    case 4: return launch_mfma<true>(W, X, Hf, nBatch, M, K, NC, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pe
