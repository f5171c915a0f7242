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
//   oracle. W stays in LDS for the workgroup's lifetime (rows padded to an odd number of complex values, so the
//   16 rows a wave reads at one k fall in distinct banks); each item's X block is staged once and read by
//   broadcast. VALU-bound.
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
__global__ __launch_bounds__(256) void fi_f32_kernel(const float2 *__restrict__ W, const float2 *__restrict__ X,
                                                     float2 *__restrict__ Hf, int nBatch, int M, int K, int NC,
                                                     int WS) {
    extern __shared__ __attribute__((aligned(16))) float2 fi_lds[];
    float2 *w = fi_lds;
    float2 *x = fi_lds + (size_t)M * WS;
    const int tid = threadIdx.x;
    for (int i = tid; i < M * K; i += 256) {
        const int m = i / K;
        w[m * WS + (i - m * K)] = W[i];
    }
    const int MN = M * NC, KN = K * NC;
    for (long b = blockIdx.x; b < nBatch; b += gridDim.x) {
        __syncthreads();  // W staged; every thread is done with the previous item's x
        const float2 *xs = X + (size_t)b * KN;
        for (int i = tid; i < KN; i += 256) x[i] = xs[i];
        __syncthreads();
        float2 *out = Hf + (size_t)b * MN;
        for (int o = tid; o < MN; o += 256) {
            const int m = o / NC, n = o - m * NC;
            const float2 *wr = w + m * WS;
            const float2 *xc = x + n;
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll 4
            for (int k = 0; k < K; ++k) acc = cma(wr[k], xc[k * NC], acc);
            out[o] = acc;
        }
    }
}

// ---- bf16 MFMA -------------------------------------------------------------------------------------------
constexpr int kFiChunks = 8;  // X complex values per thread per tile: 16 rows x 128 pilots / 256 threads

template <int KS, bool SPLIT>
__global__ __launch_bounds__(256) void fi_mfma_kernel(const float2 *__restrict__ W, const float2 *__restrict__ X,
                                                      float2 *__restrict__ Hf, int nBatch, int M, int K, int NC,
                                                      int IPT, long nTiles, int nMG) {
    using Op = mfma::Op16<true>;
    typedef Op::vec vec;
    constexpr int LV = SPLIT ? 2 : 1;   // precision levels held in LDS and registers
    constexpr int BUFV = LV * KS * 64;  // vec (16 B) per tile buffer
    extern __shared__ __attribute__((aligned(16))) unsigned char fi_ops[];
    vec *bufv = reinterpret_cast<vec *>(fi_ops);
    uint32_t *bufw = reinterpret_cast<uint32_t *>(fi_ops);

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
            const float2 wv = (active && m < M && k < K) ? W[(size_t)m * K + k] : make_float2(0.f, 0.f);
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

    // Where the tile's complex value e = tid + 256 q lands in the fragment-ordered buffer (the same for every
    // tile): item = e / (K NC), pilot k, column n; row = item NC + n; word ((t 64 + g' 16 + row) 4 + j) with
    // t = k / 16, g' = (k / 4) % 4, j = k % 4. -1: past the tile.
    const int KN = K * NC, TE = IPT * KN;
    int off[kFiChunks];
#pragma unroll
    for (int q = 0; q < kFiChunks; ++q) {
        const int e = tid + 256 * q;
        if (e < TE) {
            const int item = e / KN, rem = e - item * KN;
            const int k = rem / NC, n = rem - k * NC;
            const int row = item * NC + n;
            off[q] = (((k >> 4) * 64 + ((k >> 2) & 3) * 16 + row) << 2) + (k & 3);
        } else {
            off[q] = -1;
        }
    }
    // output rows of this lane: D[4g + j][c] -> item 4g+j / NC, column 4g+j % NC
    int oitem[4], ocol[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        oitem[j] = (4 * g + j) / NC;
        ocol[j] = (4 * g + j) - oitem[j] * NC;
    }

    // both buffers zero: pilots past K and rows past IPT NC are never written and must read as zero
    for (int i = tid; i < 2 * BUFV * 4; i += 256) bufw[i] = 0u;

    float2 rv[kFiChunks];
    auto load = [&](long tile) {
        const long b0 = tile * IPT;
        const int E = (int)min((long)IPT, (long)nBatch - b0) * KN;
        const float2 *src = X + (size_t)b0 * KN;
#pragma unroll
        for (int q = 0; q < kFiChunks; ++q) {
            const int e = tid + 256 * q;
            rv[q] = e < E ? src[e] : make_float2(0.f, 0.f);
        }
    };
    auto stage = [&](int bsel) {
        uint32_t *dst = bufw + bsel * BUFV * 4;
#pragma unroll
        for (int q = 0; q < kFiChunks; ++q) {
            if (off[q] < 0) continue;
            const uint32_t hi = mfma::pack2<true>(rv[q].x, rv[q].y);
            dst[off[q]] = hi;
            if constexpr (SPLIT) {
                const float2 h = mfma::unpack2<true>(hi);
                dst[off[q] + KS * 256] = mfma::pack2<true>(rv[q].x - h.x, rv[q].y - h.y);
            }
        }
    };

    long it = blockIdx.x / nMG;
    const long stride = gridDim.x / nMG;
    if (it < nTiles) {
        load(it);
        __syncthreads();  // zero fill before the first tile lands
        stage(0);
    }
    __syncthreads();
    int cur = 0;
    for (; it < nTiles; it += stride) {
        const long nx = it + stride;
        if (nx < nTiles) load(nx);  // in flight during the MFMAs below
        if (active) {
            const vec *a = bufv + cur * BUFV;
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
                }
                ar = Op::mfma(ah, br[0][t], ar);
                ai = Op::mfma(ah, bi[0][t], ai);
            }
            const long b0 = it * IPT;
            if (m < M) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const long b = b0 + oitem[j];
                    if (oitem[j] < IPT && b < nBatch)
                        Hf[((size_t)b * M + m) * NC + ocol[j]] = make_float2(ar[j], ai[j]);
                }
            }
        }
        if (nx < nTiles) stage(cur ^ 1);
        __syncthreads();  // next buffer written; every wave is done reading this one
        cur ^= 1;
    }
}

template <bool SPLIT>
hipError_t launch_mfma(const float2 *W, const float2 *X, float2 *Hf, int nBatch, int M, int K, int NC,
                       hipStream_t st) {
    const int KS = (K + 15) / 16;
    const int IPT = 16 / NC;
    const long nTiles = ((long)nBatch + IPT - 1) / IPT;
    const int nMT = (M + 15) / 16, nMG = (nMT + 3) / 4;
    const size_t lds = (size_t)2 * (SPLIT ? 2 : 1) * KS * 64 * 16;
    // one resident round of workgroups (item tiles are strided over them, so a second round would run alone)
    static std::atomic<int> occ_cache[9];
    auto resident = [&](const void *fn) -> long {
        int occ = occ_cache[KS].load(std::memory_order_relaxed);
        if (occ <= 0) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, lds) != hipSuccess || occ <= 0) occ = 1;
            occ_cache[KS].store(occ, std::memory_order_relaxed);
        }
        return (long)occ * device_cus();
    };
#define PE_FI_KS(N)                                                                                           \
    case N: {                                                                                                 \
        const long per_mg = std::min(nTiles, std::max(1L, resident((const void *)&fi_mfma_kernel<N, SPLIT>) / nMG)); \
        if (per_mg * nMG > 0x7fffffffL) return hipErrorInvalidConfiguration;                                    \
        hipLaunchKernelGGL((fi_mfma_kernel<N, SPLIT>), dim3((unsigned)(per_mg * nMG)), dim3(256), lds, st, W, X, Hf, \
                           nBatch, M, K, NC, IPT, nTiles, nMG);                                               \
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
    case 0: {
        const size_t lds = freq_interp_f32_lds_bytes(M, K, NC);
        if (lds > (size_t)kFiMaxLds) return hipErrorInvalidValue;
        const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / lds)));
        const long grid = std::min((long)nBatch, (long)device_cus() * per_cu);
        const int WS = (K & 1) ? K : K + 1;
        hipLaunchKernelGGL(fi_f32_kernel, dim3((unsigned)grid), dim3(256), lds, st, W, X, Hf, nBatch, M, K, NC, WS);
        return hipGetLastError();
    }
    case 2: return launch_mfma<false>(W, X, Hf, nBatch, M, K, NC, st);
    case 4: return launch_mfma<true>(W, X, Hf, nBatch, M, K, NC, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pe
