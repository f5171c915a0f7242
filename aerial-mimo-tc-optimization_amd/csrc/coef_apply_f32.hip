// coef_apply_f32.hip -- FP32 coefficient apply with the antenna-serial fused multiply-add chain, and the
// FP32 Gram product. Both reproduce the rounding sequence of the reference scalar kernels exactly
// (coef_apply_v1_scalar, aerial_coef_wmma.cu:71-104; gram_scalar, aerial_bottleneck.cu:56-80): every output
// is one chain over a = 0..A-1 in ascending order.
//
// Layout of the work (gfx950, wave64): a workgroup of 256 lanes owns SCB subcarriers. Each lane owns a 4x4
// register tile (4 layers x 4 symbols) of one subcarrier, so a subcarrier needs ceil(L/4)*ceil(S/4) lanes
// (6 at L=8, S=11). C and Y_rx rows are staged through LDS in chunks of KC antennas with coalesced
// 128-bit loads (each 128-bit load is two complex samples), so every HBM byte is read once and each staged
// sample is reused by 4 tiles from LDS instead of being re-fetched through the vector cache. The staging
// addresses of a lane are computed once per workgroup (no divisions in the chunk loop).
#include <algorithm>

#include "pe_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

constexpr int kF32Threads = 256;
constexpr int kF32MaxLds = 64 * 1024;
// Staged 16-byte slots per lane and chunk: the 64 KB LDS budget bounds the slots of a chunk by
// 4096 * KC / (KC + 2), so NQ slots per lane of 256 cover every workgroup the launcher forms.
template <int KC> struct F32Q {
    static constexpr int NQ = (4096 * KC / (KC + 2) + kF32Threads - 1) / kF32Threads;
};

template <int KC, int NPB>
__global__ __launch_bounds__(kF32Threads) void coef_apply_f32_kernel(
    const float2 *__restrict__ C, const float2 *__restrict__ Y, float2 *__restrict__ O,
    const float *__restrict__ noise_inv, float *__restrict__ llr, int nSC, int L, int S, int A, int nLG,
    int nSG, int SCB, int f16) {
    constexpr int RS = KC + 2;  // LDS row stride in complex samples (padding staggers LDS banks)
    constexpr int F4_PER_ROW = KC / 2;
    constexpr int LG_F4 = F4_PER_ROW == 2 ? 1 : F4_PER_ROW == 4 ? 2 : 3;
    constexpr int NQ = F32Q<KC>::NQ;
    extern __shared__ float4 lds_raw[];
    float2 *lds = reinterpret_cast<float2 *>(lds_raw);

    const int R = L + S;  // staged rows per subcarrier: L coefficient rows then S sample rows
    const int tid = threadIdx.x;
    const int tps = nLG * nSG;
    const int sc0 = blockIdx.x * SCB;
    const int nsc = min(SCB, nSC - sc0);

    const int scl = tid / tps;
    const int rem = tid - scl * tps;
    const int lg = rem / nSG;
    const int sg = rem - lg * nSG;
    const bool active = scl < nsc;

    int crow[4], yrow[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        crow[i] = min(4 * lg + i, L - 1);
        yrow[i] = L + min(4 * sg + i, S - 1);
    }

    // staging slots of this lane: q = tid + 256 i over [subcarrier][row][part] of a chunk. The row's source
    // offset from the workgroup's first subcarrier is computed once (bit 31 set: a Y_rx row); the part and
    // the LDS destination follow from q by shifts.
    const float2 *Cw = C + (size_t)sc0 * L * A;
    const float2 *Yw = Y + (size_t)sc0 * S * A;
    const int nf4 = nsc * R * F4_PER_ROW;
    unsigned soff[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int row = (tid + i * kF32Threads) >> LG_F4;
        const int s_l = row / R;
        const int rr = row - s_l * R;
        soff[i] = rr < L ? (unsigned)((s_l * L + rr) * A) : (unsigned)((s_l * S + rr - L) * A) | 0x80000000u;
    }
    float2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = make_float2(0.f, 0.f);

    for (int k0 = 0; k0 < A; k0 += KC) {
        const int kn = min(KC, A - k0);  // multiple of 4 (A % 4 == 0)
        // all of the lane's slot loads are issued before the first LDS write (a slot whose load sat under its
        // own branch would wait for every earlier load in turn); unused slots read the workgroup's first value
        float4 sv[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int q = tid + i * kF32Threads;
            const int p2 = 2 * (q & (F4_PER_ROW - 1));
            const unsigned o = soff[i];
            const float2 *src = (o & 0x80000000u) ? Yw + (o & 0x7fffffffu) : Cw + o;
            const bool use = q < nf4 && p2 < kn;
            sv[i] = *reinterpret_cast<const float4 *>(use ? src + k0 + p2 : Cw);
            if (!use) sv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int q = tid + i * kF32Threads;
            const int p2 = 2 * (q & (F4_PER_ROW - 1));
            if (q < nf4) *reinterpret_cast<float4 *>(lds + (q >> LG_F4) * RS + p2) = sv[i];
        }
        __syncthreads();
        if (active) {
            const float2 *base = lds + scl * R * RS;
            if (kn == KC) {
#pragma unroll
                for (int k = 0; k < KC; ++k) {
                    float2 c[4], y[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[i] = base[crow[i] * RS + k];
#pragma unroll
                    for (int j = 0; j < 4; ++j) y[j] = base[yrow[j] * RS + k];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[i][j] = cma(c[i], y[j], acc[i][j]);
                }
            } else {
                for (int k = 0; k < kn; ++k) {
                    float2 c[4], y[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[i] = base[crow[i] * RS + k];
#pragma unroll
                    for (int j = 0; j < 4; ++j) y[j] = base[yrow[j] * RS + k];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[i][j] = cma(c[i], y[j], acc[i][j]);
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;

    const size_t sc = (size_t)(sc0 + scl);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int l = 4 * lg + i;
        if (l >= L) continue;
        if constexpr (NPB == 0) {
            float2 *o = O + (sc * L + l) * S;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = 4 * sg + j;
                if (s < S) o[s] = acc[i][j];
            }
        } else {
            const float ni = noise_inv[sc * L + l];
            const size_t e0 = (sc * L + l) * S * (2 * NPB);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = 4 * sg + j;
                if (s < S) {
                    float v[2 * NPB];
                    qam_llr<NPB>(acc[i][j].x, acc[i][j].y, ni, v);
                    if (f16)  // binary16 LLR wire format, same layout
                        store_llr_f16<NPB>(reinterpret_cast<uint16_t *>(llr) + e0 + s * (2 * NPB), v);
                    else
                        store_llr<NPB>(llr + e0 + s * (2 * NPB), v);
                }
            }
        }
    }
}

// Gram product, one lane per (sc, r, c) with the antenna-serial cma chain of gram_scalar (bit-exact order). A
// workgroup takes spw subcarriers (spw L^2 lanes); M rows and H rows of the workgroup's subcarriers are staged
// through LDS in chunks of kGramKC antennas with coalesced loads (each element read from HBM once), and every
// lane reads its M row and H column from there.
constexpr int kGramKC = 32;
constexpr int kGramLds = 20 * 1024;
constexpr int kGramBatch = 8;

__global__ __launch_bounds__(256) void gram_f32_kernel(const float2 *__restrict__ M, const float2 *__restrict__ H,
                                                       float2 *__restrict__ G, int nSC, int L, int A, int spw) {
    extern __shared__ float2 g_lds[];
    // M rows padded by one value: the lanes of a wave read L different rows at the same k, which would otherwise
    // all fall in the same LDS bank
    constexpr int MS = kGramKC + 1;
    float2 *lm = g_lds;                        // [spw][L][MS]
    float2 *lh = g_lds + (size_t)spw * L * MS;  // [spw][kGramKC][L]
    const int per = L * L;
    const int s = threadIdx.x / per, rc = threadIdx.x - s * per;
    const int r = rc / L, c = rc - r * L;
    const long sc0 = (long)blockIdx.x * spw;
    const int nv = (int)min((long)spw, (long)nSC - sc0);  // subcarriers of this workgroup
    const bool mine = s < nv;
    float2 acc = make_float2(0.f, 0.f);
    for (int k0 = 0; k0 < A; k0 += kGramKC) {
        const int kc = min(kGramKC, A - k0);
        // Staging, kGramBatch loads per thread in flight before their LDS writes (a load-then-write loop would
        // wait for each load in turn). Element i < nM: M[sc][row][k0 + k], row sr = s L + row, column k (kGramKC
        // is a power of two); then H[sc][k0 .. k0 + kc)[col], kc L contiguous values per subcarrier. Addresses are
        // clamped into the operands so every load is issued unconditionally; only the LDS writes are selective.
        const int nM = nv * L * kGramKC, hs = kc * L, nT = nM + nv * hs;
        for (int b0 = 0; b0 < nT; b0 += kGramBatch * (int)blockDim.x) {
            float2 v[kGramBatch];
            int dst[kGramBatch];
#pragma unroll
            for (int b = 0; b < kGramBatch; ++b) {
                const int i = b0 + b * (int)blockDim.x + (int)threadIdx.x;
                const float2 *src = M + (size_t)sc0 * L * A;
                dst[b] = -1;
                if (i < nM) {
                    const int sr = i / kGramKC, k = i % kGramKC;
                    src = M + ((size_t)sc0 * L + sr) * A + k0 + min(k, kc - 1);
                    if (k < kc) dst[b] = sr * MS + k;
                } else if (i < nT) {
                    const int j = i - nM, ss = j / hs, e = j - ss * hs;
                    src = H + ((size_t)(sc0 + ss) * A + k0) * L + e;
                    dst[b] = (int)(lh - g_lds) + ss * kGramKC * L + e;
                }
                v[b] = *src;
            }
#pragma unroll
            for (int b = 0; b < kGramBatch; ++b)
                if (dst[b] >= 0) g_lds[dst[b]] = v[b];
        }
        __syncthreads();
        if (mine) {
            const float2 *mr = lm + (s * L + r) * MS;
            const float2 *hc = lh + s * kGramKC * L + c;
            for (int k = 0; k < kc; ++k) acc = cma(mr[k], hc[k * L], acc);
        }
        __syncthreads();
    }
    if (mine) G[((size_t)(sc0 + s) * L + r) * L + c] = acc;
}

template <int KC, int NPB>
hipError_t launch_f32_impl(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC,
                           int L, int S, int A, int f16, hipStream_t st) {
    constexpr int RS = KC + 2;
    const int nLG = (L + 3) / 4, nSG = (S + 3) / 4;
    const int tps = nLG * nSG;
    const int rowBytes = (L + S) * RS * (int)sizeof(float2);
    int scb = kF32Threads / tps;
    scb = min(scb, kF32MaxLds / rowBytes);
    scb = min(scb, F32Q<KC>::NQ * kF32Threads / ((L + S) * (KC / 2)));  // staging slots of one chunk
    if (scb < 1) return hipErrorInvalidConfiguration;
    const int grid = (nSC + scb - 1) / scb;
    const size_t lds = (size_t)scb * rowBytes;
    hipLaunchKernelGGL((coef_apply_f32_kernel<KC, NPB>), dim3(grid), dim3(kF32Threads), lds, st, C, Y, O, ni,
                       llr, nSC, L, S, A, nLG, nSG, scb, f16);
    return hipGetLastError();
}

template <int NPB>
hipError_t launch_f32_kc(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                         int S, int A, int f16, hipStream_t st) {
    const CoefTile &t = coef_tile_for(L, S, A, 0);
    if (t.f32_kc == 16) return launch_f32_impl<16, NPB>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    if (t.f32_kc == 4) return launch_f32_impl<4, NPB>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    return launch_f32_impl<8, NPB>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
}

}  // namespace

hipError_t launch_coef_apply_f32(const float2 *C, const float2 *Y, float2 *O, int nSC, int L, int S, int A,
                                 hipStream_t st) {
    return launch_f32_kc<0>(C, Y, O, nullptr, nullptr, nSC, L, S, A, 0, st);
}

hipError_t launch_coef_apply_demap_f32(const float2 *C, const float2 *Y, const float *noise_inv, float *llr,
                                       int nSC, int L, int S, int A, int nPamBits, hipStream_t st, int llr_f16) {
    switch (nPamBits) {
    case 1: return launch_f32_kc<1>(C, Y, nullptr, noise_inv, llr, nSC, L, S, A, llr_f16, st);
    case 2: return launch_f32_kc<2>(C, Y, nullptr, noise_inv, llr, nSC, L, S, A, llr_f16, st);
    case 3: return launch_f32_kc<3>(C, Y, nullptr, noise_inv, llr, nSC, L, S, A, llr_f16, st);
    case 4: return launch_f32_kc<4>(C, Y, nullptr, noise_inv, llr, nSC, L, S, A, llr_f16, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gram_f32(const float2 *M, const float2 *H, float2 *G, int nSC, int L, int A, hipStream_t st) {
    // subcarriers per workgroup: at most 256 lanes and kGramLds bytes of staged M and H chunks
    const int per = L * L;
    const int per_sc = L * (2 * kGramKC + 1) * (int)sizeof(float2);  // staged bytes of one subcarrier
    const int spw = std::max(1, std::min(256 / per, kGramLds / per_sc));
    const int threads = (spw * per + 63) / 64 * 64;
    const size_t lds = (size_t)spw * per_sc;
    const long grid = ((long)nSC + spw - 1) / spw;
    hipLaunchKernelGGL(gram_f32_kernel, dim3((unsigned)grid), dim3(threads), lds, st, M, H, G, nSC, L, A, spw);
    return hipGetLastError();
}

}  // namespace pe
