// srs_filter.hip -- SRS filter multiply: the channel-estimation step upstream of the equalizer.
//
//   H_est[b][ant][sc][port] = sum_k wf(sc, port, k) * SRS[b][ant][k],   wf = W[sc][k] * FOCC[port][k % foccLen]
//
// (srs_wmma_bench.cu:6-9, 57-90). Complex values are binary16 pairs {re, im}. Per computation block b it is a
// complex GEMM [nAnt x K] @ [K x K*nPorts] with K = 12 or 24 subcarriers: a few MB per slot, so both kernels
// are bound by launch latency rather than HBM or MFMA rate.
//
// srs_f16_kernel: the reference's binary16 chain, bit for bit. wf is formed with separately rounded binary16
//   multiplies and adds (the reference's complex_conjmul multiplies without conjugating, :43-47), then the
//   accumulator runs through two binary16 fused multiply-adds per part and k (hcmadd, :48-52) in k order.
//   One wave per (block, subcarrier, 64 antennas): lanes are antennas; the workgroup's 64 SRS rows (row
//   stride K+1 words: conflict-free lane reads) and its subcarriers' taps are staged in LDS once.
// srs_mfma_kernel: the same products on v_mfma_f32_16x16x32_f16 in the real embedding
//   A[ant][2k + {0,1}] = {Re, Im} SRS[ant][k] (the memory layout itself: 16-byte loads straight from HBM),
//   B[2k][2c] = Re wf, B[2k][2c+1] = Im wf, B[2k+1][2c] = -Im wf, B[2k+1][2c+1] = Re wf  (c = sc*nPorts+port),
//   FP32 accumulation and one rounding to binary16 per output. wf is rounded to binary16 as the reference
//   forms it; the result is the binary16 rounding of the FP32 sum of exact products. One wave per (block,
//   16-column tile, 16-antenna tile).
#include "pe_device.hpp"
#include "pe_internal.hpp"

#pragma clang fp contract(off)

namespace pe {

namespace {

typedef _Float16 h16;
typedef _Float16 f16x8v __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

struct hc {
    h16 re, im;
};

__device__ __forceinline__ hc unpack(uint32_t v) {
    hc r;
    r.re = __builtin_bit_cast(h16, (uint16_t)(v & 0xffffu));
    r.im = __builtin_bit_cast(h16, (uint16_t)(v >> 16));
    return r;
}

__device__ __forceinline__ uint32_t pack(h16 re, h16 im) {
    return (uint32_t)__builtin_bit_cast(uint16_t, re) | ((uint32_t)__builtin_bit_cast(uint16_t, im) << 16);
}

// wf = W[sc][k] * FOCC[port][k % foccLen] with binary16 roundings after each multiply and each add
__device__ __forceinline__ hc filter_tap(hc w, hc f, bool conj) {
    if (conj) w.im = -w.im;
    hc r;
    const h16 a = w.re * f.re, b = w.im * f.im, c = w.re * f.im, d = w.im * f.re;
    r.re = a - b;
    r.im = c + d;
    return r;
}

constexpr int kSrsMaxK = 48;

// ---- binary16 chain ----------------------------------------------------------------------------------
// Workgroup = 4 waves = 4 subcarriers x 64 antennas of one block. The SRS rows and the workgroup's taps
// (4 x K x P, each formed once) are staged in LDS, so the k loop issues only LDS reads and fused multiply-adds.
template <int P>
__global__ __launch_bounds__(256) void srs_f16_kernel(uint32_t *__restrict__ H, const uint32_t *__restrict__ SRS,
                                                      const uint32_t *__restrict__ W,
                                                      const uint32_t *__restrict__ FOCC, int nAnt, int K,
                                                      int foccLen, int conj, int nAntGroups, int nScGroups) {
    __shared__ uint32_t rows[64 * (kSrsMaxK + 1)];
    __shared__ __attribute__((aligned(16))) uint32_t taps[4 * kSrsMaxK * P];
    int wg = blockIdx.x;
    const int scg = wg % nScGroups;
    wg /= nScGroups;
    const int ag = wg % nAntGroups;
    const int b = wg / nAntGroups;
    const int ant0 = ag * 64;
    const int na = min(64, nAnt - ant0);
    const int RS = K + 1;
    const bool cj = conj != 0;

    // the workgroup's SRS rows are one contiguous run of na*K words
    const uint32_t *src = SRS + ((size_t)b * nAnt + ant0) * K;
    for (int i = threadIdx.x; i < na * K; i += 256) {
        const int a = i / K;
        rows[a * RS + (i - a * K)] = src[i];
    }
    // taps[w][k][p] = W[sc][k] * FOCC[p][k % foccLen] for the workgroup's subcarriers sc = 4 scg + w
    for (int i = threadIdx.x; i < 4 * K * P; i += 256) {
        const int w = i / (K * P);
        const int rem = i - w * K * P;
        const int k = rem / P;
        const int p = rem - k * P;
        const int sc = scg * 4 + w;
        if (sc < K) {
            const hc t = filter_tap(unpack(W[(size_t)sc * K + k]), unpack(FOCC[p * foccLen + k % foccLen]), cj);
            taps[i] = pack(t.re, t.im);
        }
    }
    __syncthreads();

    const int w = threadIdx.x >> 6;
    const int sc = scg * 4 + w;
    const int lane = threadIdx.x & 63;
    if (sc >= K || lane >= na) return;
    const uint32_t *row = rows + lane * RS;
    const uint32_t *tw = taps + w * K * P;

    h16 er[P], ei[P];
#pragma unroll
    for (int p = 0; p < P; ++p) er[p] = ei[p] = (h16)0.f;
#pragma unroll 4
    for (int k = 0; k < K; ++k) {
        const hc s = unpack(row[k]);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const hc t = unpack(tw[k * P + p]);
            er[p] = __builtin_fmaf16(t.re, s.re, __builtin_fmaf16(-t.im, s.im, er[p]));
            ei[p] = __builtin_fmaf16(t.re, s.im, __builtin_fmaf16(t.im, s.re, ei[p]));
        }
    }

    uint32_t *dst = H + (((size_t)b * nAnt + ant0 + lane) * K + sc) * P;
    uint32_t o[P];
#pragma unroll
    for (int p = 0; p < P; ++p) o[p] = pack(er[p], ei[p]);
    if constexpr (P == 4) {
        *reinterpret_cast<uint4 *>(dst) = make_uint4(o[0], o[1], o[2], o[3]);
    } else if constexpr (P == 2) {
        *reinterpret_cast<uint2 *>(dst) = make_uint2(o[0], o[1]);
    } else if constexpr (P == 8) {
        reinterpret_cast<uint4 *>(dst)[0] = make_uint4(o[0], o[1], o[2], o[3]);
        reinterpret_cast<uint4 *>(dst)[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int p = 0; p < P; ++p) dst[p] = o[p];
    }
}

// ---- MFMA, FP32 accumulation -------------------------------------------------------------------------
// One wave per (block, 16-column tile, 16-antenna tile): the A loads are issued first, the B fragment is
// formed in registers while they are in flight.
template <int KSTEPS>
__global__ __launch_bounds__(256) void srs_mfma_kernel(uint32_t *__restrict__ H, const uint32_t *__restrict__ SRS,
                                                       const uint32_t *__restrict__ W,
                                                       const uint32_t *__restrict__ FOCC, int nAnt, int K, int P,
                                                       int foccLen, int conj, int nColTiles, int nAntTiles,
                                                       long nWaves) {
    const long wid = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= nWaves) return;  // whole waves only: no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    long rest = wid;
    const int ct = (int)(rest % nColTiles);
    rest /= nColTiles;
    const int at = (int)(rest % nAntTiles);
    const int b = (int)(rest / nAntTiles);
    const int NBC = 2 * K * P;  // real output columns: 2 per (sc, port)
    const int n = ct * 16 + r;
    const bool col_ok = n < NBC;
    const bool cj = conj != 0;

    // A fragments: row ant0 + r, real columns 32t + 8g .. +7 = complex taps 16t + 4g .. +3 (one 16-byte load)
    const int ant0 = at * 16;
    const int ant = ant0 + r;
    const uint4 *arow = reinterpret_cast<const uint4 *>(SRS + ((size_t)b * nAnt + ant) * K);
    uint4 aq[KSTEPS];
#pragma unroll
    for (int t = 0; t < KSTEPS; ++t) {
        const int k0 = 16 * t + 4 * g;
        aq[t] = (ant < nAnt && k0 < K) ? arow[k0 >> 2] : make_uint4(0u, 0u, 0u, 0u);
    }

    // B fragments: lane (g, r) holds B[32t + 8g + i][n], i = 0..7, i.e. complex taps k = 16t + 4g + i/2
    f32x4v acc = {0.f, 0.f, 0.f, 0.f};
    {
        const int c = n >> 1;
        const int sc = col_ok ? c / P : 0;
        const int p = col_ok ? c - sc * P : 0;
        const bool im_col = n & 1;
#pragma unroll
        for (int t = 0; t < KSTEPS; ++t) {
            f16x8v bf;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * t + 4 * g + j;
                h16 v0 = (h16)0.f, v1 = (h16)0.f;
                if (col_ok && k < K) {
                    const hc wf = filter_tap(unpack(W[(size_t)sc * K + k]), unpack(FOCC[p * foccLen + k % foccLen]),
                                             cj);
                    v0 = im_col ? wf.im : wf.re;   // row 2k   (times Re SRS)
                    v1 = im_col ? wf.re : -wf.im;  // row 2k+1 (times Im SRS)
                }
                bf[2 * j] = v0;
                bf[2 * j + 1] = v1;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8v, aq[t]), bf, acc, 0, 0, 0);
        }
    }

    // acc[j] = D[4g + j][n]: even n = Re, odd n = Im of output column c = n/2 for antenna ant0 + 4g + j.
    // One xor-1 exchange pairs them: even lanes store rows 4g, 4g+1, odd lanes rows 4g+2, 4g+3.
    const bool even = (r & 1) == 0;
    const float x0 = even ? acc[2] : acc[0];
    const float x1 = even ? acc[3] : acc[1];
    const float y0 = __shfl_xor(x0, 1, 64);
    const float y1 = __shfl_xor(x1, 1, 64);
    const int j0 = even ? 0 : 2;
    const float re0 = even ? acc[0] : y0, im0 = even ? y0 : acc[2];
    const float re1 = even ? acc[1] : y1, im1 = even ? y1 : acc[3];
    if (col_ok) {
        const int c = n >> 1;
        const int a0 = ant0 + 4 * g + j0;
        uint32_t *dst = H + ((size_t)b * nAnt + a0) * (size_t)(K * P) + c;
        if (a0 < nAnt) dst[0] = pack((h16)re0, (h16)im0);
        if (a0 + 1 < nAnt) dst[K * P] = pack((h16)re1, (h16)im1);
    }
}

template <int P>
hipError_t launch_f16(uint32_t *H, const uint32_t *SRS, const uint32_t *W, const uint32_t *FOCC, int nAnt,
                      int nBlocks, int K, int foccLen, int conj, hipStream_t st) {
    const int nAG = (nAnt + 63) / 64, nSG = (K + 3) / 4;
    const long wgs = (long)nBlocks * nAG * nSG;
    if (wgs > 0x7fffffffL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((srs_f16_kernel<P>), dim3((unsigned)wgs), dim3(256), 0, st, H, SRS, W, FOCC, nAnt, K,
                       foccLen, conj, nAG, nSG);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_srs_filter(uint32_t *H, const uint32_t *SRS, const uint32_t *W, const uint32_t *FOCC, int nAnt,
                             int nBlocks, int K, int nPorts, int foccLen, bool conj, bool mfma, hipStream_t st) {
    const int cj = conj ? 1 : 0;
    if (mfma) {
        const int nCT = (2 * K * nPorts + 15) / 16, nAT = (nAnt + 15) / 16;
        const long waves = (long)nBlocks * nAT * nCT;
        const long wgs = (waves + 3) / 4;
        if (wgs > 0x7fffffffL) return hipErrorInvalidConfiguration;
        const dim3 g((unsigned)wgs), bl(256);
        if (K <= 16)
            hipLaunchKernelGGL((srs_mfma_kernel<1>), g, bl, 0, st, H, SRS, W, FOCC, nAnt, K, nPorts, foccLen, cj,
                               nCT, nAT, waves);
        else if (K <= 32)
            hipLaunchKernelGGL((srs_mfma_kernel<2>), g, bl, 0, st, H, SRS, W, FOCC, nAnt, K, nPorts, foccLen, cj,
                               nCT, nAT, waves);
        else
            hipLaunchKernelGGL((srs_mfma_kernel<3>), g, bl, 0, st, H, SRS, W, FOCC, nAnt, K, nPorts, foccLen, cj,
                               nCT, nAT, waves);
        return hipGetLastError();
    }
    switch (nPorts) {
    case 1: return launch_f16<1>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 2: return launch_f16<2>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 3: return launch_f16<3>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 4: return launch_f16<4>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 5: return launch_f16<5>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 6: return launch_f16<6>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 7: return launch_f16<7>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    case 8: return launch_f16<8>(H, SRS, W, FOCC, nAnt, nBlocks, K, foccLen, cj, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pe
