// fronthaul.hip -- block floating point decompression of fronthaul U-plane data into Y_rx [sc][sym][ant].
//
// Format (O-RAN WG4 CUS-plane block floating point, cited by README.md:214): per (symbol, antenna) row, PRBs
// consecutive; a PRB is one udCompParam byte (exponent in bits 3..0) and 24 mantissas I0 Q0 ... I11 Q11 of
// iqWidth bits, two's complement, packed most significant bit first (1 + 3 iqWidth bytes). A sample is
// mantissa * 2^exponent, times the caller's scale in Y_rx.
//
// One 256-lane workgroup per (symbol, 8 PRBs, 64 antennas), numbered so that each XCD decodes contiguous PRB
// ranges (the 8-PRB spans are not whole cache lines). The input is read row by row (each antenna row's
// 8 PRBs are contiguous bytes: coalesced loads, 4-byte words when rows and blocks are word-aligned, bytes
// otherwise) into LDS; then each lane decodes one PRB of one antenna from LDS (W, the mantissa width, is a
// template parameter, so every bit offset is a constant) and stores its 12 complex samples. For each
// subcarrier a wave's stores cover 64 consecutive antennas: contiguous 512-byte rows of Y_rx. Every product
// mantissa * 2^exponent is exact in float (at most 16 significant bits, |value| <= 2^30), so the only rounding
// is the multiplication by scale.
#include "pe_internal.hpp"

namespace pe {

namespace {

constexpr int kPrbPerWg = 8;  // measured on the 32-cell batch: 4 -> 68.5 us, 8 -> 64.5 us, 16 -> 74.4 us
constexpr unsigned kXcds = 8;  // XCDs of an MI355X

template <int W, bool WORDS>
__global__ __launch_bounds__(256) void bfp_decompress_kernel(const uint8_t *__restrict__ in, float2 *__restrict__ Y,
                                                             int nPrb, int nSym, int nAnt, size_t row_stride,
                                                             float scale) {
    constexpr int BPB = 1 + 3 * W;
    constexpr int SPAN = kPrbPerWg * BPB;              // bytes of one antenna row in this workgroup
    constexpr int PITCH = ((SPAN + 3) / 4) | 1;        // LDS row pitch in words, odd: lanes on distinct banks
    __shared__ uint32_t rows[64 * PITCH];
    // Workgroups go to the 8 XCDs round-robin by launch order; renumbered so that each XCD takes a contiguous run
    // of PRB groups of a row: neighbouring groups share the cache lines at their boundaries, which then hit in
    // that XCD's L2 instead of being fetched twice.
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned nwg = gx * gy * gridDim.z;
    const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = nwg / kXcds, r = nwg % kXcds, x = id % kXcds, k = id / kXcds;
    const unsigned nid = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
    const int prb0 = (int)(nid % gx) * kPrbPerWg, sym = (int)((nid / gx) % gy), ant0 = (int)(nid / (gx * gy)) * 64;
    const int np = min(kPrbPerWg, nPrb - prb0), na = min(64, nAnt - ant0);
    const int span = np * BPB;
    const uint8_t *base = in + ((size_t)sym * nAnt + ant0) * row_stride + (size_t)prb0 * BPB;
    if constexpr (WORDS) {  // rows and spans are whole words
        const int nw = span >> 2, tot = na * nw;
        for (int i0 = 0; i0 < tot; i0 += 8 * 256) {  // 8 loads in flight per thread before their LDS writes
            uint32_t v[8];
            int dst[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = i0 + u * 256 + (int)threadIdx.x, ic = min(i, tot - 1);
                const int a = ic / nw, k = ic - a * nw;
                v[u] = reinterpret_cast<const uint32_t *>(base + (size_t)a * row_stride)[k];
                dst[u] = i < tot ? a * PITCH + k : -1;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (dst[u] >= 0) rows[dst[u]] = v[u];
        }
    } else {
        uint8_t *rb = reinterpret_cast<uint8_t *>(rows);
        for (int i = threadIdx.x; i < na * span; i += 256) {
            const int a = i / span, k = i - a * span;
            rb[a * PITCH * 4 + k] = base[(size_t)a * row_stride + k];
        }
    }
    __syncthreads();

    const int a = threadIdx.x & 63;
    if (a >= na) return;  // no barrier follows
    const uint8_t *rb = reinterpret_cast<const uint8_t *>(rows) + a * PITCH * 4;
    for (int p = threadIdx.x >> 6; p < np; p += 4) {
        const uint8_t *blk = rb + p * BPB;
        uint32_t b[BPB + 2];
#pragma unroll
        for (int i = 0; i < BPB; ++i) b[i] = blk[i];
        b[BPB] = b[BPB + 1] = 0u;
        const int e = (int)(b[0] & 0xfu);
        const float s = ldexpf(scale, e);  // exact: a power-of-two change of scale
        float v[24];
#pragma unroll
        for (int k = 0; k < 24; ++k) {
            const int off = 8 + k * W;  // bit offset from the block start, most significant bit first
            const int bi = off >> 3, sh = off & 7;
            const uint32_t word = (b[bi] << 16) | (b[bi + 1] << 8) | b[bi + 2];
            const int32_t raw = (int32_t)((word >> (24 - sh - W)) & ((1u << W) - 1u));
            const int32_t m = raw >= (1 << (W - 1)) ? raw - (1 << W) : raw;
            v[k] = (float)m * s;
        }
        typedef float f2v __attribute__((ext_vector_type(2)));
        f2v *out = reinterpret_cast<f2v *>(Y + ((size_t)(prb0 + p) * 12 * nSym + sym) * nAnt + ant0 + a);
#pragma unroll
        for (int re = 0; re < 12; ++re) {
            const f2v val = {v[2 * re], v[2 * re + 1]};
            __builtin_nontemporal_store(val, out + (size_t)re * nSym * nAnt);  // written once, streamed out
        }
    }
}

template <int W>
hipError_t launch_w(const uint8_t *in, float2 *Y, int nPrb, int nSym, int nAnt, size_t rs, float scale,
                    hipStream_t st) {
    const dim3 g((unsigned)((nPrb + kPrbPerWg - 1) / kPrbPerWg), (unsigned)nSym, (unsigned)((nAnt + 63) / 64));
    const bool words = (1 + 3 * W) % 4 == 0 && rs % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 3u) == 0;
    if (words)
        hipLaunchKernelGGL((bfp_decompress_kernel<W, true>), g, dim3(256), 0, st, in, Y, nPrb, nSym, nAnt, rs,
                           scale);
    else
        hipLaunchKernelGGL((bfp_decompress_kernel<W, false>), g, dim3(256), 0, st, in, Y, nPrb, nSym, nAnt, rs,
                           scale);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_bfp_decompress(const uint8_t *in, float2 *Y, int nPrb, int nSym, int nAnt, int iqWidth,
                                 size_t row_stride, float scale, hipStream_t st) {
    switch (iqWidth) {
    case 1: return launch_w<1>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 2: return launch_w<2>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 3: return launch_w<3>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 4: return launch_w<4>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 5: return launch_w<5>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 6: return launch_w<6>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 7: return launch_w<7>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 8: return launch_w<8>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 9: return launch_w<9>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 10: return launch_w<10>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 11: return launch_w<11>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 12: return launch_w<12>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 13: return launch_w<13>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 14: return launch_w<14>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 15: return launch_w<15>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    case 16: return launch_w<16>(in, Y, nPrb, nSym, nAnt, row_stride, scale, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pe
