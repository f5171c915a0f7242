// fronthaul.hip -- block floating point decompression of fronthaul U-plane data into Y_rx [sc][sym][ant].
//
// Format (O-RAN WG4 CUS-plane block floating point, cited by README.md:214): per (symbol, antenna) row, PRBs
// consecutive; a PRB is one udCompParam byte (exponent in bits 3..0) and 24 mantissas I0 Q0 ... I11 Q11 of
// iqWidth bits, two's complement, packed most significant bit first (1 + 3 iqWidth bytes). A sample is
// mantissa * 2^exponent, times the caller's scale in Y_rx.
//
// One 64-lane workgroup per (PRB, symbol); lane = antenna (looping past 64). A lane loads its PRB's bytes
// into registers (byte loads: blocks of odd length are not word-aligned), extracts the 24 mantissas at
// compile-time bit offsets (W is a template parameter), and stores the 12 complex samples; for each
// subcarrier the wave's stores cover consecutive antennas, i.e. contiguous 8-byte values. Every product
// mantissa * 2^exponent is exact in float (at most 16 significant bits, |value| <= 2^30), so the only rounding
// is the multiplication by scale.
#include "pe_internal.hpp"

namespace pe {

namespace {

template <int W>
__global__ __launch_bounds__(64) void bfp_decompress_kernel(const uint8_t *__restrict__ in, float2 *__restrict__ Y,
                                                            int nPrb, int nSym, int nAnt, size_t row_stride,
                                                            float scale) {
    constexpr int BPB = 1 + 3 * W;
    const int prb = blockIdx.x, sym = blockIdx.y;
    for (int ant = threadIdx.x; ant < nAnt; ant += 64) {
        const uint8_t *blk = in + ((size_t)sym * nAnt + ant) * row_stride + (size_t)prb * BPB;
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
        float2 *out = Y + ((size_t)prb * 12 * nSym + sym) * nAnt + ant;
#pragma unroll
        for (int re = 0; re < 12; ++re) out[(size_t)re * nSym * nAnt] = make_float2(v[2 * re], v[2 * re + 1]);
    }
}

template <int W>
hipError_t launch_w(const uint8_t *in, float2 *Y, int nPrb, int nSym, int nAnt, size_t rs, float scale,
                    hipStream_t st) {
    hipLaunchKernelGGL((bfp_decompress_kernel<W>), dim3((unsigned)nPrb, (unsigned)nSym), dim3(64), 0, st, in, Y,
                       nPrb, nSym, nAnt, rs, scale);
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
