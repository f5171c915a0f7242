// coef_apply_mfma.hip -- single-launch MFMA coefficient apply and fused apply + demap (device functions in
// apply_device.hpp).
#include "apply_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mfma;

// PREC: 1 f16, 2 bf16, 3 f16 split, 4 bf16 split.  NPB: 0 -> store Y_eq, 1..4 -> fused demap.
// One-wave workgroups; each streams `spw` consecutive subcarriers and loads the next subcarrier's raw data
// into a second register set while the current one is converted and multiplied.
// One subcarrier per wave, WPB waves (each with its own LDS operand region) per workgroup: the smallest
// register footprint, so the most waves (and loads) in flight per CU.
template <int KS, int PREC, int NPB, int WPB>
__global__ __launch_bounds__(64 * WPB) void coef_apply_mfma_sc_kernel(
    const float2 *__restrict__ C, const float2 *__restrict__ Y, float2 *__restrict__ O,
    const float *__restrict__ noise_inv, float *__restrict__ llr, int nSC, int L, int S, int A, int f16) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_all[WPB * LdsLayout<KS>::bytes(PREC)];
    const int wave = threadIdx.x >> 6;
    unsigned char *lds = lds_all + wave * LdsLayout<KS>::bytes(PREC);
    const long sc = (long)blockIdx.x * WPB + wave;
    if (sc >= nSC) return;  // uniform per wave; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    if (A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
    const int nC = L * (A >> 1), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw<KS>(ra, C, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
    process_sc<KS, PREC, NPB>(ra, (size_t)sc, lds, O, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, 0, f16);
}

// `spw` consecutive subcarriers per one-wave workgroup, the next one prefetched into a second register set.
template <int KS, int PREC, int NPB>
__global__ __launch_bounds__(64) void coef_apply_mfma_kernel(
    const float2 *__restrict__ C, const float2 *__restrict__ Y, float2 *__restrict__ O,
    const float *__restrict__ noise_inv, float *__restrict__ llr, int nSC, int L, int S, int A, int spw,
    int f16) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[LdsLayout<KS>::bytes(PREC)];
    const long first = (long)blockIdx.x * spw;
    if (first >= nSC) return;  // uniform per wave
    const long last = first + spw < nSC ? first + spw : nSC;
    if (A < 16 * KS) zero_operands<KS, PREC>(lds, threadIdx.x);
    apply_range<KS, PREC, NPB, true>(C, Y, O, noise_inv, llr, first, last, L, S, A, lds, threadIdx.x, 0, f16);
}

template <int KS, int PREC, int NPB>
hipError_t launch3(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, int f16, hipStream_t st) {
    const CoefTile t = coef_tile_for(L, S, A, PREC);
    const int spw = t.mfma_spw > 1 ? t.mfma_spw : 1;
    if (spw == 1) {
        const int wpb = t.mfma_waves;
        const unsigned g = (unsigned)((nSC + wpb - 1) / wpb);
        if (wpb == 4)
            hipLaunchKernelGGL((coef_apply_mfma_sc_kernel<KS, PREC, NPB, 4>), dim3(g), dim3(256), 0, st, C, Y, O, ni,
                               llr, nSC, L, S, A, f16);
        else if (wpb == 2)
            hipLaunchKernelGGL((coef_apply_mfma_sc_kernel<KS, PREC, NPB, 2>), dim3(g), dim3(128), 0, st, C, Y, O, ni,
                               llr, nSC, L, S, A, f16);
        else
            hipLaunchKernelGGL((coef_apply_mfma_sc_kernel<KS, PREC, NPB, 1>), dim3((unsigned)nSC), dim3(64), 0, st,
                               C, Y, O, ni, llr, nSC, L, S, A, f16);
    } else {
        const long grid = (nSC + spw - 1) / spw;
        hipLaunchKernelGGL((coef_apply_mfma_kernel<KS, PREC, NPB>), dim3((unsigned)grid), dim3(64), 0, st, C, Y,
                           O, ni, llr, nSC, L, S, A, spw, f16);
    }
    return hipGetLastError();
}

template <int KS, int PREC>
hipError_t launch2(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, int npb, int f16, hipStream_t st) {
    switch (npb) {
    case 0: return launch3<KS, PREC, 0>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 1: return launch3<KS, PREC, 1>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 2: return launch3<KS, PREC, 2>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 3: return launch3<KS, PREC, 3>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    case 4: return launch3<KS, PREC, 4>(C, Y, O, ni, llr, nSC, L, S, A, f16, st);
    default: return hipErrorInvalidValue;
    }
}

template <int KS>
hipError_t launch1(const float2 *C, const float2 *Y, float2 *O, const float *ni, float *llr, int nSC, int L,
                   int S, int A, int npb, int prec, int f16, hipStream_t st) {
    switch (prec) {
    case 1: return launch2<KS, 1>(C, Y, O, ni, llr, nSC, L, S, A, npb, f16, st);
    case 2: return launch2<KS, 2>(C, Y, O, ni, llr, nSC, L, S, A, npb, f16, st);
    case 3: return launch2<KS, 3>(C, Y, O, ni, llr, nSC, L, S, A, npb, f16, st);
    case 4: return launch2<KS, 4>(C, Y, O, ni, llr, nSC, L, S, A, npb, f16, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_coef_apply_mfma(const float2 *C, const float2 *Y, float2 *O, const float *noise_inv,
                                  float *llr, int nSC, int L, int S, int A, int nPamBits, int prec,
                                  hipStream_t st, int llr_f16) {
    const int npb = llr ? nPamBits : 0;
    if (A <= 16) return launch1<1>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, llr_f16, st);
    if (A <= 32) return launch1<2>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, llr_f16, st);
    if (A <= 64) return launch1<4>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, llr_f16, st);
    if (A <= 128) return launch1<8>(C, Y, O, noise_inv, llr, nSC, L, S, A, npb, prec, llr_f16, st);
    return hipErrorInvalidValue;
}

}  // namespace pe
