// mmse_apply.hip -- channel estimates to LLRs without the coefficient matrix in HBM: the MMSE stage writes
// its compact solve per subcarrier (Ginv as a lower triangle plus the L output scales: L (L + 2) floats,
// 320 B at 8 layers, instead of the 4 KB coefficient block) and the fused apply forms the coefficients from
// it and the antenna rows of H while the Y block is in flight.
//
//   1. mmse_solve_kernel: the MMSE kernel's Gram and L x L algebra (tile_solve) per tile; the lane that solves
//      a subcarrier stores its record at the start of that subcarrier's block of the caller's C workspace; the
//      subcarriers the FP32 algebra flags are solved in FP64 by the same workgroup, their coefficient blocks
//      written to C; noise_inv and flags are written as by pe_gram_mmse. Tiles are sized so that the batch
//      spreads evenly over the workgroups resident at once (the launch lasts as long as its busiest compute
//      unit; the tile changes no result).
//   2. apply_solve_kernel, one subcarrier per wave: lane a reads antenna row a of H and forms the column
//      C[.][a] with the operation sequence of the MMSE kernel's c_values (Ginv entries expanded once into a
//      full matrix in LDS and read there by broadcast); the column goes straight into the operand layout of
//      the fused apply (process_sc, column form), which runs the MFMA product and the demapper as the
//      standalone apply kernel does. Subcarriers solved in FP64 take the coefficient block from C instead.
// Per subcarrier at 8 layers x 64 antennas x 11 symbols this moves 4.4 + 12.7 KB instead of 8.2 + 12.6 KB
// (H is read by both kernels, the coefficient block neither written nor read). The coefficient values, the
// apply and the LLRs are bit-identical to pe_gram_mmse followed by pe_coef_apply_demap at the same precision.
// Shapes for which this does not pay (fewer than 32 antennas or 4 layers: measured slower, single cells being
// launch- and latency-bound) or cannot be done (more than 64 antennas: one column per lane) run pe_gram_mmse's
// and the apply's kernels with C in HBM: the same kernels, hence the same results.
#include <atomic>
#include <mutex>

#include "pe_internal.hpp"
#include "pusch_eq.h"
#include "solve_record.hpp"

// Coefficient workspace of pe_mmse_apply_demap_noalloc: solve records (and the coefficient blocks of subcarriers
// solved in FP64) for up to PE_NOALLOC_SCRATCH_BYTES of coefficients, one copy per device, held until the process
// ends. Host code never passes `scratch` itself (its host-side name is not a device address): it passes the
// address noalloc_scratch() looked up for the current device.
static __device__ __align__(256) unsigned char scratch[PE_NOALLOC_SCRATCH_BYTES];

namespace pe {

hipError_t noalloc_scratch(void **ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {  // once per device, with that device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

namespace {

using namespace mmse;
using namespace mfma;

constexpr int kWpb = 4;  // apply: waves (subcarriers) per workgroup

template <int KS, int PREC>
__host__ __device__ constexpr int wave_lds_bytes() {
    return LdsLayout<KS>::bytes(PREC) + rec::kGinvBytes;
}

template <int L>
__global__ __launch_bounds__(512) void mmse_solve_kernel(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                                         float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                         uint8_t *__restrict__ flags, int nSC, int A, int P, int T,
                                                         int unbiased, float thr) {
    extern __shared__ __attribute__((aligned(16))) float lds_f[];
    const int W = blockDim.x >> 6;
    const long sc0 = (long)blockIdx.x * T;
    const int n = (int)min((long)T, (long)nSC - sc0);
    float kr[1][L], ki[1][L];
    tile_solve<L, 1, true>(H, sigma2, noise_inv, flags, sc0, n, A, P, unbiased != 0, thr, lds_f, threadIdx.x, W,
                           false, kr, ki, C);  // records of the subcarriers solved in FP32
    tile_fp64_inline<L>(H, sigma2, C, noise_inv, flags, sc0, n, A, unbiased != 0, lds_f, threadIdx.x, W);
}

template <int L, int KS, int PREC>
__global__ __launch_bounds__(64 * kWpb) void apply_solve_kernel(
    const float2 *__restrict__ H, const float2 *__restrict__ Y, float2 *C, const float *__restrict__ noise_inv,
    const uint8_t *__restrict__ flags, float *__restrict__ llr, int nSC, int S, int A, int npb, int f16,
    int write_c) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_all[kWpb * wave_lds_bytes<KS, PREC>()];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned char *ops = lds_all + wave * wave_lds_bytes<KS, PREC>();
    float2 *gfull = reinterpret_cast<float2 *>(ops + LdsLayout<KS>::bytes(PREC));
    const long sc = (long)blockIdx.x * kWpb + wave;
    if (sc >= nSC) return;  // uniform per wave; no workgroup barrier follows
    if (A < 16 * KS) zero_operands<KS, PREC>(ops, lane);
    if (flags[sc] & PE_SCFLAG_FP64) {  // solved in FP64: the coefficient block is in C
        const int nC = L * (A >> 1), rc = (nC + 63) >> 6, nY = S * (A >> 1), ry = (nY + 63) >> 6;
        RawSC<KS> ra;
        load_raw<KS>(ra, C, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
        process_sc<KS, PREC, -1>(ra, (size_t)sc, ops, nullptr, noise_inv, llr, L, S, A, lane, nC, nY, rc, ry, npb,
                                 f16);
        return;
    }
    rec::apply_from_record<L, KS, PREC>(L, H, Y, C, noise_inv, llr, sc, S, A, npb, f16, write_c, ops, gfull, lane);
}

// Subcarriers per tile for nSC subcarriers on tiles of at most tmax: tmax when the launch has at most one
// workgroup per compute unit; else the size that spreads the batch evenly over the rounds of `resident`
// workgroups (e.g. 50,688 subcarriers, at most 128 per tile, 512 resident: 99 each, where tiles of 128 would
// leave 140 compute units with two and 116 with one).
int balanced_tile(long nSC, int tmax, long resident) {
    const long tiles = (nSC + tmax - 1) / tmax;
    if (tiles <= device_cus() || resident <= 0) return tmax;
    const long rounds = (tiles + resident - 1) / resident;
    return (int)((nSC + resident * rounds - 1) / (resident * rounds));
}

// The apply of the one-pass chain: coefficients from the records in C (or the FP64 blocks) and H.
template <int L, int KS, int PREC>
hipError_t launch_apply(const float2 *H, const float2 *Y, float2 *C, const float *ni, const uint8_t *flags, float *llr,
                        int nSC, int S, int A, int npb, int f16, int write_c, hipStream_t st) {
    hipLaunchKernelGGL((apply_solve_kernel<L, KS, PREC>), dim3((unsigned)((nSC + kWpb - 1) / kWpb)), dim3(64 * kWpb),
                       0, st, H, Y, C, ni, flags, llr, nSC, S, A, npb, f16, write_c);
    return hipGetLastError();
}

template <int L, int KS, int PREC>
hipError_t launch_l(const float2 *H, const float *s2, const float2 *Y, float2 *C, float *ni, uint8_t *flags,
                    float *llr, int nSC, int S, int A, int npb, bool unbiased, float thr, int f16, int write_c,
                    hipStream_t st) {
    const int P = mmse_lanes_for(L, A);
    const MmseTile mt = solve_tile_for(L, A, nSC);
    const int W = mt.waves;
    if (W < 1 || W > 16 || mt.tile < 1 || mt.tile > 128 || mt.tile > 64 * W) return hipErrorInvalidValue;
    // workgroups of this shape resident per compute unit, per waves-per-workgroup (cached; LDS at the largest tile)
    static std::atomic<int> occ_cache[17][129];
    std::atomic<int> &oc = occ_cache[W][mt.tile];
    int occ = oc.load(std::memory_order_relaxed);
    if (occ <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)&mmse_solve_kernel<L>, 64 * W,
                                                         (size_t)tile_lds_bytes(L, mt.tile)) != hipSuccess ||
            occ <= 0)
            occ = 1;
        oc.store(occ, std::memory_order_relaxed);
    }
    const int T = balanced_tile(nSC, mt.tile, (long)occ * device_cus());
    hipLaunchKernelGGL((mmse_solve_kernel<L>), dim3((unsigned)((nSC + T - 1) / T)), dim3(64 * W),
                       (size_t)tile_lds_bytes(L, T), st, H, s2, C, ni, flags, nSC, A, P, T, unbiased ? 1 : 0, thr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_apply<L, KS, PREC>(H, Y, C, ni, flags, llr, nSC, S, A, npb, f16, write_c, st);
}

template <int KS, int PREC>
hipError_t launch_kp(const float2 *H, const float *s2, const float2 *Y, float2 *C, float *ni, uint8_t *flags,
                     float *llr, int nSC, int L, int S, int A, int npb, bool unbiased, float thr, int f16,
                     int write_c, hipStream_t st) {
#define PE_MA_L(LL)                                                                                               \
    case LL:                                                                                                      \
        return launch_l<LL, KS, PREC>(H, s2, Y, C, ni, flags, llr, nSC, S, A, npb, unbiased, thr, f16, write_c, st);
    switch (L) {  // 4 <= L (record_pays)
        PE_MA_L(4)
        PE_MA_L(5)
        PE_MA_L(6)
        PE_MA_L(7)
        PE_MA_L(8)
    default: return hipErrorInvalidValue;
    }
#undef PE_MA_L
}

template <int KS>
hipError_t launch_k(const float2 *H, const float *s2, const float2 *Y, float2 *C, float *ni, uint8_t *flags,
                    float *llr, int nSC, int L, int S, int A, int npb, int prec, bool unbiased, float thr, int f16,
                    int write_c, hipStream_t st) {
    switch (prec) {
    case 1: return launch_kp<KS, 1>(H, s2, Y, C, ni, flags, llr, nSC, L, S, A, npb, unbiased, thr, f16, write_c, st);
    case 2: return launch_kp<KS, 2>(H, s2, Y, C, ni, flags, llr, nSC, L, S, A, npb, unbiased, thr, f16, write_c, st);
    case 3: return launch_kp<KS, 3>(H, s2, Y, C, ni, flags, llr, nSC, L, S, A, npb, unbiased, thr, f16, write_c, st);
    case 4: return launch_kp<KS, 4>(H, s2, Y, C, ni, flags, llr, nSC, L, S, A, npb, unbiased, thr, f16, write_c, st);
    default: return hipErrorInvalidValue;
    }
}

template <int KS>
hipError_t launch_apply_k(const float2 *H, const float2 *Y, float2 *C, const float *ni, const uint8_t *flags,
                          float *llr, int nSC, int L, int S, int A, int npb, int prec, int f16, int write_c,
                          hipStream_t st) {
#define PE_AP_LP(LL, PP)                                                                                          \
    case PP:                                                                                                      \
        return launch_apply<LL, KS, PP>(H, Y, C, ni, flags, llr, nSC, S, A, npb, f16, write_c, st);
#define PE_AP_L(LL)                                                                                               \
    case LL:                                                                                                      \
        switch (prec) {                                                                                           \
            PE_AP_LP(LL, 1)                                                                                       \
            PE_AP_LP(LL, 2)                                                                                       \
            PE_AP_LP(LL, 3)                                                                                       \
            PE_AP_LP(LL, 4)                                                                                       \
        default: return hipErrorInvalidValue;                                                                     \
        }
    switch (L) {  // 4 <= L (record_pays)
        PE_AP_L(4)
        PE_AP_L(5)
        PE_AP_L(6)
        PE_AP_L(7)
        PE_AP_L(8)
    default: return hipErrorInvalidValue;
    }
#undef PE_AP_L
#undef PE_AP_LP
}

}  // namespace

hipError_t launch_apply_records(const float2 *H, const float2 *Y, float2 *C, const float *noise_inv,
                                const uint8_t *flags, float *llr, int nSC, int L, int S, int A, int nPamBits, int prec,
                                int llr_f16, int write_c, hipStream_t st) {
    if (A > 64 || !rec::record_pays(L, A)) return hipErrorInvalidValue;
    if (A <= 32)
        return launch_apply_k<2>(H, Y, C, noise_inv, flags, llr, nSC, L, S, A, nPamBits, prec, llr_f16, write_c, st);
    return launch_apply_k<4>(H, Y, C, noise_inv, flags, llr, nSC, L, S, A, nPamBits, prec, llr_f16, write_c, st);
}

hipError_t launch_mmse_apply_demap(const float2 *H, const float *sigma2, const float2 *Y, float2 *C,
                                   float *noise_inv, uint8_t *flags, void *llr, int nSC, int L, int S, int A,
                                   int nPamBits, int prec, bool unbiased, float thr, int llr_f16, int write_c,
                                   hipStream_t st) {
    float *l = reinterpret_cast<float *>(llr);
    if (A > 128) return hipErrorInvalidValue;
    if (!rec::record_pays(L, A)) {  // the two kernels with C in HBM
        hipError_t e = launch_gram_mmse(H, sigma2, C, noise_inv, flags, nSC, L, A, unbiased, thr, st);
        if (e != hipSuccess) return e;
        return launch_coef_apply_mfma(C, Y, nullptr, noise_inv, l, nSC, L, S, A, nPamBits, prec, st, llr_f16);
    }
    if (A <= 32) return launch_k<2>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased, thr,
                                    llr_f16, write_c, st);
    return launch_k<4>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased, thr, llr_f16,
                       write_c, st);
}

}  // namespace pe
