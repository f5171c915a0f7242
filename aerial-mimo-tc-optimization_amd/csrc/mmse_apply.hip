// mmse_apply.hip -- channel estimates to LLRs in one pass: MMSE coefficients and the fused MFMA apply + demap
// in one kernel, so the coefficients of a subcarrier go from registers into the MFMA operand layout in LDS
// and never travel through HBM (per subcarrier at 8 layers x 64 antennas x 11 symbols: 8 KB of coefficient
// write + read saved out of 20.6 KB).
//
// Per tile of T subcarriers (one workgroup of W waves; the MMSE tile of the same shape):
//   1-2. Gram and the L x L algebra exactly as the MMSE kernel (tile_solve): Ginv, scales and ok flag per
//        subcarrier in LDS slots, noise_inv and flags written;
//   3.   per subcarrier, one wave: lane q computes the coefficient column C[.][q] (and C[.][q + 64]) with the
//        MMSE kernel's c_values, while the Y block of the subcarrier is in flight; the columns go straight
//        into the operand layout of the fused apply (process_sc, column form), which runs the MFMA product and
//        the demapper as the standalone apply kernel does.
// Subcarriers the FP32 pass hands to FP64 are skipped here; the MMSE FP64 pass writes their coefficients to C
// and a third kernel applies them (it leaves after one flag load per lane when nothing was flagged).
// The coefficient values, the apply and the LLRs are therefore bit-identical to pe_gram_mmse followed by
// pe_coef_apply_demap at the same precision.
#include "apply_device.hpp"
#include "mmse_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mmse;
using namespace mfma;

constexpr int kNiBytes = 32;           // per-wave LDS copy of the subcarrier's noise_inv (8 floats)
constexpr size_t kMaxLds = 64 * 1024;  // LDS per workgroup kept within 64 KB

template <int KS, int PREC>
__host__ __device__ constexpr int wave_lds_bytes() {
    return LdsLayout<KS>::bytes(PREC) + kNiBytes;
}

// Phase 3 for one subcarrier by one wave (slot: its LDS slot; ops, ni: the wave's operand region and
// noise_inv copy). A separate function: inlined into the loop over the tile's subcarriers, the compiler keeps
// per-lane address arithmetic of the apply alive across the MMSE phases and nearly doubles the registers.
template <int L, int KS, int PREC>
__device__ __attribute__((noinline)) void fused_one_sc(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                                      const float2 *__restrict__ Y, float2 *__restrict__ C,
                                                      float *__restrict__ llr, const float *__restrict__ slot,
                                                      long sc, int S, int A, unsigned char *ops, float *ni, int lane,
                                                      int npb, int f16, int write_c) {
    constexpr int NT = Dim<L>::NT;
    constexpr int NU = KS > 4 ? 2 : 1;
    float vr[NT], vi[NT], scl[L];
    {
        float in[2 * NT + L];
        slot_load<2 * NT + L>(slot, in);
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            vr[k] = in[k];
            vi[k] = in[NT + k];
        }
#pragma unroll
        for (int l = 0; l < L; ++l) scl[l] = in[2 * NT + l];
    }
    CCols cc;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int a = lane + 64 * u;
        float cr[L], ci[L];
#pragma unroll
        for (int l = 0; l < L; ++l) cr[l] = ci[l] = 0.f;
        if (a < A) {
            float hr[L], hi[L];
            load_row<float, L>(H + ((size_t)sc * A + a) * L, hr, hi);
            c_values<float, L>(hr, hi, vr, vi, scl, cr, ci);
            if (write_c) {
                typedef float f2v __attribute__((ext_vector_type(2)));
                f2v *cb = reinterpret_cast<f2v *>(C + (size_t)sc * L * A) + a;
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    const f2v v = {cr[l], ci[l]};
                    __builtin_nontemporal_store(v, cb + (size_t)l * A);
                }
            }
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            cc.r[u][l] = l < L ? cr[l < L ? l : 0] : 0.f;
            cc.i[u][l] = l < L ? ci[l < L ? l : 0] : 0.f;
        }
    }
    const int nY = S * (A >> 1), ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw_y<KS>(ra, Y, (size_t)sc, S, A, lane, nY, ry);
    lds_order_wave();  // the previous subcarrier's noise_inv reads are done
    // noise_inv of the layers as the algebra computed them (x_l = sigma2 Ginv_ll)
    if (lane < L) ni[lane] = noise_inv_of(sigma2[sc] * slot[tri(lane, lane)]);
    const size_t off = (size_t)sc * L * S * 2 * npb;
    float *llr_sc = f16 ? reinterpret_cast<float *>(reinterpret_cast<uint16_t *>(llr) + off) : llr + off;
    process_sc<KS, PREC, -1, false, true>(ra, 0, ops, nullptr, ni, llr_sc, L, S, A, lane, 0, nY, 0, ry, npb, f16, 0,
                                          &cc);
}

template <int L, int KS, int PREC>
__global__ __launch_bounds__(512) void mmse_apply_kernel(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                                         const float2 *__restrict__ Y, float2 *__restrict__ C,
                                                         float *__restrict__ noise_inv, uint8_t *__restrict__ flags,
                                                         float *__restrict__ llr, int nSC, int S, int A, int P, int T,
                                                         int unbiased, float thr, int npb, int f16, int write_c) {
    constexpr int NT = Dim<L>::NT;
    constexpr int SLOT = Dim<L>::SLOT;
    extern __shared__ __attribute__((aligned(16))) float lds_f[];
    const int W = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sc0 = (long)blockIdx.x * T;
    const int n = (int)min((long)T, (long)nSC - sc0);
    float kr[1][L], ki[1][L];
    tile_solve<L, 1>(H, sigma2, noise_inv, flags, sc0, n, A, P, unbiased != 0, thr, lds_f, threadIdx.x, W, nullptr,
                     false, kr, ki);

    unsigned char *ops = reinterpret_cast<unsigned char *>(lds_f + T * SLOT) + wave * wave_lds_bytes<KS, PREC>();
    float *ni = reinterpret_cast<float *>(ops + LdsLayout<KS>::bytes(PREC));
    if (A < 16 * KS) zero_operands<KS, PREC>(ops, lane);
    for (int t = wave; t < n; t += W) {  // t is uniform per wave
        if (lds_f[t * SLOT + 2 * NT + L] == 0.f) continue;  // flagged: the FP64 pass and the flagged apply take it
        fused_one_sc<L, KS, PREC>(H, sigma2, Y, C, llr, lds_f + t * SLOT, sc0 + t, S, A, ops, ni, lane, npb, f16,
                                  write_c);
    }
}

// The subcarriers the FP64 pass solved (flag PE_SCFLAG_FP64): their coefficients are in C; one wave per 64
// subcarriers, one flag load per lane, the flagged ones applied one after another.
template <int KS, int PREC>
__global__ __launch_bounds__(64) void apply_fp64_kernel(const float2 *__restrict__ C, const float2 *__restrict__ Y,
                                                        const float *__restrict__ noise_inv,
                                                        const uint8_t *__restrict__ flags, float *__restrict__ llr,
                                                        int nSC, int L, int S, int A, int npb, int f16) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[LdsLayout<KS>::bytes(PREC)];
    const long sc0 = (long)blockIdx.x * 64;
    const int lane = threadIdx.x;
    const bool mine = sc0 + lane < nSC && (flags[sc0 + lane] & PE_SCFLAG_FP64);
    unsigned long long m = __ballot(mine);
    if (!m) return;  // uniform
    if (A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
    while (m) {
        const int idx = __builtin_ctzll(m);
        m &= m - 1;
        apply_one_sc<KS, PREC, -1>(C, Y, nullptr, noise_inv, llr, sc0 + idx, L, S, A, lds, lane, npb, f16);
    }
}

template <int L, int KS, int PREC>
hipError_t launch_l(const float2 *H, const float *s2, const float2 *Y, float2 *C, float *ni, uint8_t *flags,
                    float *llr, int nSC, int S, int A, int npb, bool unbiased, float thr, int f16, int write_c,
                    hipStream_t st) {
    const int P = mmse_lanes_for(L, A);
    const MmseTile mt = mmse_tile_for(L, A, nSC);
    int T = mt.tile, W = mt.waves;
    auto lds_of = [&](int t, int w) { return (size_t)tile_lds_bytes(L, t) + (size_t)w * wave_lds_bytes<KS, PREC>(); };
    while (W > 1 && lds_of(T, W) > kMaxLds) {  // fewer waves (and at most 64 per wave subcarriers) per tile
        W >>= 1;
        if (T > 64 * W) T = 64 * W;
    }
    const size_t lds = lds_of(T, W);
    hipLaunchKernelGGL((mmse_apply_kernel<L, KS, PREC>), dim3((unsigned)((nSC + T - 1) / T)), dim3(64 * W), lds, st,
                       H, s2, Y, C, ni, flags, llr, nSC, S, A, P, T, unbiased ? 1 : 0, thr, npb, f16, write_c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_gram_mmse_fp64_pass(H, s2, C, ni, flags, nSC, L, A, unbiased, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((apply_fp64_kernel<KS, PREC>), dim3((unsigned)((nSC + 63) / 64)), dim3(64), 0, st, C, Y, ni,
                       flags, llr, nSC, L, S, A, npb, f16);
    return hipGetLastError();
}

template <int KS, int PREC>
hipError_t launch_kp(const float2 *H, const float *s2, const float2 *Y, float2 *C, float *ni, uint8_t *flags,
                     float *llr, int nSC, int L, int S, int A, int npb, bool unbiased, float thr, int f16,
                     int write_c, hipStream_t st) {
#define PE_MA_L(LL)                                                                                               \
    case LL:                                                                                                      \
        return launch_l<LL, KS, PREC>(H, s2, Y, C, ni, flags, llr, nSC, S, A, npb, unbiased, thr, f16, write_c, st);
    switch (L) {
        PE_MA_L(1)
        PE_MA_L(2)
        PE_MA_L(3)
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

}  // namespace

hipError_t launch_mmse_apply_demap(const float2 *H, const float *sigma2, const float2 *Y, float2 *C,
                                   float *noise_inv, uint8_t *flags, void *llr, int nSC, int L, int S, int A,
                                   int nPamBits, int prec, bool unbiased, float thr, int llr_f16, int write_c,
                                   hipStream_t st) {
    float *l = reinterpret_cast<float *>(llr);
    if (A <= 16) return launch_k<1>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased, thr,
                                    llr_f16, write_c, st);
    if (A <= 32) return launch_k<2>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased, thr,
                                    llr_f16, write_c, st);
    if (A <= 64) return launch_k<4>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased, thr,
                                    llr_f16, write_c, st);
    if (A <= 128) return launch_k<8>(H, sigma2, Y, C, noise_inv, flags, l, nSC, L, S, A, nPamBits, prec, unbiased,
                                     thr, llr_f16, write_c, st);
    return hipErrorInvalidValue;
}

}  // namespace pe
