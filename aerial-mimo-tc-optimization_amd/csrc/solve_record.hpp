// solve_record.hpp -- the compact MMSE solve of a subcarrier kept in HBM instead of its coefficient block, and
// the fused apply that forms the coefficients on chip from it. Shared by the one-pass entry point
// (mmse_apply.hip) and the slot executor (multicell.hip).
//
// Record: Ginv as a lower triangle (NT real parts, NT imaginary parts) and the L output scales, L (L + 2)
// floats (320 B at 8 layers against the 4 KB coefficient block at 64 antennas), stored at the start of the
// subcarrier's coefficient block (which must hold it: L + 2 <= 2 A). The apply forms the column C[.][a] of
// antenna a on lane a with the operation sequence of the MMSE kernel's c_values, so the coefficients, the
// apply and the LLRs are bit-identical to the two-kernel path with the coefficient block in HBM.
#pragma once
#include "apply_device.hpp"
#include "mmse_device.hpp"

namespace pe {
namespace rec {

__host__ __device__ constexpr int record_floats(int L) { return L * (L + 2); }
// shapes whose coefficients can be formed on chip: one antenna per lane, and the record fits the block
__host__ __device__ constexpr bool record_shape(int L, int A) { return A <= 64 && L + 2 <= 2 * A; }
// shapes for which it pays: measured over the deployment sweep (one cell per launch, f16), the one-pass chain ran
// at 0.93-0.99x of the two launches below 32 antennas or 4 layers, 0.99-1.06x at 32 antennas and 1.04-1.14x at
// 64 antennas with 4-8 layers (1.2x on 32-cell batches of 8 layers x 64 antennas)
__host__ __device__ constexpr bool record_pays(int L, int A) { return record_shape(L, A) && A >= 32 && L >= 4; }
constexpr int kGinvBytes = 8 * 8 * 8;  // per-wave full Ginv, complex [8][8], in LDS

// The coefficient columns of subcarrier sc from its record, by one wave: lane a forms C[.][a] (gfull: the wave's
// Ginv buffer in LDS). write_c: the columns are stored over the block as well.
template <int L>
__device__ __forceinline__ mfma::CCols form_columns(const float2 *__restrict__ H, float2 *C, long sc, int A,
                                                    int write_c, float2 *gfull, int lane) {
#pragma clang fp contract(off)
    constexpr int NT = mmse::Dim<L>::NT;
    mfma::CCols cc;
    float hr[L], hi[L];
    if (lane < A) {
        mmse::load_row<float, L>(H + ((size_t)sc * A + lane) * L, hr, hi);
    } else {
#pragma unroll
        for (int l = 0; l < L; ++l) hr[l] = hi[l] = 0.f;
    }
    const float *rec = reinterpret_cast<const float *>(C + (size_t)sc * L * A);
    lds_order_wave();  // a previous subcarrier's reads of gfull are done
    {
        const int l = lane >> 3, j = lane & 7;
        if (l < L && j < L) {  // Ginv[l][j]: stored lower; for j > l it is conj(V[j][l])
            const int k = j <= l ? mmse::tri(l, j) : mmse::tri(j, l);
            const float im = rec[NT + k];
            gfull[lane] = make_float2(rec[k], j <= l ? im : -im);
        }
    }
    float scl[L];
#pragma unroll
    for (int l = 0; l < L; ++l) scl[l] = rec[2 * NT + l];
    lds_order_wave();
    // C[l][lane] = s_l sum_j Ginv[l][j] conj(H[lane][j]); zero on lanes past A (they then change neither the
    // block maximum nor the operands)
#pragma unroll
    for (int l = 0; l < L; ++l) {
        float re = 0.f, im = 0.f;
#pragma unroll
        for (int j = 0; j < L; ++j) {
            const float2 g = gfull[l * 8 + j];
            re = mmse::fma_(g.x, hr[j], mmse::fma_(g.y, hi[j], re));  // (g_r + i g_i)(h_r - i h_i)
            im = mmse::fma_(g.y, hr[j], mmse::fma_(-g.x, hi[j], im));
        }
        cc.r[0][l] = re * scl[l];
        cc.i[0][l] = im * scl[l];
    }
#pragma unroll
    for (int l = L; l < 8; ++l) cc.r[0][l] = cc.i[0][l] = 0.f;
    if (write_c && lane < A) {  // over the record, whose reads the values above depend on
        typedef float f2v __attribute__((ext_vector_type(2)));
        f2v *cb = reinterpret_cast<f2v *>(C + (size_t)sc * L * A) + lane;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const f2v v = {cc.r[0][l], cc.i[0][l]};
            __builtin_nontemporal_store(v, cb + (size_t)l * A);
        }
    }
    return cc;
}

// Fused apply + demap of subcarrier sc from its record, by one wave (ops: the wave's operand region). LT > 0:
// the layer count is LT; LT = 0: it is the run-time L (only the column formation depends on it).
template <int LT, int KS, int PREC>
__device__ __forceinline__ void apply_from_record(int L, const float2 *__restrict__ H, const float2 *__restrict__ Y,
                                                  float2 *C, const float *__restrict__ noise_inv,
                                                  float *__restrict__ llr, long sc, int S, int A, int npb, int f16,
                                                  int write_c, unsigned char *ops, float2 *gfull, int lane) {
    static_assert(KS <= 4, "one antenna per lane");
    using namespace mfma;
    const int nY = S * (A >> 1), ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw_y<KS>(ra, Y, (size_t)sc, S, A, lane, nY, ry);  // in flight during the coefficients
    CCols cc;
    if constexpr (LT > 0) {
        cc = form_columns<LT>(H, C, sc, A, write_c, gfull, lane);
        L = LT;
    } else {
        switch (L) {
        case 1: cc = form_columns<1>(H, C, sc, A, write_c, gfull, lane); break;
        case 2: cc = form_columns<2>(H, C, sc, A, write_c, gfull, lane); break;
        case 3: cc = form_columns<3>(H, C, sc, A, write_c, gfull, lane); break;
        case 4: cc = form_columns<4>(H, C, sc, A, write_c, gfull, lane); break;
        case 5: cc = form_columns<5>(H, C, sc, A, write_c, gfull, lane); break;
        case 6: cc = form_columns<6>(H, C, sc, A, write_c, gfull, lane); break;
        case 7: cc = form_columns<7>(H, C, sc, A, write_c, gfull, lane); break;
        default: cc = form_columns<8>(H, C, sc, A, write_c, gfull, lane); break;
        }
    }
    process_sc<KS, PREC, -1, false, true>(ra, (size_t)sc, ops, nullptr, noise_inv, llr, L, S, A, lane, 0, nY, 0, ry,
                                          npb, f16, 0, &cc);
}

}  // namespace rec
}  // namespace pe
