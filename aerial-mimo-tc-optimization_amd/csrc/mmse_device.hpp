// mmse_device.hpp -- per-subcarrier MMSE solve as device functions, shared by the single-cell kernels
// (gram_mmse.hip) and the multi-cell tile kernels (multicell.hip).
//
// The reference only names this stage (eqMmseCoefCompHighMimoKernel, aerial_bottleneck.cu:51-53); the
// arithmetic here is the regularized solve that defines C in the reference's headline workload:
//   G = H^H H + sigma2 I (Hermitian L x L),  Cholesky G = R R^H,  W = R^-1,  Ginv = W^H W,
//   C = diag(s) Ginv H^H  with s_l = 1 / (1 - sigma2 Ginv_ll) when unbiased (else 1),
//   noise_inv_l = (1 - x_l) / x_l  with x_l = sigma2 Ginv_ll,
//   condition estimate kappa_1(G) = ||G||_1 ||Ginv||_1; above the threshold, or when the factorization
//   breaks down, the subcarrier is solved again in FP64 by the same workgroup (wave_fp64).
//
// Lanes and subcarriers. P lanes (P = 4, 8 or 16) share a subcarrier wherever antennas are summed: lane p of
// the group takes antenna rows a = p, p + P, ... of H[sc] (rows of L complex values, contiguous, so a group's
// loads are contiguous). The partial Gram matrices are summed across the group by a DPP butterfly, after
// which every lane of the group holds the same G.
//   * tile solve (FP32 pass): one wave owns a tile of up to 64 subcarriers. Gram: 64/P subcarriers per
//     round, the group sum of each written to the wave's LDS slot of that subcarrier. Algebra: one lane per
//     subcarrier reads its G from LDS and writes Ginv, the output scales and an ok flag back. Coefficients:
//     one lane per antenna (up to 64 lanes per subcarrier), reading Ginv from LDS, so stores cover whole C
//     rows. With 64/P subcarriers per tile the Gram is a single round (lowest latency); with 64 the algebra
//     runs on every lane without redundancy. The tile size changes no result.
//   * FP64 solve of a flagged subcarrier: one wave, lane 8 i + j holding element (i, j) of the L x L matrices;
//     independent of P and of the tile.
// Results depend on P only through the FP32 Gram summation order; every path takes P from the same per-(L, A)
// choice, and the arithmetic is written with explicit fused multiply-adds and contraction off, so the
// single-cell kernels and the multi-cell kernels produce bit-identical coefficients.
#pragma once
#include "pe_device.hpp"
#include "pusch_eq.h"

namespace pe {
namespace mmse {

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j
// Floats per subcarrier slot in LDS: Ginv (2 NT) + scales (L) + ok flag, rounded up to 4 x an odd number so
// that 16 lanes reading 16-byte vectors from consecutive slots hit distinct LDS banks.
__host__ __device__ constexpr int slot_quads(int L) { return (L * (L + 1) + L + 4) / 4; }
__host__ __device__ constexpr int slot_floats(int L) {
    return 4 * (slot_quads(L) % 2 ? slot_quads(L) : slot_quads(L) + 1);
}
// LDS bytes of a tile of T subcarriers with L layers
__host__ __device__ constexpr int tile_lds_bytes(int L, int T) { return 4 * T * slot_floats(L); }

template <int L> struct Dim {
    static constexpr int NT = L * (L + 1) / 2;  // stored entries of a Hermitian L x L (lower triangle)
    static constexpr int SLOT = slot_floats(L);
};

template <typename T> __device__ __forceinline__ T dev_sqrt(T x);
template <> __device__ __forceinline__ float dev_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double dev_sqrt<double>(double x) { return __builtin_sqrt(x); }
template <typename T> __device__ __forceinline__ T dev_rcp(T x);
template <> __device__ __forceinline__ float dev_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double dev_rcp<double>(double x) { return 1.0 / x; }
__device__ __forceinline__ float abs_(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double abs_(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---- group sums -------------------------------------------------------------------------------------------
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF,
                                                                 true));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Sum over each aligned group of P lanes (P = 4, 8, 16; P is wave-uniform), result in every lane of the
// group: xor-1 and xor-2 quad permutes, then the half-row mirror (lane i <-> 7 - i) and the row mirror
// (i <-> 15 - i), each pairing the two halves of the next larger group. a + b == b + a, so every lane ends
// with the same bits.
template <typename T> __device__ __forceinline__ T group_sum(T v, int P) {
    v += dpp<0xB1>(v);
    v += dpp<0x4E>(v);
    if (P >= 8) v += dpp<0x141>(v);
    if (P >= 16) v += dpp<0x140>(v);
    return v;
}

#pragma clang fp contract(off)

// Group sums of a Gram matrix's lower triangle (diagonal imaginary parts set to zero).
template <typename T, int L>
__device__ __forceinline__ void group_sums(T (&gr)[Dim<L>::NT], T (&gi)[Dim<L>::NT], int P) {
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            gr[tri(i, j)] = group_sum(gr[tri(i, j)], P);
            gi[tri(i, j)] = j < i ? group_sum(gi[tri(i, j)], P) : T(0);
        }
}

// ---- Gram: this lane's partial sum G[i][j] = sum_a conj(H[a][i]) H[a][j], i >= j ---------------------------
// The diagonal is real: its imaginary parts are neither accumulated nor summed (zero; the solve sets them to zero
// anyway), which leaves every other entry's bits as they are.
template <typename T, int L>
__device__ __forceinline__ void load_row(const float2 *__restrict__ row, T (&vr)[L], T (&vi)[L]) {
    if constexpr (L % 2 == 0) {
#pragma unroll
        for (int k = 0; k < L / 2; ++k) {
            const float4 q = reinterpret_cast<const float4 *>(row)[k];
            vr[2 * k] = (T)q.x;
            vi[2 * k] = (T)q.y;
            vr[2 * k + 1] = (T)q.z;
            vi[2 * k + 1] = (T)q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const float2 q = row[k];
            vr[k] = (T)q.x;
            vi[k] = (T)q.y;
        }
    }
}

// Partial Gram of KR antenna rows per lane (A <= KR P), the rows kept in registers for the coefficients.
template <int L, int KR>
__device__ __forceinline__ void gram_group_keep(const float2 *__restrict__ h, int A, int p, int P,
                                                float (&gr)[Dim<L>::NT], float (&gi)[Dim<L>::NT],
                                                float (&hr)[KR][L], float (&hi)[KR][L]) {
#pragma unroll
    for (int k = 0; k < Dim<L>::NT; ++k) gr[k] = gi[k] = 0.f;
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        const int a = p + r * P;
        if (a < A) {
            load_row<float, L>(h + (size_t)a * L, hr[r], hi[r]);
        } else {
#pragma unroll
            for (int k = 0; k < L; ++k) hr[r][k] = hi[r][k] = 0.f;
        }
    }
#pragma unroll
    for (int r = 0; r < KR; ++r) {
        if (p + r * P >= A) break;  // same rows, same order as gram_group: bit-identical sums
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                gr[tri(i, j)] = fma_(hr[r][i], hr[r][j], fma_(hi[r][i], hi[r][j], gr[tri(i, j)]));
                if (j < i) gi[tri(i, j)] = fma_(hr[r][i], hi[r][j], fma_(-hi[r][i], hr[r][j], gi[tri(i, j)]));
            }
    }
    group_sums<float, L>(gr, gi, P);
}

template <typename T, int L>
__device__ __forceinline__ void gram_group(const float2 *__restrict__ h, int A, int p, int P,
                                           T (&gr)[Dim<L>::NT], T (&gi)[Dim<L>::NT]) {
#pragma unroll
    for (int k = 0; k < Dim<L>::NT; ++k) gr[k] = gi[k] = T(0);
    // unrolled so several antenna rows are in flight at once
#pragma unroll 4
    for (int a = p; a < A; a += P) {
        T vr[L], vi[L];
        load_row<T, L>(h + (size_t)a * L, vr, vi);
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                gr[tri(i, j)] = fma_(vr[i], vr[j], fma_(vi[i], vi[j], gr[tri(i, j)]));
                if (j < i) gi[tri(i, j)] = fma_(vr[i], vi[j], fma_(-vi[i], vr[j], gi[tri(i, j)]));
            }
    }
    group_sums<T, L>(gr, gi, P);
}

// ---- the L x L algebra --------------------------------------------------------------------------------------
// 1-norm of a Hermitian matrix stored as its lower triangle.
template <typename T, int L>
__device__ __forceinline__ T herm_norm1(const T (&mr)[Dim<L>::NT], const T (&mi)[Dim<L>::NT]) {
    T mag[Dim<L>::NT];
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            const int k = tri(i, j);
            mag[k] = (i == j) ? abs_(mr[k]) : dev_sqrt<T>(fma_(mr[k], mr[k], mi[k] * mi[k]));
        }
    T nrm = T(0);
#pragma unroll
    for (int j = 0; j < L; ++j) {
        T cs = T(0);
#pragma unroll
        for (int i = 0; i < L; ++i) cs += mag[i >= j ? tri(i, j) : tri(j, i)];
        nrm = cs > nrm ? cs : nrm;
    }
    return nrm;
}

// On entry g holds the summed Gram matrix (without sigma2); on return v holds Ginv (lower triangle), scl the
// output scale of each layer and x_l = sigma2 Ginv_ll. Returns false when the factorization breaks down or,
// with check_cond, when kappa_1 exceeds thr or is not finite.
template <typename T, int L>
__device__ __forceinline__ bool solve(T (&gr)[Dim<L>::NT], T (&gi)[Dim<L>::NT], T s2, bool unbiased, T thr,
                                      bool check_cond, T (&vr)[Dim<L>::NT], T (&vi)[Dim<L>::NT], T (&scl)[L],
                                      T (&x)[L]) {
    constexpr int NT = Dim<L>::NT;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        gr[tri(i, i)] += s2;
        gi[tri(i, i)] = T(0);
    }
    T normG = check_cond ? herm_norm1<T, L>(gr, gi) : T(0);
    // pin ||G||_1 here: computed next to its use at the end, it would keep all of G alive through the solve
    asm volatile("" : "+v"(normG));

    // Cholesky in place: R (lower) overwrites G; rd[j] = 1 / R_jj
    bool ok = true;
    T rd[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        T d = gr[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) d = fma_(-gr[tri(j, k)], gr[tri(j, k)], fma_(-gi[tri(j, k)], gi[tri(j, k)], d));
        ok = ok && (d > T(0));
        const T rjj = dev_sqrt<T>(d > T(0) ? d : T(1));
        rd[j] = dev_rcp<T>(rjj);
        gr[tri(j, j)] = rjj;
#pragma unroll
        for (int i = j + 1; i < L; ++i) {
            T re = gr[tri(i, j)], im = gi[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) {  // R_ij -= R_ik conj(R_jk)
                re = fma_(-gr[tri(i, k)], gr[tri(j, k)], fma_(-gi[tri(i, k)], gi[tri(j, k)], re));
                im = fma_(-gi[tri(i, k)], gr[tri(j, k)], fma_(gr[tri(i, k)], gi[tri(j, k)], im));
            }
            gr[tri(i, j)] = re * rd[j];
            gi[tri(i, j)] = im * rd[j];
        }
    }

    // W = R^-1 (lower), from W R = I: W_jj = 1 / R_jj, W_ij = -(sum_{k=j+1}^{i} W_ik R_kj) / R_jj. Columns
    // are formed from the last to the first and rows from the bottom up, so column j of R is dead once
    // column j of W is done (R and W never need to be held at once).
    T wr[NT], wi[NT];
#pragma unroll
    for (int j = L - 1; j >= 0; --j) {
        wr[tri(j, j)] = rd[j];
        wi[tri(j, j)] = T(0);
#pragma unroll
        for (int i = L - 1; i > j; --i) {
            T re = wr[tri(i, i)] * gr[tri(i, j)], im = wr[tri(i, i)] * gi[tri(i, j)];
#pragma unroll
            for (int k = j + 1; k < i; ++k) {
                re = fma_(wr[tri(i, k)], gr[tri(k, j)], fma_(-wi[tri(i, k)], gi[tri(k, j)], re));
                im = fma_(wr[tri(i, k)], gi[tri(k, j)], fma_(wi[tri(i, k)], gr[tri(k, j)], im));
            }
            wr[tri(i, j)] = -re * rd[j];
            wi[tri(i, j)] = -im * rd[j];
        }
    }

    // Ginv = W^H W, lower triangle: V_ij = sum_{k >= i} conj(W_ki) W_kj (i >= j; W_ii is real), formed
    // column by column so that column j of W is dead once column j of V is done.
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
        for (int i = j; i < L; ++i) {
            T re = wr[tri(i, i)] * wr[tri(i, j)], im = wr[tri(i, i)] * wi[tri(i, j)];
#pragma unroll
            for (int k = i + 1; k < L; ++k) {
                re = fma_(wr[tri(k, i)], wr[tri(k, j)], fma_(wi[tri(k, i)], wi[tri(k, j)], re));
                im = fma_(wr[tri(k, i)], wi[tri(k, j)], fma_(-wi[tri(k, i)], wr[tri(k, j)], im));
            }
            vr[tri(i, j)] = re;
            vi[tri(i, j)] = (i == j) ? T(0) : im;
        }

    if (check_cond) {
        const T kappa = normG * herm_norm1<T, L>(vr, vi);
        ok = ok && (kappa <= thr);  // false for NaN
    }
#pragma unroll
    for (int l = 0; l < L; ++l) {
        x[l] = s2 * vr[tri(l, l)];
        const T mu = T(1) - x[l];
        scl[l] = (unbiased && mu != T(0)) ? dev_rcp<T>(mu) : T(1);
    }
    return ok;
}

template <typename T>
__device__ __forceinline__ float noise_inv_of(T x) {
    const T xf = x > T(1e-30) ? x : T(1e-30);
    return (float)((T(1) - x) * dev_rcp<T>(xf));
}

// ---- coefficients: C[l][a] = s_l sum_j Ginv[l][j] conj(H[a][j]) for this lane's antennas --------------------
// C[l][a] for all l from one antenna row of H, as the complex64 values stored
template <typename T, int L>
__device__ __forceinline__ void c_values(const T (&hr)[L], const T (&hi)[L], const T (&vr)[Dim<L>::NT],
                                         const T (&vi)[Dim<L>::NT], const T (&scl)[L], float (&cr)[L],
                                         float (&ci)[L]) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
        T re = T(0), im = T(0);
#pragma unroll
        for (int j = 0; j < L; ++j) {
            // Ginv[l][j]: stored lower; for j > l it is conj(V[j][l])
            const T g_r = j <= l ? vr[tri(l, j)] : vr[tri(j, l)];
            const T g_i = j <= l ? vi[tri(l, j)] : -vi[tri(j, l)];
            re = fma_(g_r, hr[j], fma_(g_i, hi[j], re));   // (g_r + i g_i)(h_r - i h_i)
            im = fma_(g_i, hr[j], fma_(-g_r, hi[j], im));
        }
        cr[l] = (float)(re * scl[l]);
        ci[l] = (float)(im * scl[l]);
    }
}

template <typename T, int L>
__device__ __forceinline__ void c_column(const T (&hr)[L], const T (&hi)[L], const T (&vr)[Dim<L>::NT],
                                         const T (&vi)[Dim<L>::NT], const T (&scl)[L], int A, int a,
                                         float2 *__restrict__ c_out, bool store) {
    float cr[L], ci[L];
    c_values<T, L>(hr, hi, vr, vi, scl, cr, ci);
    if (store) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            typedef float f2v __attribute__((ext_vector_type(2)));
            const f2v v = {cr[l], ci[l]};
            __builtin_nontemporal_store(v, reinterpret_cast<f2v *>(c_out) + (size_t)l * A + a);
        }
    }
}

// Binary16 coefficient output (C16 pairs and the subcarrier's exponent, rule in c16_exp) by a group of Q
// lanes (lane q: antennas q, q + Q, ..., A <= NA Q): the block's values stay in registers until the group's
// maximum is known. Every lane of the wave calls it (the maximum is a butterfly over lane bits below Q).
template <int L, int NA>
__device__ __forceinline__ void write_c16_group(const float2 *__restrict__ h, int A, int q, int Q,
                                                const float (&vr)[Dim<L>::NT], const float (&vi)[Dim<L>::NT],
                                                const float (&scl)[L], uint32_t *__restrict__ c16,
                                                int *__restrict__ cexp, bool store) {
    float cr[NA][L], ci[NA][L];
    float m = 0.f;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = q + u * Q;
        if (a < A) {
            float hr[L], hi[L];
            load_row<float, L>(h + (size_t)a * L, hr, hi);
            c_values<float, L>(hr, hi, vr, vi, scl, cr[u], ci[u]);
#pragma unroll
            for (int l = 0; l < L; ++l) m = fmaxf(m, fmaxf(abs_(cr[u][l]), abs_(ci[u][l])));
        }
    }
    const int e = c16_exp(wave_max_xor(m, 1, Q));
    const float s = ldexpf(1.f, -e);
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = q + u * Q;
        if (store && a < A) {
#pragma unroll
            for (int l = 0; l < L; ++l)
                __builtin_nontemporal_store(pack_half2(cr[u][l] * s, ci[u][l] * s), c16 + (size_t)l * A + a);
        }
    }
    if (store && q == 0) *cexp = e;
}

template <typename T, int L>
__device__ __forceinline__ void write_c(const float2 *__restrict__ h, int A, int p, int P,
                                        const T (&vr)[Dim<L>::NT], const T (&vi)[Dim<L>::NT], const T (&scl)[L],
                                        float2 *__restrict__ c_out, bool store = true) {
#pragma unroll 2
    for (int a = p; a < A; a += P) {
        T hr[L], hi[L];
        load_row<T, L>(h + (size_t)a * L, hr, hi);
        c_column<T, L>(hr, hi, vr, vi, scl, A, a, c_out, store);
    }
}

// ---- FP64 solve of one flagged subcarrier by one wave ---------------------------------------------------------
// Lane 8 i + j holds element (i, j) of the L x L matrices (lower triangle, i >= j, in use) as complex doubles; values
// of other lanes are read by shuffles. G = H^H H + sigma2 I (antenna-serial sums), right-looking Cholesky G = R R^H,
// W = R^-1 row by row, Ginv = W^H W, then one lane per antenna forms C[l][a] = s_l sum_j Ginv[l][j] conj(H[a][j]).
// Loops are kept rolled (this path is rare): about 60 registers, so it runs inside the FP32 kernels for the
// subcarriers they flag (no separate launch) without raising their register count much.
// Writes C (or the binary16 block and its exponent), noise_inv and the final flag (FP64, plus SINGULAR with zero C
// and noise_inv when the factorization breaks down). Every lane of the wave calls it.
__device__ __forceinline__ double shfl_d(double v, int src) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__shfl((int)(unsigned)u, src, 64);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)(u >> 32), src, 64);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int L, bool C16>
__device__ __forceinline__ void wave_fp64(const float2 *__restrict__ h, float s2f, float2 *__restrict__ c,
                                                    float *__restrict__ ni, uint8_t *__restrict__ flag, int A,
                                                    bool unbiased, int lane, uint32_t *__restrict__ c16,
                                                    int *__restrict__ cexp) {
    const int i = lane >> 3, j = lane & 7;
    const bool low = i < L && j <= i;  // this lane's element is in the lower triangle
    double rr = 0.0, ri = 0.0;
    if (low) {
#pragma unroll 1
        for (int a = 0; a < A; ++a) {
            const float2 hi_ = h[a * L + i], hj = h[a * L + j];
            rr = fma_((double)hi_.x, (double)hj.x, fma_((double)hi_.y, (double)hj.y, rr));
            ri = fma_((double)hi_.x, (double)hj.y, fma_(-(double)hi_.y, (double)hj.x, ri));
        }
    }
    if (i == j) {
        rr += (double)s2f;
        ri = 0.0;
    }
    // Cholesky, right-looking: column k scaled by 1 / R_kk, then the trailing lower triangle updated
    bool ok = true;
#pragma unroll 1
    for (int k = 0; k < L; ++k) {
        const double d = shfl_d(rr, 9 * k);
        if (!(d > 0.0)) {  // uniform
            ok = false;
            break;
        }
        const double rk = dev_sqrt<double>(d), rinv = 1.0 / rk;
        if (j == k && i == k) {
            rr = rk;
            ri = 0.0;
        } else if (j == k && i > k && i < L) {
            rr *= rinv;
            ri *= rinv;
        }
        const double aik_r = shfl_d(rr, 8 * i + k), aik_i = shfl_d(ri, 8 * i + k);  // R[i][k]
        const double ajk_r = shfl_d(rr, 8 * j + k), ajk_i = shfl_d(ri, 8 * j + k);  // R[j][k]
        if (low && j > k) {  // G[i][j] -= R[i][k] conj(R[j][k])
            rr = fma_(-aik_r, ajk_r, fma_(-aik_i, ajk_i, rr));
            ri = fma_(-aik_i, ajk_r, fma_(aik_r, ajk_i, ri));
        }
    }
    double scale = 1.0, x = 0.0, vr = 0.0, vi = 0.0;
    if (ok) {
        // W = R^-1: W_ii = 1 / R_ii, W_ij = -(sum_{k=j}^{i-1} R_ik W_kj) / R_ii, one row at a time
        const double rdiag = shfl_d(rr, 9 * (i < L ? i : 0));
        double wr = (i == j && i < L) ? 1.0 / rdiag : 0.0, wi = 0.0;
#pragma unroll 1
        for (int row = 1; row < L; ++row) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < row; ++k) {
                const double Rr = shfl_d(rr, 8 * row + k), Ri = shfl_d(ri, 8 * row + k);  // R[row][k]
                const double Wr = shfl_d(wr, 8 * k + j), Wi = shfl_d(wi, 8 * k + j);      // W[k][j]
                if (j <= k) {
                    sr = fma_(Rr, Wr, fma_(-Ri, Wi, sr));
                    si = fma_(Rr, Wi, fma_(Ri, Wr, si));
                }
            }
            if (i == row && j < row) {
                const double dinv = 1.0 / rdiag;
                wr = -sr * dinv;
                wi = -si * dinv;
            }
        }
        // Ginv = W^H W: V_ij = sum_{k >= i} conj(W_ki) W_kj
#pragma unroll 1
        for (int k = 0; k < L; ++k) {
            const double a_r = shfl_d(wr, 8 * k + i), a_i = shfl_d(wi, 8 * k + i);  // W[k][i]
            const double b_r = shfl_d(wr, 8 * k + j), b_i = shfl_d(wi, 8 * k + j);  // W[k][j]
            if (low && k >= i) {
                vr = fma_(a_r, b_r, fma_(a_i, b_i, vr));
                vi = fma_(a_r, b_i, fma_(-a_i, b_r, vi));
            }
        }
        if (i == j) vi = 0.0;  // Ginv is Hermitian: a real diagonal
        if (i == j && i < L) {
            x = (double)s2f * vr;
            const double mu = 1.0 - x;
            scale = (unbiased && mu != 0.0) ? 1.0 / mu : 1.0;
            ni[i] = noise_inv_of(x);
        }
    } else if (i == j && i < L) {
        ni[i] = 0.f;
    }
    // coefficients: one lane per antenna (a = lane, lane + 64, ...), one layer row at a time (a rolled loop, so
    // that only that row of Ginv is in registers); C16: a first sweep for the block maximum, a second to store
    auto coef = [&](int l, const float (&hr)[L], const float (&hi)[L], float &cr, float &ci) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int jj = 0; jj < L; ++jj) {  // Ginv[l][jj]: V[l][jj] for jj <= l, else conj(V[jj][l])
            const int src = jj <= l ? 8 * l + jj : 8 * jj + l;
            const double g_r = shfl_d(vr, src), g_i0 = shfl_d(vi, src);
            const double g_i = jj <= l ? g_i0 : -g_i0;
            re = fma_(g_r, (double)hr[jj], fma_(g_i, (double)hi[jj], re));
            im = fma_(g_i, (double)hr[jj], fma_(-g_r, (double)hi[jj], im));
        }
        const double s = shfl_d(scale, 9 * l);
        cr = ok ? (float)(re * s) : 0.f;
        ci = ok ? (float)(im * s) : 0.f;
    };
    auto row = [&](int a, float (&hr)[L], float (&hi)[L]) {
#pragma unroll
        for (int jj = 0; jj < L; ++jj) {
            const float2 v = a < A ? h[a * L + jj] : make_float2(0.f, 0.f);
            hr[jj] = v.x;
            hi[jj] = v.y;
        }
    };
    if constexpr (!C16) {
        for (int a = lane; a < A; a += 64) {
            float hr[L], hi[L];
            row(a, hr, hi);
#pragma unroll 1
            for (int l = 0; l < L; ++l) {
                float cr, ci;
                coef(l, hr, hi, cr, ci);
                c[(size_t)l * A + a] = make_float2(cr, ci);
            }
        }
    } else {  // A <= 128
        float m = 0.f;
        for (int a = lane; a < A; a += 64) {
            float hr[L], hi[L];
            row(a, hr, hi);
#pragma unroll 1
            for (int l = 0; l < L; ++l) {
                float cr, ci;
                coef(l, hr, hi, cr, ci);
                m = fmaxf(m, fmaxf(abs_(cr), abs_(ci)));
            }
        }
        const int e = c16_exp(wave_max_xor(m, 1, 64));
        const float sc = ldexpf(1.f, -e);
        for (int a = lane; a < A; a += 64) {
            float hr[L], hi[L];
            row(a, hr, hi);
#pragma unroll 1
            for (int l = 0; l < L; ++l) {
                float cr, ci;
                coef(l, hr, hi, cr, ci);
                c16[(size_t)l * A + a] = pack_half2(cr * sc, ci * sc);
            }
        }
        if (lane == 0) *cexp = e;
    }
    if (lane == 0) *flag = ok ? PE_SCFLAG_FP64 : (PE_SCFLAG_FP64 | PE_SCFLAG_SINGULAR);
}

// The same as a called function: in the multi-cell kernels, whose registers are set by the largest layer count,
// a call keeps the kernel's register count where it is without the FP64 code (the call itself is only made for a
// flagged subcarrier).
template <int L, bool C16>
__device__ __noinline__ void wave_fp64_call(const float2 *__restrict__ h, float s2f, float2 *__restrict__ c,
                                            float *__restrict__ ni, uint8_t *__restrict__ flag, int A, bool unbiased,
                                            int lane, uint32_t *__restrict__ c16, int *__restrict__ cexp) {
    wave_fp64<L, C16>(h, s2f, c, ni, flag, A, unbiased, lane, c16, cexp);
}

// The FP64 solves of a tile's flagged subcarriers (those whose slot ok flag is 0), spread over the workgroup's W
// waves; called by every thread after the tile's algebra, while the slots are intact. CALL: through
// wave_fp64_call.
template <int L, bool C16 = false, bool CALL = false>
__device__ __forceinline__ void tile_fp64_inline(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                                 float2 *__restrict__ C, float *__restrict__ noise_inv,
                                                 uint8_t *__restrict__ flags, long sc0, int n, int A, bool unbiased,
                                                 const float *__restrict__ lds, int tid, int W,
                                                 uint32_t *__restrict__ c16 = nullptr, int *__restrict__ cexp = nullptr) {
    constexpr int NT = Dim<L>::NT;
    constexpr int SLOT = Dim<L>::SLOT;
    const int lane = tid & 63, wave = tid >> 6;
    int idx = 0;
    for (int t0 = 0; t0 < n; t0 += 64) {
        const bool f = t0 + lane < n && lds[(t0 + lane) * SLOT + 2 * NT + L] == 0.f;
        unsigned long long m = __ballot(f);  // the same in every wave
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            if (idx++ % W != wave) continue;
            const long sc = sc0 + t0 + b;
            float2 *c = C ? C + (size_t)sc * L * A : nullptr;
            uint32_t *c16s = c16 ? c16 + (size_t)sc * L * A : nullptr;
            int *ce = cexp ? cexp + sc : nullptr;
            if constexpr (CALL)
                wave_fp64_call<L, C16>(H + (size_t)sc * A * L, sigma2[sc], c, noise_inv + (size_t)sc * L, flags + sc,
                                       A, unbiased, lane, c16s, ce);
            else
                wave_fp64<L, C16>(H + (size_t)sc * A * L, sigma2[sc], c, noise_inv + (size_t)sc * L, flags + sc, A,
                                  unbiased, lane, c16s, ce);
        }
    }
}

// ---- LDS slot access (16-byte vectors; SLOT is a multiple of 4 floats, slots 16-byte aligned) ----------------
template <int K>
__device__ __forceinline__ void slot_store(float *__restrict__ s, const float (&v)[K]) {
#pragma unroll
    for (int k = 0; k + 3 < K; k += 4)
        *reinterpret_cast<float4 *>(s + k) = make_float4(v[k], v[k + 1], v[k + 2], v[k + 3]);
#pragma unroll
    for (int k = K / 4 * 4; k < K; ++k) s[k] = v[k];
}
template <int K>
__device__ __forceinline__ void slot_load(const float *__restrict__ s, float (&v)[K]) {
#pragma unroll
    for (int k = 0; k + 3 < K; k += 4) {
        const float4 q = *reinterpret_cast<const float4 *>(s + k);
        v[k] = q.x;
        v[k + 1] = q.y;
        v[k + 2] = q.z;
        v[k + 3] = q.w;
    }
#pragma unroll
    for (int k = K / 4 * 4; k < K; ++k) v[k] = s[k];
}

// Orders the workgroup's LDS accesses between the phases of a tile (one-wave workgroups need no barrier).
__device__ __forceinline__ void tile_barrier(int W) {
    if (W > 1)
        __syncthreads();
    else
        lds_order_wave();
}

// Stores the first K floats of v at dst, 16 bytes at a time when dst is 16-byte aligned (a16) and K is a multiple
// of 4, else 8 bytes at a time (dst is always 8-byte aligned here) and a last float for odd K.
template <int K, int N>
__device__ __forceinline__ void store_floats(float *__restrict__ dst, const float (&v)[N], bool a16) {
    static_assert(K <= N, "short vector");
    typedef float f4v __attribute__((ext_vector_type(4)));
    typedef float f2v __attribute__((ext_vector_type(2)));
    if (K % 4 == 0 && a16) {
#pragma unroll
        for (int k = 0; k < K / 4; ++k) {
            const f4v q = {v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
            reinterpret_cast<f4v *>(dst)[k] = q;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K / 2; ++k) {
            const f2v q = {v[2 * k], v[2 * k + 1]};
            reinterpret_cast<f2v *>(dst)[k] = q;
        }
        if (K % 2) dst[K - 1] = v[K - 1];
    }
}

// ---- FP32 pass over one tile: subcarriers sc0 .. sc0 + n - 1 of one cell, by a workgroup of W waves -------
// Phases 1 and 2 of a tile (Gram, algebra), ending at a barrier with Ginv, the scales and the ok flag of
// every subcarrier in its LDS slot; noise_inv and flags (0) written for the subcarriers solved in FP32, the
// flagged ones left to tile_fp64_inline. keep: the lane's KR1 antenna rows stay in kr / ki (single Gram round
// only). REC: the solve is kept as a record instead (Ginv lower triangle, real parts then imaginary parts, then
// the L scales: the slot's first 2 NT + L floats) at the start of the subcarrier's coefficient block in C, stored
// by the lane that solved it; the slot then holds only the ok flag (flagged subcarriers get no record: their block
// receives the FP64 coefficients).
template <int L, int KR1, bool REC = false>
__device__ __forceinline__ void tile_solve(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                           float *__restrict__ noise_inv, uint8_t *__restrict__ flags, long sc0,
                                           int n, int A, int P, bool unbiased, float thr, float *__restrict__ lds,
                                           int tid, int W, bool keep, float (&kr)[KR1][L],
                                           float (&ki)[KR1][L], float2 *__restrict__ C = nullptr) {
    constexpr int NT = Dim<L>::NT;
    constexpr int SLOT = Dim<L>::SLOT;
    const int lane = tid & 63, wave = tid >> 6;
    const int lgP = __builtin_ctz((unsigned)P);
    const int p = lane & (P - 1), g = lane >> lgP, G = 64 >> lgP;

    if (keep) {
        const int t = min(wave * G + g, n - 1);
        float gr[NT], gi[NT];
        gram_group_keep<L, KR1>(H + (size_t)(sc0 + t) * A * L, A, p, P, gr, gi, kr, ki);
        if (p == 0 && wave * G + g < n) {
            float buf[2 * NT];
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                buf[k] = gr[k];
                buf[NT + k] = gi[k];
            }
            slot_store<2 * NT>(lds + t * SLOT, buf);
        }
    }
    // Gram, 64 / P subcarriers per wave round (lanes past n repeat the last subcarrier and store nothing)
    for (int t0 = wave * G; !keep && t0 < n; t0 += W * G) {
        const int t = min(t0 + g, n - 1);
        float gr[NT], gi[NT];
        gram_group<float, L>(H + (size_t)(sc0 + t) * A * L, A, p, P, gr, gi);
        if (p == 0 && t0 + g < n) {
            float buf[2 * NT];
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                buf[k] = gr[k];
                buf[NT + k] = gi[k];
            }
            slot_store<2 * NT>(lds + t * SLOT, buf);
        }
    }
    tile_barrier(W);

    // algebra, one thread per subcarrier
    if (tid < n) {
        const long sc = sc0 + tid;
        float buf[2 * NT];
        slot_load<2 * NT>(lds + tid * SLOT, buf);
        float gr[NT], gi[NT], vr[NT], vi[NT], scl[L], x[L];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            gr[k] = buf[k];
            gi[k] = buf[NT + k];
        }
        const bool ok = solve<float, L>(gr, gi, sigma2[sc], unbiased, thr, true, vr, vi, scl, x);
        float out[2 * NT + L + 1];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            out[k] = vr[k];
            out[NT + k] = vi[k];
        }
#pragma unroll
        for (int l = 0; l < L; ++l) out[2 * NT + l] = scl[l];
        out[2 * NT + L] = ok ? 1.f : 0.f;
        if constexpr (REC)
            lds[tid * SLOT + 2 * NT + L] = out[2 * NT + L];
        else
            slot_store<2 * NT + L + 1>(lds + tid * SLOT, out);
        if (ok) {
#pragma unroll
            for (int l = 0; l < L; ++l) noise_inv[sc * L + l] = noise_inv_of(x[l]);
            flags[sc] = 0;
            if constexpr (REC)  // the block starts 16-byte aligned when L A is even
                store_floats<2 * NT + L>(reinterpret_cast<float *>(C + (size_t)sc * L * A), out, (L * A) % 2 == 0);
        }
    }
    tile_barrier(W);
}

// n <= 64 W; lds: the workgroup's region of at least n * Dim<L>::SLOT floats; tid = thread in the workgroup.
// Gram and coefficient rounds are spread over the W waves; the algebra runs on threads 0 .. n - 1, i.e. on as
// few waves as possible, while the others wait at the barrier and leave their issue slots to other
// workgroups. Flagged subcarriers are then solved in FP64 by the workgroup's waves (tile_fp64_inline), which write
// their C, noise_inv and final flags; the tile ends at a barrier, so a workgroup may go on to another tile.
// KR > 0: when every wave has a single Gram round (n <= 64 W / P) and A <= KR P, each lane keeps its KR antenna
// rows in registers from the Gram to the coefficients instead of reading H a second time (the coefficient
// stores then cover P consecutive antennas of a C row per lane group). It trades the second read of H for
// registers (fewer waves per SIMD); which is faster depends on the shape.
// C16: coefficients go to c16 / cexp as binary16 pairs and exponents (write_c16_group with NA antennas per
// lane: A <= 64 NA; no kept rows) instead of complex64 to C. CALL: the FP64 solves through a function call.
template <int L, int KR, bool C16 = false, int NA = 2, bool CALL = false>
__device__ __forceinline__ void tile_fp32(const float2 *__restrict__ H, const float *__restrict__ sigma2,
                                          float2 *__restrict__ C, float *__restrict__ noise_inv,
                                          uint8_t *__restrict__ flags, long sc0, int n, int A, int P,
                                          bool unbiased, float thr, float *__restrict__ lds, int tid, int W,
                                          uint32_t *__restrict__ c16 = nullptr,
                                          int *__restrict__ cexp = nullptr) {
    static_assert(!C16 || KR == 0, "binary16 output reads H again for the coefficients");
    constexpr int NT = Dim<L>::NT;
    constexpr int SLOT = Dim<L>::SLOT;
    const int lane = tid & 63, wave = tid >> 6;
    const int lgP = __builtin_ctz((unsigned)P);
    const int p = lane & (P - 1), g = lane >> lgP, G = 64 >> lgP;
    const bool keep = KR > 0 && A <= KR * P && n <= W * G;
    constexpr int KR1 = KR > 0 ? KR : 1;
    float kr[KR1][L], ki[KR1][L];
    tile_solve<L, KR1>(H, sigma2, noise_inv, flags, sc0, n, A, P, unbiased, thr, lds, tid, W, keep, kr, ki);

    // coefficients: Q = min(64, 2^ceil(log2 A)) lanes per subcarrier, lane q writing antennas q, q + Q, ..., so
    // a store instruction covers up to 512 contiguous bytes of a C row whatever P is (C[l][a] is computed by
    // one lane in a fixed order either way). Lanes past n repeat the last subcarrier and store nothing; two
    // rounds per iteration keep two subcarriers' loads in flight.
    if (keep) {
        const int t = min(wave * G + g, n - 1);
        float in[2 * NT + L + 1];
        slot_load<2 * NT + L + 1>(lds + t * SLOT, in);
        const bool store = wave * G + g < n && in[2 * NT + L] != 0.f;  // flagged: solved in FP64 below
        float vr[NT], vi[NT], scl[L];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            vr[k] = in[k];
            vi[k] = in[NT + k];
        }
#pragma unroll
        for (int l = 0; l < L; ++l) scl[l] = in[2 * NT + l];
#pragma unroll
        for (int r = 0; r < KR1; ++r) {
            const int a = p + r * P;
            if (a < A) c_column<float, L>(kr[r], ki[r], vr, vi, scl, A, a, C + (size_t)(sc0 + t) * L * A, store);
        }
    }
    const int lgQ = A >= 64 ? 6 : A <= 1 ? 0 : 32 - __builtin_clz((unsigned)A - 1u);
    const int q = lane & ((1 << lgQ) - 1), g3 = lane >> lgQ, R3 = 64 >> lgQ;
    auto coef_round = [&](int t0) {
        const int t = min(t0 + g3, n - 1);
        float in[2 * NT + L + 1];
        slot_load<2 * NT + L + 1>(lds + t * SLOT, in);
        const bool store = t0 + g3 < n && in[2 * NT + L] != 0.f;  // flagged: solved in FP64 below
        float vr[NT], vi[NT], scl[L];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            vr[k] = in[k];
            vi[k] = in[NT + k];
        }
#pragma unroll
        for (int l = 0; l < L; ++l) scl[l] = in[2 * NT + l];
        if constexpr (C16)
            write_c16_group<L, NA>(H + (size_t)(sc0 + t) * A * L, A, q, 1 << lgQ, vr, vi, scl,
                                   c16 + (size_t)(sc0 + t) * L * A, cexp + sc0 + t, store);
        else
            write_c<float, L>(H + (size_t)(sc0 + t) * A * L, A, q, 1 << lgQ, vr, vi, scl,
                              C + (size_t)(sc0 + t) * L * A, store);
    };
    if constexpr (C16) {  // the group maximum is a cross-lane butterfly, which the unroller does not duplicate
        for (int t0 = wave * R3; t0 < n; t0 += W * R3) coef_round(t0);
    } else {
#pragma unroll 2
        for (int t0 = wave * R3; !keep && t0 < n; t0 += W * R3) coef_round(t0);
    }
    // subcarriers the FP32 pass flagged, solved in FP64 by this workgroup
    tile_fp64_inline<L, C16, CALL>(H, sigma2, C, noise_inv, flags, sc0, n, A, unbiased, lds, tid, W, c16, cexp);
    tile_barrier(W);  // slot reads done before the workgroup's next tile rewrites them
}

#pragma clang fp contract(fast)  // the HIP default again for code after this header

}  // namespace mmse
}  // namespace pe
