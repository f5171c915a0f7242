// mmse_device.hpp -- per-subcarrier MMSE solve as a device function, shared by the single-cell kernels
// (gram_mmse.hip) and the multi-cell persistent kernels (multicell.hip).
//
// The reference only names this stage (eqMmseCoefCompHighMimoKernel, aerial_bottleneck.cu:51-53); the
// arithmetic here is the regularized solve that defines C in the reference's headline workload.
//
// Mapping: a group of P consecutive lanes owns one subcarrier (64/P subcarriers per wave). Lane p of a
// group reads antenna rows a = p, p+P, ... of H[sc] (L complex each, contiguous, so the group's loads are
// contiguous) and accumulates the lower triangle of the Hermitian Gram; a butterfly of xor-shuffles inside
// the group sums the P partial Grams, after which every lane of the group holds the same G. The L x L
// algebra (Cholesky G = L L^H, W = L^-1, Ginv = W^H W, 1-norm condition number) then runs redundantly in
// each lane's registers with fully unrolled static indexing, and each lane writes its own antennas' columns
// of C = diag(s) Ginv H^H.
//
// Precision: pass 1 runs in FP32 and flags subcarriers whose kappa_1(G) exceeds the threshold or whose
// Cholesky breaks down; pass 2 (same code in FP64) recomputes only the flagged subcarriers.
#pragma once
#include "pe_device.hpp"
#include "pusch_eq.h"

namespace pe {
namespace mmse {

constexpr int kMmseP = 8;              // lanes per subcarrier
constexpr uint8_t kPendingFp64 = 0x80;  // internal flag: FP32 pass handed the subcarrier to the FP64 pass

__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x);
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return sqrtf(x); }
template <>
__device__ __forceinline__ double dev_sqrt<double>(double x) { return sqrt(x); }

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int m) { return __shfl_xor(v, m, 64); }

// Solves one subcarrier. T = float (pass 1) or double (pass 2). Returns false when the factorization or
// the condition test fails (pass 1: the subcarrier is handed to pass 2).
template <typename T, int L, int P>
__device__ __forceinline__ bool mmse_one(const float2 *__restrict__ h, T s2, int A, int p, bool unbiased,
                                         T thr, bool check_cond, float2 *__restrict__ c_out,
                                         float *__restrict__ ni_out) {
    constexpr int NT = L * (L + 1) / 2;
    T gr[NT], gi[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) gr[k] = gi[k] = T(0);

    // partial Gram over this lane's antennas: G[i][j] = sum_a conj(H[a][i]) H[a][j], i >= j
    // (unrolled so several antenna rows are in flight at once instead of one memory latency per row)
#pragma unroll 4
    for (int a = p; a < A; a += P) {
        T vr[L], vi[L];
        const float2 *row = h + (size_t)a * L;
        if constexpr (L % 2 == 0) {
#pragma unroll
            for (int k = 0; k < L / 2; ++k) {
                const float4 q = reinterpret_cast<const float4 *>(row)[k];
                vr[2 * k] = q.x; vi[2 * k] = q.y; vr[2 * k + 1] = q.z; vi[2 * k + 1] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < L; ++k) {
                const float2 q = row[k];
                vr[k] = q.x; vi[k] = q.y;
            }
        }
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                gr[tri(i, j)] += vr[i] * vr[j] + vi[i] * vi[j];
                gi[tri(i, j)] += vr[i] * vi[j] - vi[i] * vr[j];
            }
    }
#pragma unroll
    for (int m = 1; m < P; m <<= 1)
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            gr[k] += shfl_xor_t(gr[k], m);
            gi[k] += shfl_xor_t(gi[k], m);
        }
#pragma unroll
    for (int i = 0; i < L; ++i) {
        gr[tri(i, i)] += s2;
        gi[tri(i, i)] = T(0);
    }

    // ||G||_1 (Hermitian: column j sums |G[i][j]| over the stored lower triangle and its mirror)
    T normG = T(0);
#pragma unroll
    for (int j = 0; j < L; ++j) {
        T cs = T(0);
#pragma unroll
        for (int i = 0; i < L; ++i) {
            const int k = i >= j ? tri(i, j) : tri(j, i);
            cs += dev_sqrt<T>(gr[k] * gr[k] + gi[k] * gi[k]);
        }
        normG = cs > normG ? cs : normG;
    }

    // Cholesky in place: lower L overwrites the lower triangle of G
    bool ok = true;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        T d = gr[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= gr[tri(j, k)] * gr[tri(j, k)] + gi[tri(j, k)] * gi[tri(j, k)];
        ok = ok && (d > T(0));
        const T ljj = dev_sqrt<T>(d > T(0) ? d : T(1));
        const T inv = T(1) / ljj;
        gr[tri(j, j)] = ljj;
#pragma unroll
        for (int i = j + 1; i < L; ++i) {
            T re = gr[tri(i, j)], im = gi[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) {
                re -= gr[tri(i, k)] * gr[tri(j, k)] + gi[tri(i, k)] * gi[tri(j, k)];
                im -= gi[tri(i, k)] * gr[tri(j, k)] - gr[tri(i, k)] * gi[tri(j, k)];
            }
            gr[tri(i, j)] = re * inv;
            gi[tri(i, j)] = im * inv;
        }
    }

    // W = L^-1 (lower)
    T wr[NT], wi[NT];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        wr[tri(j, j)] = T(1) / gr[tri(j, j)];
        wi[tri(j, j)] = T(0);
    }
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
        for (int i = j + 1; i < L; ++i) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int k = j; k < i; ++k) {
                re += gr[tri(i, k)] * wr[tri(k, j)] - gi[tri(i, k)] * wi[tri(k, j)];
                im += gr[tri(i, k)] * wi[tri(k, j)] + gi[tri(i, k)] * wr[tri(k, j)];
            }
            wr[tri(i, j)] = -re * wr[tri(i, i)];
            wi[tri(i, j)] = -im * wr[tri(i, i)];
        }

    // Ginv = W^H W, lower triangle: V[i][j] = sum_{k >= i} conj(W[k][i]) W[k][j]   (i >= j)
    T vr_[NT], vi_[NT];
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int k = i; k < L; ++k) {
                re += wr[tri(k, i)] * wr[tri(k, j)] + wi[tri(k, i)] * wi[tri(k, j)];
                im += wr[tri(k, i)] * wi[tri(k, j)] - wi[tri(k, i)] * wr[tri(k, j)];
            }
            vr_[tri(i, j)] = re;
            vi_[tri(i, j)] = (i == j) ? T(0) : im;
        }

    if (check_cond) {
        T normV = T(0);
#pragma unroll
        for (int j = 0; j < L; ++j) {
            T cs = T(0);
#pragma unroll
            for (int i = 0; i < L; ++i) {
                const int k = i >= j ? tri(i, j) : tri(j, i);
                cs += dev_sqrt<T>(vr_[k] * vr_[k] + vi_[k] * vi_[k]);
            }
            normV = cs > normV ? cs : normV;
        }
        const T kappa = normG * normV;
        ok = ok && (kappa <= thr);  // false for NaN
    }
    if (!ok) return false;

    T scl[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const T x = s2 * vr_[tri(l, l)];
        const T mu = T(1) - x;
        scl[l] = (unbiased && mu != T(0)) ? T(1) / mu : T(1);
        if (p == (l % P)) {
            const T xf = x > T(1e-30) ? x : T(1e-30);
            ni_out[l] = (float)((T(1) - x) / xf);
        }
    }

    // C[l][a] = s_l * sum_j Ginv[l][j] conj(H[a][j])
#pragma unroll 4
    for (int a = p; a < A; a += P) {
        T hr[L], hi[L];
        const float2 *row = h + (size_t)a * L;
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const float2 q = row[k];
            hr[k] = q.x;
            hi[k] = q.y;
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int j = 0; j < L; ++j) {
                // Ginv[l][j]: stored lower; for j > l use conj(V[j][l])
                const T gr_ = j <= l ? vr_[tri(l, j)] : vr_[tri(j, l)];
                const T gi_ = j <= l ? vi_[tri(l, j)] : -vi_[tri(j, l)];
                // (gr_ + i gi_) * (hr - i hi)
                re += gr_ * hr[j] + gi_ * hi[j];
                im += gi_ * hr[j] - gr_ * hi[j];
            }
            c_out[(size_t)l * A + a] = make_float2((float)(re * scl[l]), (float)(im * scl[l]));
        }
    }
    return true;
}

}  // namespace mmse
}  // namespace pe
