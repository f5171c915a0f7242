/*
 * pe_oracle.c -- CPU oracle for the PUSCH equalization path (plain C, gcc).
 *
 * TEST INFRASTRUCTURE ONLY. Loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * as the checker or the timed CPU baseline; never linked into or called by the product library.
 *
 * Build: oracle/Makefile (gcc -O2 -mfma -ffp-contract=off -fopenmp). -ffp-contract=off keeps every
 * multiply and add that is written separately rounded separately; fused multiply-adds appear only where
 * fmaf()/fma() are written explicitly.
 *
 * Reference files restated (paths relative to the reference repository):
 *   src/aerial_coef_wmma.cu   cma() :60-64, coef_apply_v1_scalar :71-104, inputs :383-387
 *   src/coef_rigor.cu         inputs :101-104
 *   src/aerial_bottleneck.cu  gram_scalar :56-80
 *   src/llr/aerial_llr_bench_v2.cu  LUT :47-64, cpu_computePamLlr_ref :72-107, llr_v1_baseline :114-156,
 *                                   inputs :289-304
 *   src/llr/aerial_llr_v3.cu  cpu_ref :61-79, inputs :268-273
 *   src/srs/srs_wmma_bench.cu complex_conjmul :43-47, hcmadd :48-52, srs_filter_v1_scalar :57-90,
 *                             inputs :262-268
 *   fronthaul block floating point: O-RAN WG4 CUS-plane format, cited only (README.md:214)
 *   DMRS frequency interpolation GEMM: named only (README.md:200); cma() order of aerial_coef_wmma.cu:60-64
 */
#include "pe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------------------------------------
 * generators
 * ------------------------------------------------------------------------------------------------------ */

/* One glibc rand() stream: out[i] = ((float)rand()/RAND_MAX - 0.5f) [* scale].
 * aerial_coef_wmma.cu:383-387 (srand(42), no scale; C then Y from one stream, re before im) and
 * coef_rigor.cu:101-104 / coef_scaling.cu:182-185 (scale 1.4f). Callers ask for |C| + |Y| floats in one call
 * and split, which reproduces "all of C first, then all of Y". */
void po_glibc_uniform(unsigned seed, float scale, int apply_scale, float *out, size_t nfloats) {
    srand(seed);
    for (size_t i = 0; i < nfloats; i++) {
        float v = (float)rand() / RAND_MAX - 0.5f;
        out[i] = apply_scale ? v * scale : v;
    }
}

/* aerial_llr_bench_v2.cu:289-304 and aerial_llr_v3.cu:268-273: per symbol re, im = U(+-0.7) rounded to
 * binary16, then noiseInv = powf(10, (rand() % 26) / 10) * 2. */
void po_llr_bench_inputs(unsigned seed, int N, uint16_t *soft_half, float *noise_inv) {
    srand(seed);
    for (int i = 0; i < N; i++) {
        float re = ((float)rand() / RAND_MAX - 0.5f) * 1.4f;
        float im = ((float)rand() / RAND_MAX - 0.5f) * 1.4f;
        soft_half[2 * i] = po_f32_to_f16(re);
        soft_half[2 * i + 1] = po_f32_to_f16(im);
        float snr_lin = powf(10.f, (float)(rand() % 26) / 10.f);
        noise_inv[i] = snr_lin * 2.f;
    }
}

/* ------------------------------------------------------------------------------------------------------
 * IEEE binary16, round to nearest even (the rounding of __float2half)
 * ------------------------------------------------------------------------------------------------------ */
uint16_t po_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) { /* inf / nan */
        if (ax == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (ax < 0x38800000u) {                                    /* below 2^-14: subnormal half */
        float a;
        memcpy(&a, &ax, 4);
        float m = nearbyintf(a * 16777216.0f); /* exact scaling by 2^24, then RNE to an integer */
        return (uint16_t)(sign | (uint32_t)m);
    }
    uint32_t mant = ax & 0x7fffffu;
    uint32_t e = (ax >> 23) - 127u + 15u;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float po_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    float f;
    if (e == 0) {
        f = (float)m * (1.0f / 16777216.0f);
        uint32_t b;
        memcpy(&b, &f, 4);
        b |= sign;
        memcpy(&f, &b, 4);
        return f;
    }
    uint32_t b = (e == 31) ? (sign | 0x7f800000u | (m << 13)) : (sign | ((e - 15u + 127u) << 23) | (m << 13));
    memcpy(&f, &b, 4);
    return f;
}

void po_f32_to_f16_array(const float *in, uint16_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = po_f32_to_f16(in[i]);
}
void po_f16_to_f32_array(const uint16_t *in, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = po_f16_to_f32(in[i]);
}

/* ------------------------------------------------------------------------------------------------------
 * coefficient apply  Y_eq[sc][l][s] = sum_a C[sc][l][a] * Y[sc][s][a]
 * ------------------------------------------------------------------------------------------------------ */

/* coef_apply_v1_scalar, aerial_coef_wmma.cu:71-104: antenna-serial chain a = 0..A-1 of
 * cma(a, b, acc) = {acc.re + a.re*b.re - a.im*b.im, acc.im + a.re*b.im + a.im*b.re} (:60-64), written in
 * the contracted form a CUDA compiler emits for that expression:
 *   re = fma(-c.im, y.im, fma(c.re, y.re, re));  im = fma(c.im, y.re, fma(c.re, y.im, im)). */
void po_coef_apply_v1_f32(const float *C, const float *Y, float *O, int nSC, int L, int S, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            const float *c = C + 2 * ((size_t)sc * L + l) * A;
            for (int s = 0; s < S; s++) {
                const float *y = Y + 2 * ((size_t)sc * S + s) * A;
                float re = 0.f, im = 0.f;
                for (int a = 0; a < A; a++) {
                    re = fmaf(c[2 * a], y[2 * a], re);
                    re = fmaf(-c[2 * a + 1], y[2 * a + 1], re);
                    im = fmaf(c[2 * a], y[2 * a + 1], im);
                    im = fmaf(c[2 * a + 1], y[2 * a], im);
                }
                float *o = O + 2 * (((size_t)sc * L + l) * S + s);
                o[0] = re;
                o[1] = im;
            }
        }
}

/* Same chain with every product and sum rounded separately (the uncontracted reading of :60-64). */
void po_coef_apply_v1_f32_nofma(const float *C, const float *Y, float *O, int nSC, int L, int S, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            const float *c = C + 2 * ((size_t)sc * L + l) * A;
            for (int s = 0; s < S; s++) {
                const float *y = Y + 2 * ((size_t)sc * S + s) * A;
                float re = 0.f, im = 0.f;
                for (int a = 0; a < A; a++) {
                    float p0 = c[2 * a] * y[2 * a];
                    float p1 = c[2 * a + 1] * y[2 * a + 1];
                    float p2 = c[2 * a] * y[2 * a + 1];
                    float p3 = c[2 * a + 1] * y[2 * a];
                    re = (re + p0) - p1;
                    im = (im + p2) + p3;
                }
                float *o = O + 2 * (((size_t)sc * L + l) * S + s);
                o[0] = re;
                o[1] = im;
            }
        }
}

/* FP64 evaluation of the same product (inputs widened exactly). */
void po_coef_apply_f64(const float *C, const float *Y, double *O, int nSC, int L, int S, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            const float *c = C + 2 * ((size_t)sc * L + l) * A;
            for (int s = 0; s < S; s++) {
                const float *y = Y + 2 * ((size_t)sc * S + s) * A;
                double re = 0.0, im = 0.0;
                for (int a = 0; a < A; a++) {
                    double cr = c[2 * a], ci = c[2 * a + 1], yr = y[2 * a], yi = y[2 * a + 1];
                    re += cr * yr - ci * yi;
                    im += cr * yi + ci * yr;
                }
                double *o = O + 2 * (((size_t)sc * L + l) * S + s);
                o[0] = re;
                o[1] = im;
            }
        }
}

/* FP64 product with FP64 coefficients (the FP64 chain: C from po_mmse_f64). */
void po_coef_apply_f64d(const double *C, const float *Y, double *O, int nSC, int L, int S, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            const double *c = C + 2 * ((size_t)sc * L + l) * A;
            for (int s = 0; s < S; s++) {
                const float *y = Y + 2 * ((size_t)sc * S + s) * A;
                double re = 0.0, im = 0.0;
                for (int a = 0; a < A; a++) {
                    double cr = c[2 * a], ci = c[2 * a + 1], yr = y[2 * a], yi = y[2 * a + 1];
                    re += cr * yr - ci * yi;
                    im += cr * yi + ci * yr;
                }
                double *o = O + 2 * (((size_t)sc * L + l) * S + s);
                o[0] = re;
                o[1] = im;
            }
        }
}

/* Host OpenMP restatement of the scalar loop (the CPU baseline): the contracted chain of
 * po_coef_apply_v1_f32, `omp parallel for` over (subcarrier, layer). Returns the thread count used. */
int po_coef_apply_v1_f32_omp(const float *C, const float *Y, float *O, int nSC, int L, int S, int A,
                             int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    long total = (long)nSC * L;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long t = 0; t < total; t++) {
        long sc = t / L;
        int l = (int)(t % L);
        const float *c = C + 2 * ((size_t)sc * L + l) * A;
        for (int s = 0; s < S; s++) {
            const float *y = Y + 2 * ((size_t)sc * S + s) * A;
            float re = 0.f, im = 0.f;
            for (int a = 0; a < A; a++) {
                re = fmaf(c[2 * a], y[2 * a], re);
                re = fmaf(-c[2 * a + 1], y[2 * a + 1], re);
                im = fmaf(c[2 * a], y[2 * a + 1], im);
                im = fmaf(c[2 * a + 1], y[2 * a], im);
            }
            float *o = O + 2 * (((size_t)sc * L + l) * S + s);
            o[0] = re;
            o[1] = im;
        }
    }
    return nthreads;
}

/* ------------------------------------------------------------------------------------------------------
 * Gram product  G[sc][r][c] = sum_k M[sc][r][k] * H[sc][k][c]   (gram_scalar, aerial_bottleneck.cu:56-80)
 * ------------------------------------------------------------------------------------------------------ */
void po_gram_f32(const float *M, const float *H, float *G, int nSC, int L, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int r = 0; r < L; r++)
            for (int c = 0; c < L; c++) {
                const float *m = M + 2 * ((size_t)sc * L + r) * A;
                const float *h = H + 2 * ((size_t)sc * A * L + c);
                float re = 0.f, im = 0.f;
                for (int k = 0; k < A; k++) {
                    float mr = m[2 * k], mi = m[2 * k + 1];
                    float hr = h[2 * (size_t)k * L], hi = h[2 * (size_t)k * L + 1];
                    re = fmaf(mr, hr, re);
                    re = fmaf(-mi, hi, re);
                    im = fmaf(mr, hi, im);
                    im = fmaf(mi, hr, im);
                }
                float *g = G + 2 * (((size_t)sc * L + r) * L + c);
                g[0] = re;
                g[1] = im;
            }
}

void po_gram_f64(const float *M, const float *H, double *G, int nSC, int L, int A) {
    for (int sc = 0; sc < nSC; sc++)
        for (int r = 0; r < L; r++)
            for (int c = 0; c < L; c++) {
                const float *m = M + 2 * ((size_t)sc * L + r) * A;
                const float *h = H + 2 * ((size_t)sc * A * L + c);
                double re = 0.0, im = 0.0;
                for (int k = 0; k < A; k++) {
                    double mr = m[2 * k], mi = m[2 * k + 1];
                    double hr = h[2 * (size_t)k * L], hi = h[2 * (size_t)k * L + 1];
                    re += mr * hr - mi * hi;
                    im += mr * hi + mi * hr;
                }
                double *g = G + 2 * (((size_t)sc * L + r) * L + c);
                g[0] = re;
                g[1] = im;
            }
}

/* ------------------------------------------------------------------------------------------------------
 * MMSE coefficients in FP64.  The reference has no implementation (aerial_bottleneck.cu:51-53 names
 * eqMmseCoefCompHighMimoKernel only); this is the textbook regularized solve:
 *   G = H^H H + sigma2 I,  G = L L^H (Cholesky),  Ginv = L^-H L^-1,  W = Ginv H^H,
 *   mu_l = 1 - sigma2 Ginv[l][l],  noise_inv_l = mu_l / (1 - mu_l) = (1 - x) / x with x = sigma2 Ginv[l][l],
 *   cond = ||G||_1 ||Ginv||_1,  C = W or diag(1/mu) W (unbiased).
 * H is [nSC][A][L]; C is [nSC][L][A] (complex double); noise_inv, mu are [nSC][L]; cond is [nSC].
 * Returns the number of subcarriers whose Gram is not positive definite (their outputs are zero).
 * ------------------------------------------------------------------------------------------------------ */
#define PO_MAXL 8
int po_mmse_f64(const float *H, const float *sigma2, double *C, double *noise_inv, double *mu, double *cond,
                int nSC, int L, int A, int unbiased) {
    int bad = 0;
    for (int sc = 0; sc < nSC; sc++) {
        const float *h = H + 2 * (size_t)sc * A * L;
        double s2 = sigma2[sc];
        double gr[PO_MAXL][PO_MAXL], gi[PO_MAXL][PO_MAXL];
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                double re = 0.0, im = 0.0;
                for (int a = 0; a < A; a++) {
                    double ar = h[2 * (a * L + i)], ai = h[2 * (a * L + i) + 1];
                    double br = h[2 * (a * L + j)], bi = h[2 * (a * L + j) + 1];
                    re += ar * br + ai * bi; /* conj(a) * b */
                    im += ar * bi - ai * br;
                }
                gr[i][j] = re + (i == j ? s2 : 0.0);
                gi[i][j] = (i == j) ? 0.0 : im;
            }
        /* Cholesky, lower L */
        double lr[PO_MAXL][PO_MAXL] = {{0}}, li[PO_MAXL][PO_MAXL] = {{0}};
        int ok = 1;
        for (int j = 0; j < L && ok; j++) {
            double d = gr[j][j];
            for (int k = 0; k < j; k++) d -= lr[j][k] * lr[j][k] + li[j][k] * li[j][k];
            if (!(d > 0.0)) {
                ok = 0;
                break;
            }
            double ljj = sqrt(d);
            lr[j][j] = ljj;
            for (int i = j + 1; i < L; i++) {
                /* L[i][j] = (G[i][j] - sum_k L[i][k] conj(L[j][k])) / L[j][j] */
                double re = gr[i][j], im = gi[i][j];
                for (int k = 0; k < j; k++) {
                    re -= lr[i][k] * lr[j][k] + li[i][k] * li[j][k];
                    im -= li[i][k] * lr[j][k] - lr[i][k] * li[j][k];
                }
                lr[i][j] = re / ljj;
                li[i][j] = im / ljj;
            }
        }
        double *csc = C + 2 * (size_t)sc * L * A;
        if (!ok) {
            bad++;
            memset(csc, 0, sizeof(double) * 2 * (size_t)L * A);
            for (int l = 0; l < L; l++) {
                noise_inv[(size_t)sc * L + l] = 0.0;
                if (mu) mu[(size_t)sc * L + l] = 0.0;
            }
            if (cond) cond[sc] = INFINITY;
            continue;
        }
        /* W = L^-1 (lower) */
        double wr[PO_MAXL][PO_MAXL] = {{0}}, wi[PO_MAXL][PO_MAXL] = {{0}};
        for (int j = 0; j < L; j++) {
            wr[j][j] = 1.0 / lr[j][j];
            for (int i = j + 1; i < L; i++) {
                double re = 0.0, im = 0.0;
                for (int k = j; k < i; k++) {
                    re += lr[i][k] * wr[k][j] - li[i][k] * wi[k][j];
                    im += lr[i][k] * wi[k][j] + li[i][k] * wr[k][j];
                }
                wr[i][j] = -re / lr[i][i];
                wi[i][j] = -im / lr[i][i];
            }
        }
        /* Ginv = W^H W */
        double vr[PO_MAXL][PO_MAXL], vi[PO_MAXL][PO_MAXL];
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                double re = 0.0, im = 0.0;
                int k0 = i > j ? i : j;
                for (int k = k0; k < L; k++) {
                    re += wr[k][i] * wr[k][j] + wi[k][i] * wi[k][j];
                    im += wr[k][i] * wi[k][j] - wi[k][i] * wr[k][j];
                }
                vr[i][j] = re;
                vi[i][j] = im;
            }
        if (cond) {
            double ng = 0.0, nv = 0.0;
            for (int j = 0; j < L; j++) {
                double cg = 0.0, cv = 0.0;
                for (int i = 0; i < L; i++) {
                    cg += hypot(gr[i][j], gi[i][j]);
                    cv += hypot(vr[i][j], vi[i][j]);
                }
                if (cg > ng) ng = cg;
                if (cv > nv) nv = cv;
            }
            cond[sc] = ng * nv;
        }
        for (int l = 0; l < L; l++) {
            double x = s2 * vr[l][l];
            double m = 1.0 - x;
            if (mu) mu[(size_t)sc * L + l] = m;
            double xf = x > 1e-30 ? x : 1e-30;
            double ni = (1.0 - x) / xf;
            noise_inv[(size_t)sc * L + l] = ni;
            double scale = (unbiased && m != 0.0) ? 1.0 / m : 1.0;
            for (int a = 0; a < A; a++) {
                double re = 0.0, im = 0.0;
                for (int j = 0; j < L; j++) {
                    double hr = h[2 * (a * L + j)], hi = -h[2 * (a * L + j) + 1]; /* conj(H[a][j]) */
                    re += vr[l][j] * hr - vi[l][j] * hi;
                    im += vr[l][j] * hi + vi[l][j] * hr;
                }
                csc[2 * ((size_t)l * A + a)] = re * scale;
                csc[2 * ((size_t)l * A + a) + 1] = im * scale;
            }
        }
    }
    return bad;
}

/* ------------------------------------------------------------------------------------------------------
 * max-log LLR (computePamLlr)
 * ------------------------------------------------------------------------------------------------------ */

/* LUT of aerial_llr_bench_v2.cu:47-54 (PAM2, PAM4, PAM8, PAM16 distances at offsets 0, 1, 3, 6), kept as the
 * same decimal literals; the tests check them against 1/sqrt(2), {2,1}/sqrt(10), {4,2,1}/sqrt(42),
 * {8,4,2,1}/sqrt(170). */
static const double LUT_D[10] = {0.707106781186548, 0.632455532033676, 0.316227766016838, 0.617213399848368,
                                 0.308606699924184, 0.154303349962092, 0.613571991077897, 0.306785995538948,
                                 0.153392997769474, 0.076696498884737};
static const float LUT_F[10] = {0.707106781186548f, 0.632455532033676f, 0.316227766016838f,
                                0.617213399848368f, 0.308606699924184f, 0.154303349962092f,
                                0.613571991077897f, 0.306785995538948f, 0.153392997769474f,
                                0.076696498884737f};
static const int PAM_OFF[4] = {0, 1, 3, 6};

double po_llr_lut(int idx) { return (idx >= 0 && idx < 10) ? LUT_D[idx] : NAN; }

/* cpu_computePamLlr_ref, aerial_llr_bench_v2.cu:72-107 (= cpu_ref, aerial_llr_v3.cu:61-79): FP64 folding
 * chain per axis, out[i*2 + iq] as float. */
void po_llr_ref_f64(float pamI, float pamQ, float noise_inv, int nPamBits, float *out) {
    int off = PAM_OFF[nPamBits - 1];
    double dist = LUT_D[off + nPamBits - 1];
    for (int iq = 0; iq < 2; iq++) {
        double sb[4], md[4];
        sb[0] = iq == 0 ? (double)pamI : (double)pamQ;
        unsigned sign = 0;
        for (int i = 0; i < nPamBits - 1; i++) {
            sb[i + 1] = LUT_D[off + i] - fabs(sb[i]);
            md[i] = dist + fabs(sb[i]);
            md[i] *= md[i];
            if (sb[i] < 0.0) sign |= 1u << i;
        }
        md[nPamBits - 1] = dist + fabs(sb[nPamBits - 1]);
        md[nPamBits - 1] *= md[nPamBits - 1];
        if (sb[nPamBits - 1] < 0.0) sign |= 1u << (nPamBits - 1);
        double m1 = fabs(sb[nPamBits - 1]) - dist;
        m1 *= m1;
        for (int i = 0; i < nPamBits; i++) {
            double llr = (md[i] - m1) * (double)noise_inv;
            if (sign & (1u << i)) llr = -llr;
            if (llr < -65504.0) llr = -65504.0;
            if (llr > 65504.0) llr = 65504.0;
            out[i * 2 + iq] = (float)llr;
        }
    }
}

/* llr_v1_baseline arithmetic, aerial_llr_bench_v2.cu:134-155 (= aerial_llr_v3.cu:93-107): the same chain in
 * FP32 with the float LUT, clamp to +-65504; out[i*2 + iq] before the binary16 conversion. */
void po_llr_f32(float pamI, float pamQ, float noise, int nPamBits, float *out) {
    int off = PAM_OFF[nPamBits - 1];
    float dist = LUT_F[off + nPamBits - 1];
    for (int iq = 0; iq < 2; iq++) {
        float sb[4], md[4];
        sb[0] = iq == 0 ? pamI : pamQ;
        unsigned sign = 0;
        for (int i = 0; i < nPamBits - 1; i++) {
            sb[i + 1] = LUT_F[off + i] - fabsf(sb[i]);
            md[i] = dist + fabsf(sb[i]);
            md[i] *= md[i];
            if (sb[i] < 0.f) sign |= 1u << i;
        }
        md[nPamBits - 1] = dist + fabsf(sb[nPamBits - 1]);
        md[nPamBits - 1] *= md[nPamBits - 1];
        if (sb[nPamBits - 1] < 0.f) sign |= 1u << (nPamBits - 1);
        float m1 = fabsf(sb[nPamBits - 1]) - dist;
        m1 *= m1;
        for (int i = 0; i < nPamBits; i++) {
            float llr = (md[i] - m1) * noise;
            if (sign & (1u << i)) llr = -llr;
            llr = fmaxf(-65504.f, fminf(65504.f, llr));
            out[i * 2 + iq] = llr;
        }
    }
}

/* Whole llr_v1_baseline (aerial_llr_bench_v2.cu:114-156): binary16 in, FP32 math, binary16 out. */
void po_llr_demap_f16(const uint16_t *soft_half, const float *noise_inv, uint16_t *out, int N, int nPamBits) {
    float tmp[8];
    for (int i = 0; i < N; i++) {
        po_llr_f32(po_f16_to_f32(soft_half[2 * i]), po_f16_to_f32(soft_half[2 * i + 1]), noise_inv[i],
                   nPamBits, tmp);
        for (int k = 0; k < 2 * nPamBits; k++) out[(size_t)i * 2 * nPamBits + k] = po_f32_to_f16(tmp[k]);
    }
}

/* The FP64 reference loop of aerial_llr_bench_v2.cu:364-371 over all symbols. */
void po_llr_demap_ref_f64(const uint16_t *soft_half, const float *noise_inv, float *out, int N, int nPamBits) {
    for (int i = 0; i < N; i++)
        po_llr_ref_f64(po_f16_to_f32(soft_half[2 * i]), po_f16_to_f32(soft_half[2 * i + 1]), noise_inv[i],
                       nPamBits, out + (size_t)i * 2 * nPamBits);
}

/* FP32 demap on the equalizer layout: Y_eq[sc][l][s] complex, noise_inv[sc][l] ->
 * llr[sc][l][s][2q], each symbol through po_llr_f32. */
void po_llr_demap_f32(const float *Yeq, const float *noise_inv, float *llr, int nSC, int L, int S,
                      int nPamBits) {
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            float ni = noise_inv[(size_t)sc * L + l];
            for (int s = 0; s < S; s++) {
                size_t e = ((size_t)sc * L + l) * S + s;
                po_llr_f32(Yeq[2 * e], Yeq[2 * e + 1], ni, nPamBits, llr + e * 2 * nPamBits);
            }
        }
}

/* FP64 folding chain on double inputs (the FP64 end of the chain), same layout as po_llr_demap_f32. */
void po_llr_demap_f64(const double *Yeq, const double *noise_inv, double *llr, int nSC, int L, int S,
                      int nPamBits) {
    int off = PAM_OFF[nPamBits - 1];
    double dist = LUT_D[off + nPamBits - 1];
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            double ni = noise_inv[(size_t)sc * L + l];
            for (int s = 0; s < S; s++) {
                size_t e = ((size_t)sc * L + l) * S + s;
                for (int iq = 0; iq < 2; iq++) {
                    double sb[4], md[4];
                    sb[0] = Yeq[2 * e + iq];
                    unsigned sign = 0;
                    for (int i = 0; i < nPamBits - 1; i++) {
                        sb[i + 1] = LUT_D[off + i] - fabs(sb[i]);
                        md[i] = dist + fabs(sb[i]);
                        md[i] *= md[i];
                        if (sb[i] < 0.0) sign |= 1u << i;
                    }
                    md[nPamBits - 1] = dist + fabs(sb[nPamBits - 1]);
                    md[nPamBits - 1] *= md[nPamBits - 1];
                    if (sb[nPamBits - 1] < 0.0) sign |= 1u << (nPamBits - 1);
                    double m1 = fabs(sb[nPamBits - 1]) - dist;
                    m1 *= m1;
                    for (int i = 0; i < nPamBits; i++) {
                        double v = (md[i] - m1) * ni;
                        if (sign & (1u << i)) v = -v;
                        if (v < -65504.0) v = -65504.0;
                        if (v > 65504.0) v = 65504.0;
                        llr[e * 2 * nPamBits + i * 2 + iq] = v;
                    }
                }
            }
        }
}

/* Brute-force max-log LLR of one PAM axis under the TS 38.211 section 5.1 Gray mapping: the 2^m levels are
 * t_0 / sqrt(2(4^m - 1)/3) with t_{m-1} = 1 - 2c_{m-1}, t_i = (1 - 2c_i)(2^{m-1-i} - t_{i+1}); bit c_i of
 * the axis is 38.211 bit b_{2i} (I) or b_{2i+1} (Q).  out[i] = (min_{c_i=1} d^2 - min_{c_i=0} d^2) * noise. */
void po_llr_bruteforce_f64(double x, double noise, int nPamBits, double *out) {
    int m = nPamBits;
    double norm = sqrt(2.0 * (pow(4.0, m) - 1.0) / 3.0);
    double best0[4], best1[4];
    for (int i = 0; i < m; i++) best0[i] = best1[i] = INFINITY;
    for (int code = 0; code < (1 << m); code++) {
        int c[4] = {0, 0, 0, 0};
        for (int i = 0; i < m; i++) c[i] = (code >> i) & 1;
        double t = 1.0 - 2.0 * c[m - 1];
        for (int i = m - 2; i >= 0; i--) t = (1.0 - 2.0 * c[i]) * ((double)(1 << (m - 1 - i)) - t);
        double amp = t / norm;
        double d2 = (x - amp) * (x - amp);
        for (int i = 0; i < m; i++) {
            if (c[i]) {
                if (d2 < best1[i]) best1[i] = d2;
            } else if (d2 < best0[i])
                best0[i] = d2;
        }
    }
    for (int i = 0; i < m; i++) out[i] = (best1[i] - best0[i]) * noise;
}

/* ------------------------------------------------------------------------------------------------------
 * CPU baseline of the fused stage: the scalar coef-apply chain then the FP32 demapper per symbol.
 * ------------------------------------------------------------------------------------------------------ */
int po_coef_apply_demap_f32_omp(const float *C, const float *Y, const float *noise_inv, float *llr,
                                int nSC, int L, int S, int A, int nPamBits, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    long total = (long)nSC * L;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long t = 0; t < total; t++) {
        long sc = t / L;
        int l = (int)(t % L);
        const float *c = C + 2 * ((size_t)sc * L + l) * A;
        float ni = noise_inv[t];
        for (int s = 0; s < S; s++) {
            const float *y = Y + 2 * ((size_t)sc * S + s) * A;
            float re = 0.f, im = 0.f;
            for (int a = 0; a < A; a++) {
                re = fmaf(c[2 * a], y[2 * a], re);
                re = fmaf(-c[2 * a + 1], y[2 * a + 1], re);
                im = fmaf(c[2 * a], y[2 * a + 1], im);
                im = fmaf(c[2 * a + 1], y[2 * a], im);
            }
            size_t e = ((size_t)sc * L + l) * S + s;
            po_llr_f32(re, im, ni, nPamBits, llr + e * 2 * nPamBits);
        }
    }
    return nthreads;
}

int po_omp_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------------------------------------
 * binary16 arithmetic: every operation rounded once to nearest even, the semantics of __hmul, __hadd,
 * __hsub and __hfma. Products of two binary16 values (22 significant bits) and sums of two (at most 40 bits
 * between 2^15 and 2^-24) are exact in double; the fused multiply-add carries the double sum's rounding
 * error (TwoSum) into the final rounding, so ties are broken on the exact value.
 * ------------------------------------------------------------------------------------------------------ */

/* binary16 nearest to hi + lo (|lo| at most half an ulp of hi in double), ties to even */
static uint16_t f64x2_to_f16(double hi, double lo) {
    uint16_t sign = signbit(hi) ? 0x8000u : 0u;
    double a = fabs(hi);
    if (isnan(a)) return (uint16_t)(sign | 0x7e00u);
    if (isinf(a)) return (uint16_t)(sign | 0x7c00u);
    if (a == 0.0) return sign; /* an exact zero sum has no rounding error */
    if (signbit(hi)) lo = -lo; /* lo relative to |hi| */
    int e2;
    frexp(a, &e2); /* a = f * 2^e2, f in [0.5, 1) */
    int e = e2 - 1;   /* a in [2^e, 2^(e+1)) */
    int sub = e < -14;
    double q = sub ? ldexp(a, 24) : ldexp(a, 10 - e); /* exact scaling: integer part = kept significand */
    double fl = floor(q), fr = q - fl;
    double n;
    if (fr > 0.5 || (fr == 0.5 && (lo > 0 || (lo == 0 && fmod(fl, 2.0) != 0.0)))) n = fl + 1.0;
    else n = fl;
    if (sub) return (uint16_t)(sign | (uint16_t)n); /* n == 1024 encodes the smallest normal */
    if (n == 2048.0) { n = 1024.0; e += 1; }
    if (e > 15) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | ((uint32_t)(e + 15) << 10) | ((uint32_t)n - 1024u));
}

static double f16_to_f64(uint16_t h) { return (double)po_f16_to_f32(h); }

uint16_t po_hmul(uint16_t a, uint16_t b) { return f64x2_to_f16(f16_to_f64(a) * f16_to_f64(b), 0.0); }
uint16_t po_hadd(uint16_t a, uint16_t b) { return f64x2_to_f16(f16_to_f64(a) + f16_to_f64(b), 0.0); }
uint16_t po_hsub(uint16_t a, uint16_t b) { return f64x2_to_f16(f16_to_f64(a) - f16_to_f64(b), 0.0); }
uint16_t po_hfma(uint16_t a, uint16_t b, uint16_t c) {
    double p = f16_to_f64(a) * f16_to_f64(b), x = f16_to_f64(c);
    double s = p + x;
    double bp = s - p;
    double err = (p - (s - bp)) + (x - bp);
    if (isinf(s) || isnan(s)) err = 0.0;
    return f64x2_to_f16(s, err);
}

void po_half_ops(const uint16_t *a, const uint16_t *b, const uint16_t *c, uint16_t *mul, uint16_t *add,
                 uint16_t *sub, uint16_t *fma_, size_t n) {
    for (size_t i = 0; i < n; i++) {
        mul[i] = po_hmul(a[i], b[i]);
        add[i] = po_hadd(a[i], b[i]);
        sub[i] = po_hsub(a[i], b[i]);
        fma_[i] = po_hfma(a[i], b[i], c[i]);
    }
}

/* ------------------------------------------------------------------------------------------------------
 * SRS filter multiply (channel estimation from sounding reference signals)
 *   H_est[b][ant][sc][port] = sum_k (W[sc][k] * FOCC[port][k % foccLen]) * SRS[b][ant][k]
 * srs_wmma_bench.cu srs_filter_v1_scalar :57-90 with complex_conjmul :43-47 and hcmadd :48-52. The
 * helper named complex_conjmul multiplies without conjugating (its real part is re*re - im*im); conj_w != 0
 * conjugates W first, the form the operation's comment (:6) states. Complex values are binary16 pairs
 * {re, im}.
 * ------------------------------------------------------------------------------------------------------ */
void po_srs_filter_v1(uint16_t *H_est, const uint16_t *SRS, const uint16_t *W, const uint16_t *FOCC, int nAnt,
                      int nBlocks, int K, int nPorts, int foccLen, int conj_w) {
    for (int b = 0; b < nBlocks; b++)
        for (int ant = 0; ant < nAnt; ant++)
            for (int sc = 0; sc < K; sc++)
                for (int p = 0; p < nPorts; p++) {
                    const uint16_t *srs = SRS + 2 * ((size_t)b * nAnt + ant) * K;
                    const uint16_t *w = W + 2 * (size_t)sc * K;
                    const uint16_t *f = FOCC + 2 * (size_t)p * foccLen;
                    uint16_t er = 0, ei = 0;
                    for (int k = 0; k < K; k++) {
                        uint16_t wr = w[2 * k], wi = w[2 * k + 1];
                        if (conj_w) wi ^= 0x8000u;
                        uint16_t fr = f[2 * (k % foccLen)], fi = f[2 * (k % foccLen) + 1];
                        uint16_t pr = po_hsub(po_hmul(wr, fr), po_hmul(wi, fi));
                        uint16_t pi = po_hadd(po_hmul(wr, fi), po_hmul(wi, fr));
                        uint16_t sr = srs[2 * k], si = srs[2 * k + 1];
                        er = po_hfma(pr, sr, po_hfma((uint16_t)(pi ^ 0x8000u), si, er));
                        ei = po_hfma(pr, si, po_hfma(pi, sr, ei));
                    }
                    size_t o = (((size_t)b * nAnt + ant) * K + sc) * nPorts + p;
                    H_est[2 * o] = er;
                    H_est[2 * o + 1] = ei;
                }
}

/* Same operation in double from the binary16 inputs (no intermediate rounding); prod_f16 != 0 rounds the
 * W * FOCC products to binary16 as the reference forms them before the sum. */
void po_srs_filter_f64(double *H_est, const uint16_t *SRS, const uint16_t *W, const uint16_t *FOCC, int nAnt,
                       int nBlocks, int K, int nPorts, int foccLen, int conj_w, int prod_f16) {
    for (int b = 0; b < nBlocks; b++)
        for (int ant = 0; ant < nAnt; ant++)
            for (int sc = 0; sc < K; sc++)
                for (int p = 0; p < nPorts; p++) {
                    const uint16_t *srs = SRS + 2 * ((size_t)b * nAnt + ant) * K;
                    const uint16_t *w = W + 2 * (size_t)sc * K;
                    const uint16_t *f = FOCC + 2 * (size_t)p * foccLen;
                    double er = 0, ei = 0;
                    for (int k = 0; k < K; k++) {
                        double wr = f16_to_f64(w[2 * k]), wi = f16_to_f64(w[2 * k + 1]);
                        if (conj_w) wi = -wi;
                        double fr = f16_to_f64(f[2 * (k % foccLen)]), fi = f16_to_f64(f[2 * (k % foccLen) + 1]);
                        double pr = wr * fr - wi * fi, pi = wr * fi + wi * fr;
                        if (prod_f16) {
                            uint16_t wr16 = w[2 * k], wi16 = w[2 * k + 1];
                            if (conj_w) wi16 ^= 0x8000u;
                            uint16_t fr16 = f[2 * (k % foccLen)], fi16 = f[2 * (k % foccLen) + 1];
                            pr = f16_to_f64(po_hsub(po_hmul(wr16, fr16), po_hmul(wi16, fi16)));
                            pi = f16_to_f64(po_hadd(po_hmul(wr16, fi16), po_hmul(wi16, fr16)));
                        }
                        double sr = f16_to_f64(srs[2 * k]), si = f16_to_f64(srs[2 * k + 1]);
                        er += pr * sr - pi * si;
                        ei += pr * si + pi * sr;
                    }
                    size_t o = (((size_t)b * nAnt + ant) * K + sc) * nPorts + p;
                    H_est[2 * o] = er;
                    H_est[2 * o + 1] = ei;
                }
}

/* Inputs of srs_wmma_bench.cu run_config :262-268: srand(42), then SRS, W and FOCC from one stream, each
 * complex value {re, im} = two draws of __float2half(((float)rand()/RAND_MAX - 0.5f) * 1.4f). */
void po_srs_bench_inputs(unsigned seed, int nAnt, int nBlocks, int K, int nPorts, int foccLen, uint16_t *SRS,
                         uint16_t *W, uint16_t *FOCC) {
    srand(seed);
    size_t n[3] = {(size_t)nBlocks * nAnt * K, (size_t)K * K, (size_t)nPorts * foccLen};
    uint16_t *dst[3] = {SRS, W, FOCC};
    for (int t = 0; t < 3; t++)
        for (size_t i = 0; i < 2 * n[t]; i++)
            dst[t][i] = po_f32_to_f16(((float)rand() / RAND_MAX - 0.5f) * 1.4f);
}

/* ------------------------------------------------------------------------------------------------------
 * Fronthaul block floating point (O-RAN WG4 CUS-plane U-plane compression, cited by README.md:214; no code
 * in the reference). A compressed PRB is one udCompParam byte (exponent in bits 3..0) followed by 24
 * mantissas I0 Q0 I1 Q1 ... I11 Q11, each an iqWidth-bit two's complement integer, packed most significant
 * bit first; sample = mantissa * 2^exponent. Blocks are stored per (symbol, antenna) row, PRBs consecutive:
 * in[(sym * nAnt + ant) * row_stride + prb * (1 + 3 * iqWidth)].
 * ------------------------------------------------------------------------------------------------------ */
static int bfp_get(const uint8_t *blk, int bitoff, int w) {
    uint32_t v = 0;
    for (int b = 0; b < w; b++) {
        const int bit = bitoff + b;
        v = (v << 1) | ((blk[bit >> 3] >> (7 - (bit & 7))) & 1u);
    }
    return (v & (1u << (w - 1))) ? (int)v - (1 << w) : (int)v; /* sign extension */
}

static void bfp_put(uint8_t *blk, int bitoff, int w, int m) {
    const uint32_t v = (uint32_t)m & ((1u << w) - 1u);
    for (int b = 0; b < w; b++) {
        const int bit = bitoff + b;
        const uint8_t one = (uint8_t)(1u << (7 - (bit & 7)));
        if ((v >> (w - 1 - b)) & 1u) blk[bit >> 3] |= one;
        else blk[bit >> 3] &= (uint8_t)~one;
    }
}

/* Y[(sc * nSym + sym) * nAnt + ant] (complex float, sc = prb * 12 + re) = scale * mantissa * 2^exponent */
void po_bfp_decompress(const uint8_t *in, float *Y, int nPrb, int nSym, int nAnt, int iqWidth, size_t row_stride,
                       float scale) {
    const int bpb = 1 + 3 * iqWidth;
    for (int sym = 0; sym < nSym; sym++)
        for (int ant = 0; ant < nAnt; ant++)
            for (int prb = 0; prb < nPrb; prb++) {
                const uint8_t *blk = in + ((size_t)sym * nAnt + ant) * row_stride + (size_t)prb * bpb;
                const int e = blk[0] & 0xf;
                for (int re = 0; re < 12; re++) {
                    const int mi = bfp_get(blk, 8 + (2 * re) * iqWidth, iqWidth);
                    const int mq = bfp_get(blk, 8 + (2 * re + 1) * iqWidth, iqWidth);
                    const size_t o = (((size_t)(prb * 12 + re) * nSym + sym) * nAnt + ant) * 2;
                    Y[o] = (float)ldexp((double)mi, e) * scale;
                    Y[o + 1] = (float)ldexp((double)mq, e) * scale;
                }
            }
}

/* Test-vector encoder: the smallest exponent e for which every value of the PRB, shifted right by e
   (arithmetic shift, i.e. rounded toward minus infinity), fits iqWidth-bit two's complement. Samples are
   int32 [sym][ant][sc][2] (sc = prb * 12 + re). */
void po_bfp_compress(const int32_t *x, uint8_t *out, int nPrb, int nSym, int nAnt, int iqWidth, size_t row_stride) {
    const int bpb = 1 + 3 * iqWidth;
    const int32_t hi = (1 << (iqWidth - 1)) - 1, lo = -(1 << (iqWidth - 1));
    for (int sym = 0; sym < nSym; sym++)
        for (int ant = 0; ant < nAnt; ant++)
            for (int prb = 0; prb < nPrb; prb++) {
                const int32_t *v = x + (((size_t)sym * nAnt + ant) * nPrb * 12 + (size_t)prb * 12) * 2;
                int e = 0;
                for (;; e++) {
                    int fits = 1;
                    for (int k = 0; k < 24; k++) {
                        const int32_t m = v[k] >> e;
                        if (m > hi || m < lo) fits = 0;
                    }
                    if (fits || e == 15) break;
                }
                uint8_t *blk = out + ((size_t)sym * nAnt + ant) * row_stride + (size_t)prb * bpb;
                blk[0] = (uint8_t)e;
                for (int k = 0; k < 24; k++) {
                    int32_t m = v[k] >> e;
                    m = m > hi ? hi : (m < lo ? lo : m);
                    bfp_put(blk, 8 + k * iqWidth, iqWidth, m);
                }
            }
}

/* ------------------------------------------------------------------------------------------------------
 * DMRS frequency interpolation: H_f[b][m][n] = sum_k W[m][k] X[b][k][n]
 * The reference names this GEMM ([64 x 96] @ [96 x 4], batched over OFDM symbols and UEs, README.md:200)
 * without code; the chain below uses the complex multiply-accumulate of cma() (aerial_coef_wmma.cu:60-64) in
 * its fused form, k ascending, the order every FP32 kernel of the reference uses for its K loop.
 * ------------------------------------------------------------------------------------------------------ */
static void freq_interp_item_f32(const float *W, const float *X, float *Hf, int M, int K, int NC) {
    for (int m = 0; m < M; m++)
        for (int n = 0; n < NC; n++) {
            float re = 0.f, im = 0.f;
            for (int k = 0; k < K; k++) {
                const float ar = W[2 * ((size_t)m * K + k)], ai = W[2 * ((size_t)m * K + k) + 1];
                const float br = X[2 * ((size_t)k * NC + n)], bi = X[2 * ((size_t)k * NC + n) + 1];
                re = fmaf(ar, br, re);
                re = fmaf(-ai, bi, re);
                im = fmaf(ar, bi, im);
                im = fmaf(ai, br, im);
            }
            Hf[2 * ((size_t)m * NC + n)] = re;
            Hf[2 * ((size_t)m * NC + n) + 1] = im;
        }
}

void po_freq_interp_f32(const float *W, const float *X, float *Hf, int nBatch, int M, int K, int NC) {
    for (int b = 0; b < nBatch; b++)
        freq_interp_item_f32(W, X + 2 * (size_t)b * K * NC, Hf + 2 * (size_t)b * M * NC, M, K, NC);
}

/* The same chain, OpenMP over the batch (the timed CPU baseline); returns the thread count used. */
int po_freq_interp_f32_omp(const float *W, const float *X, float *Hf, int nBatch, int M, int K, int NC,
                           int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int used = 1;
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(static)
        for (int b = 0; b < nBatch; b++)
            freq_interp_item_f32(W, X + 2 * (size_t)b * K * NC, Hf + 2 * (size_t)b * M * NC, M, K, NC);
    }
    return used;
}

/* FP64 sum of the exact products of the given operands (double W, so that operands rounded to a narrower
 * format by the caller are taken as they are). */
void po_freq_interp_f64(const double *W, const double *X, double *Hf, int nBatch, int M, int K, int NC) {
    for (int b = 0; b < nBatch; b++) {
        const double *x = X + 2 * (size_t)b * K * NC;
        double *h = Hf + 2 * (size_t)b * M * NC;
        for (int m = 0; m < M; m++)
            for (int n = 0; n < NC; n++) {
                double re = 0.0, im = 0.0;
                for (int k = 0; k < K; k++) {
                    const double ar = W[2 * ((size_t)m * K + k)], ai = W[2 * ((size_t)m * K + k) + 1];
                    const double br = x[2 * ((size_t)k * NC + n)], bi = x[2 * ((size_t)k * NC + n) + 1];
                    re += ar * br - ai * bi;
                    im += ar * bi + ai * br;
                }
                h[2 * ((size_t)m * NC + n)] = re;
                h[2 * ((size_t)m * NC + n) + 1] = im;
            }
    }
}

/* ------------------------------------------------------------------------------------------------------
 * Descrambling and rate-recovery layout of the demapper output. The reference hands its binary16 LLRs to the
 * decoder in the per-symbol order [b0_I, b0_Q, b1_I, ...] (aerial_llr_bench_v2.cu:117,221-227) and stops there;
 * the receiver steps that follow are fixed by the 3GPP specifications, restated here bit by bit:
 *   TS 38.211 5.2.1   pseudo-random (Gold) sequence c(n) = (x1(n + Nc) + x2(n + Nc)) mod 2, Nc = 1600,
 *                     x1(n + 31) = (x1(n + 3) + x1(n)) mod 2,  x1(0) = 1, x1(1..30) = 0,
 *                     x2(n + 31) = (x2(n + 3) + x2(n + 2) + x2(n + 1) + x2(n)) mod 2,  x2(i) = bit i of c_init;
 *   TS 38.211 6.3.1.1 scrambling b~(i) = (b(i) + c(i)) mod 2 over the codeword's bits (c_init = n_RNTI 2^15 +
 *                     n_ID for PUSCH), so the receiver negates the LLR of bit i where c(i) = 1;
 *   TS 38.211 6.3.1.2-6.3.1.6 modulation (Qm bits per symbol, bit i of the symbol = codeword bit m Qm + i),
 *                     layer mapping of one codeword onto nu layers (layer v carries codeword symbol m = nu i + v)
 *                     and mapping to resource elements frequency first (layer symbol i = s nSC + sc over the
 *                     allocation's nSC subcarriers and data symbols s);
 *   TS 38.212 5.4.2.1 rate-matching output length of code block r of C: E_r = nu Qm floor(G / (nu Qm C)) for
 *                     r < C - mod(G / (nu Qm), C), else nu Qm ceil(G / (nu Qm C)); blocks concatenated (5.5);
 *   TS 38.212 5.4.2.2 bit interleaving f(i + j Qm) = e(i E/Qm + j), undone here: e(i E/Qm + j) = f(i + j Qm).
 * ------------------------------------------------------------------------------------------------------ */

/* c(0 .. n_bits - 1), one byte per bit, by running both registers bit by bit from their initial states. */
void po_gold_sequence(uint32_t c_init, int n_bits, uint8_t *c) {
    const int NC = 1600, len = n_bits + NC + 31;
    uint8_t *x1 = (uint8_t *)calloc((size_t)len, 1), *x2 = (uint8_t *)calloc((size_t)len, 1);
    x1[0] = 1;
    for (int i = 0; i < 31; i++) x2[i] = (uint8_t)((c_init >> i) & 1u);
    for (int n = 0; n + 31 < len; n++) {
        x1[n + 31] = (uint8_t)((x1[n + 3] + x1[n]) & 1);
        x2[n + 31] = (uint8_t)((x2[n + 3] + x2[n + 2] + x2[n + 1] + x2[n]) & 1);
    }
    for (int n = 0; n < n_bits; n++) c[n] = (uint8_t)((x1[n + NC] + x2[n + NC]) & 1);
    free(x1);
    free(x2);
}

/* E_r of the C code blocks of a codeword of G bits on nu layers with Qm bits per symbol (38.212 5.4.2.1, all C
 * blocks scheduled). Returns 0, or -1 when C is out of range (C < 1 or more blocks than symbols per layer). */
int po_cw_block_sizes(int G, int nu, int qm, int C, int *E) {
    const int per = nu * qm, N = G / per;
    if (C < 1 || C > N || G % per) return -1;
    for (int r = 0; r < C; r++) E[r] = (r < C - N % C) ? per * (N / C) : per * ((N + C - 1) / C);
    return 0;
}

/* The demapper's LLRs llr[sc][l][s][2 nPamBits] moved into per-codeword arrays: codeword k holds the layers l with
 * layer_cw[l] == k (their rank v in increasing l), G_k = nSC S nu_k Qm LLRs. Codeword bit g = m Qm + i with
 * m = (s nSC + sc) nu_k + v is descrambled (negated where c_k[g] = 1; c_k == NULL: no descrambling) and stored at
 * g (nCodeBlocks[k] == 0: codeword order) or, with C = nCodeBlocks[k] blocks, de-interleaved within its block. */
void po_cw_layout(const float *llr, int nSC, int L, int S, int nPamBits, const int *layer_cw, int nCodewords,
                  const uint8_t *const *c, const int *nCodeBlocks, float *const *out) {
    const int qm = 2 * nPamBits;
    int nu[8] = {0}, v[8], start[8][64], E[8][64];
    for (int l = 0; l < L; l++) v[l] = nu[layer_cw[l]]++;
    for (int k = 0; k < nCodewords; k++) {
        const int G = nSC * S * nu[k] * qm, C = nCodeBlocks[k];
        if (C > 0) {
            po_cw_block_sizes(G, nu[k], qm, C, E[k]);
            start[k][0] = 0;
            for (int r = 1; r < C; r++) start[k][r] = start[k][r - 1] + E[k][r - 1];
        }
    }
    for (int sc = 0; sc < nSC; sc++)
        for (int l = 0; l < L; l++) {
            const int k = layer_cw[l];
            for (int s = 0; s < S; s++) {
                const int m = (s * nSC + sc) * nu[k] + v[l];
                for (int i = 0; i < qm; i++) {
                    const int g = m * qm + i;
                    float x = llr[(((size_t)sc * L + l) * S + s) * qm + i];
                    if (c && c[k] && c[k][g]) x = -x;
                    int pos = g;
                    if (nCodeBlocks[k] > 0) {
                        int r = 0;
                        while (r + 1 < nCodeBlocks[k] && g >= start[k][r + 1]) r++;
                        const int t = g - start[k][r];
                        pos = start[k][r] + (t % qm) * (E[k][r] / qm) + t / qm;
                    }
                    out[k][pos] = x;
                }
            }
        }
}
