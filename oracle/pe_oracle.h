/*
 * pe_oracle.h -- CPU oracle for the PUSCH equalization path. TEST INFRASTRUCTURE ONLY: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it as the checker. The product library never links,
 * loads or calls it.
 *
 * Each function restates the arithmetic of a reference file (cited per function) in plain C. Complex
 * arrays are interleaved {re, im} floats (or doubles) in the reference layouts.
 */
#ifndef PE_ORACLE_H
#define PE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

/* -- input generators that reproduce the reference's glibc rand() streams -- */
void po_glibc_uniform(unsigned seed, float scale, int apply_scale, float *out, size_t nfloats);
void po_llr_bench_inputs(unsigned seed, int N, uint16_t *soft_half, float *noise_inv);

/* -- IEEE binary16 -- */
uint16_t po_f32_to_f16(float f);
float po_f16_to_f32(uint16_t h);
void po_f32_to_f16_array(const float *in, uint16_t *out, size_t n);
void po_f16_to_f32_array(const uint16_t *in, float *out, size_t n);

/* -- coefficient apply -- */
void po_coef_apply_v1_f32(const float *C, const float *Y, float *O, int nSC, int L, int S, int A);
void po_coef_apply_v1_f32_nofma(const float *C, const float *Y, float *O, int nSC, int L, int S, int A);
void po_coef_apply_f64(const float *C, const float *Y, double *O, int nSC, int L, int S, int A);
void po_coef_apply_f64d(const double *C, const float *Y, double *O, int nSC, int L, int S, int A);
int po_coef_apply_v1_f32_omp(const float *C, const float *Y, float *O, int nSC, int L, int S, int A,
                             int nthreads);

/* -- Gram product M @ H -- */
void po_gram_f32(const float *M, const float *H, float *G, int nSC, int L, int A);
void po_gram_f64(const float *M, const float *H, double *G, int nSC, int L, int A);

/* -- MMSE coefficients in FP64 -- */
int po_mmse_f64(const float *H, const float *sigma2, double *C, double *noise_inv, double *mu, double *cond,
                int nSC, int L, int A, int unbiased);

/* -- max-log LLR -- */
void po_llr_ref_f64(float pamI, float pamQ, float noise_inv, int nPamBits, float *out);
void po_llr_f32(float pamI, float pamQ, float noise, int nPamBits, float *out);
void po_llr_demap_f16(const uint16_t *soft_half, const float *noise_inv, uint16_t *out, int N, int nPamBits);
void po_llr_demap_ref_f64(const uint16_t *soft_half, const float *noise_inv, float *out, int N, int nPamBits);
void po_llr_demap_f32(const float *Yeq, const float *noise_inv, float *llr, int nSC, int L, int S,
                      int nPamBits);
void po_llr_demap_f64(const double *Yeq, const double *noise_inv, double *llr, int nSC, int L, int S,
                      int nPamBits);
void po_llr_bruteforce_f64(double x, double noise, int nPamBits, double *out);
double po_llr_lut(int idx);

/* -- CPU baseline of the fused chain stage (coef apply + demap), OpenMP -- */
int po_coef_apply_demap_f32_omp(const float *C, const float *Y, const float *noise_inv, float *llr,
                                int nSC, int L, int S, int A, int nPamBits, int nthreads);

int po_omp_max_threads(void);

/* -- binary16 arithmetic, one rounding to nearest even per operation (__hmul/__hadd/__hsub/__hfma) -- */
uint16_t po_hmul(uint16_t a, uint16_t b);
uint16_t po_hadd(uint16_t a, uint16_t b);
uint16_t po_hsub(uint16_t a, uint16_t b);
uint16_t po_hfma(uint16_t a, uint16_t b, uint16_t c);
void po_half_ops(const uint16_t *a, const uint16_t *b, const uint16_t *c, uint16_t *mul, uint16_t *add,
                 uint16_t *sub, uint16_t *fma_, size_t n);

/* -- SRS filter multiply (srs_wmma_bench.cu) -- */
void po_srs_filter_v1(uint16_t *H_est, const uint16_t *SRS, const uint16_t *W, const uint16_t *FOCC, int nAnt,
                      int nBlocks, int K, int nPorts, int foccLen, int conj_w);
void po_srs_filter_f64(double *H_est, const uint16_t *SRS, const uint16_t *W, const uint16_t *FOCC, int nAnt,
                       int nBlocks, int K, int nPorts, int foccLen, int conj_w, int prod_f16);
void po_srs_bench_inputs(unsigned seed, int nAnt, int nBlocks, int K, int nPorts, int foccLen, uint16_t *SRS,
                         uint16_t *W, uint16_t *FOCC);
/* -- fronthaul block floating point (O-RAN U-plane) -- */
void po_bfp_decompress(const uint8_t *in, float *Y, int nPrb, int nSym, int nAnt, int iqWidth, size_t row_stride,
                       float scale);
void po_bfp_compress(const int32_t *x, uint8_t *out, int nPrb, int nSym, int nAnt, int iqWidth, size_t row_stride);
/* -- DMRS frequency interpolation (README.md:200, named without code) -- */
void po_freq_interp_f32(const float *W, const float *X, float *Hf, int nBatch, int M, int K, int NC);
int po_freq_interp_f32_omp(const float *W, const float *X, float *Hf, int nBatch, int M, int K, int NC,
                           int nthreads);
void po_freq_interp_f64(const double *W, const double *X, double *Hf, int nBatch, int M, int K, int NC);
/* -- descrambling and rate-recovery layout of the demapper output (TS 38.211 5.2.1, 6.3.1; TS 38.212 5.4.2) -- */
void po_gold_sequence(uint32_t c_init, int n_bits, uint8_t *c);
int po_cw_block_sizes(int G, int nu, int qm, int C, int *E);
void po_cw_layout(const float *llr, int nSC, int L, int S, int nPamBits, const int *layer_cw, int nCodewords,
                  const uint8_t *const *c, const int *nCodeBlocks, float *const *out);
#endif
