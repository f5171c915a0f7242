/*
 * pusch_eq.h -- C ABI of the MI355X (gfx950) PUSCH receiver equalization library.
 *
 * The library implements the equalization hot path of a 5G NR PUSCH receiver:
 *
 *     H --gram_mmse--> C, noise_inv --coef_apply--> Y_eq --llr_demap--> LLR
 *                                  \_________ coef_apply_demap (fused) ___/
 *
 * Tensor layouts are the ones the reference kernels use (complex64 = interleaved {re, im} float pair):
 *
 *     C     [nSC][nLayer][nAnt]        coefficients           (aerial_coef_wmma.cu:72)
 *     Y_rx  [nSC][nSym][nAnt]          received samples       (aerial_coef_wmma.cu:73)
 *     Y_eq  [nSC][nLayer][nSym]        equalized symbols      (aerial_coef_wmma.cu:74)
 *     H     [nSC][nAnt][nLayer]        channel estimate       (aerial_bottleneck.cu:58)
 *     M     [nSC][nLayer][nAnt]        Gram left operand      (aerial_bottleneck.cu:57)
 *     G     [nSC][nLayer][nLayer]      Gram product           (aerial_bottleneck.cu:59)
 *     llr   [nSC][nLayer][nSym][2*q]   max-log LLRs, bit order b0(I) b1(Q) b2(I) ... (aerial_llr_bench_v2.cu:221-227)
 *
 * Conventions (all entry points):
 *   - every pointer is caller-owned device memory; nothing here allocates except pe_slot_plan_create and
 *     pe_serve_run, which say so;
 *   - work is enqueued on `stream` (a hipStream_t passed as void*, NULL = the null stream) with no implicit
 *     synchronisation; calls on different streams are independent;
 *   - errors come back as pe_status codes with a message in pe_last_error() (thread-local); nothing calls
 *     exit() (the reference's CUDA_CHECK does, aerial_coef_wmma.cu:57-58);
 *   - nSC == 0 is a valid empty call and returns PE_OK without launching anything;
 *   - alignment: every buffer aligned to its element size, and further to the widest access the kernels make:
 *     C, Y_rx and C16 16 bytes (read 16 bytes at a time), H 16 bytes for an even nLayer (8 for odd), float LLR
 *     buffers 16 bytes for nPamBits 2 and 4 (8 for 1 and 3), binary16 LLR buffers 16 / 8 / 4 bytes for
 *     nPamBits 4 / 2 / 1 and 3. hipMalloc's 256-byte alignment always qualifies; a misaligned pointer is
 *     rejected with PE_ERR_INVALID (naming the buffer) before anything is launched.
 */
#ifndef PUSCH_EQ_H
#define PUSCH_EQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_ABI_VERSION 1

/* complex64: same bytes as the reference's `struct cf32 { float re, im; }` (aerial_coef_wmma.cu:60) */
typedef struct { float re, im; } pe_c64;
/* complex binary16 (raw IEEE half bits): same bytes as `struct half2c { __half re, im; }`
   (aerial_llr_bench_v2.cu:66) */
typedef struct { uint16_t re, im; } pe_c16;

typedef enum {
    PE_OK = 0,
    PE_ERR_INVALID = 1,      /* bad argument: null pointer, negative size, shape out of range */
    PE_ERR_UNSUPPORTED = 2,  /* valid request this build does not implement (e.g. nSym > 16 on MFMA) */
    PE_ERR_DEVICE = 3,       /* HIP runtime error (launch failure, wrong architecture, ...) */
    PE_ERR_NOMEM = 4,        /* a plan/serving allocation failed */
    PE_ERR_CAPACITY = 5      /* request exceeds a stated fixed capacity */
} pe_status;

/* Arithmetic of the coefficient-apply product. */
typedef enum {
    PE_PREC_F32 = 0,         /* FP32 VALU, antenna-serial fused multiply-add chain in the order of the
                                reference scalar loop (aerial_coef_wmma.cu:96-102); bit-reproducible */
    PE_PREC_F16 = 1,         /* MFMA fp16 x fp16 -> fp32, per-subcarrier power-of-two scaling */
    PE_PREC_BF16 = 2,        /* MFMA bf16 x bf16 -> fp32 */
    PE_PREC_F16_SPLIT = 3,   /* MFMA fp16 hi+lo compensated product (3 MFMAs per k-step) */
    PE_PREC_BF16_SPLIT = 4,  /* MFMA bf16 hi+lo compensated product */
    PE_PREC_F16_C16 = 5      /* slot plans and the serving loop only: PE_PREC_F16 with the cell's coefficients
                                kept in the binary16 operand format (pe_gram_mmse_c16); the cell's C buffer holds
                                C16[nSC][nLayer][nAnt] (4 bytes each) followed by the exponents int32[nSC],
                                nSC * (4 nLayer nAnt + 4) bytes, 16-byte aligned */
} pe_prec;

/* ---- library ------------------------------------------------------------------------------------- */

/* No reference counterpart (the reference has no library). */
const char *pe_version(void);
const char *pe_last_error(void);
const char *pe_status_string(pe_status s);
/* Selects `device`, checks that it is a gfx950 part and loads the tile table (file named by the
   environment variable PE_TUNE_TABLE if set). Calling it is optional; entry points work on the current
   device without it. */
pe_status pe_init(int device);
/* Replaces the tile table with the file at `path` (NULL: no table, built-in defaults). Returns the number of
   coef, mmse and solve entries read (0 for NULL), or -1 if the file cannot be read; malformed lines are
   skipped. Lines, first match wins, -1 = any:
     coef <nLayer> <nSym> <nAnt> <prec> <f32_kc: 4|8|16> <mfma_wpb: 1|2|4> <mfma_spw: 1..8>
     mmse <nLayer> <nAnt> <nSC_max> <lanes_per_subcarrier: 4|8|16> [<subcarriers_per_tile> [<waves: 1|2|4|8>
          [<keep_rows: 0|1>]]]
     solve <nLayer> <nAnt> <nSC_max> <subcarriers_per_tile> <waves: 1|2|4|8>   (pe_mmse_apply_demap's solve kernel)
     slot <mmse_subcarriers_per_tile> [<mmse_waves: 1|2|4|8> [<keep_rows: 0|1>]]   (multi-cell MMSE kernel)
   The lanes of an mmse line set the summation order of the Gram matrix and are taken from the first line
   matching (nLayer, nAnt) whatever nSC_max, so every entry point computes the same coefficients for a
   subcarrier; tile sizes and waves change only speed. Without an explicit table the library reads
   tune_table.txt beside itself (written by autotune.py). */
int pe_tune_load(const char *path);

/* ---- coefficient apply:  Y_eq[sc][l][s] = sum_a C[sc][l][a] * Y_rx[sc][s][a]  (no conjugation) ------ */

/* General form. Replaces coef_apply_v1_scalar (aerial_coef_wmma.cu:71-75, nAnt/nLayer compile-time there),
   coef_apply_v2_wmma (aerial_coef_wmma.cu:146-150), coef_apply_v3_wmma_multiw (aerial_coef_wmma.cu:250-254),
   v1/v2 (coef_rigor.cu:38,54) and v1_scalar/v2_wmma<N_ANT,N_SYM> (coef_scaling.cu:41,77).
   Shapes: 1 <= nLayer <= 8, 1 <= nSym <= 16, nAnt % 4 == 0 with 4 <= nAnt <= 256 (PE_PREC_F32) or
   4 <= nAnt <= 128 (MFMA precisions). */
pe_status pe_coef_apply(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq,
                        int nSC, int nLayer, int nSym, int nAnt, pe_prec prec, void *stream);

/* The reference's exact argument list (nLayer = 8, nAnt = 64 fixed as in aerial_coef_wmma.cu:42-47),
   FP32 path. Replaces coef_apply_v1_scalar(C, Y_rx, Y_eq, nSC, nSym) (aerial_coef_wmma.cu:71-75). */
pe_status pe_coef_apply_v1(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC, int nSym,
                           void *stream);

/* Subcarrier range [scOffset, scOffset + nSC) of full-size tensors. Replaces k_coef_offset
   (adaptive_smpartition.cu:52-54). */
pe_status pe_coef_apply_range(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC, int scOffset,
                              int nLayer, int nSym, int nAnt, pe_prec prec, void *stream);

/* ---- Gram product G[sc] = M[sc] @ H[sc] (plain product, no conjugation, FP32 antenna-serial chain) -- */

/* Replaces gram_scalar(M, H, G, nSC) (aerial_bottleneck.cu:56-60). 1 <= nLayer <= 8, 1 <= nAnt <= 1024. */
pe_status pe_gram(const pe_c64 *M, const pe_c64 *H, pe_c64 *G, int nSC, int nLayer, int nAnt, void *stream);

/* ---- MMSE coefficient compute ---------------------------------------------------------------------- */

#define PE_MMSE_UNBIASED 1u    /* scale row l of C by 1/mu_l (unbiased estimate, the LLR convention) */

#define PE_SCFLAG_FP64 1u      /* subcarrier solved in FP64 (condition estimate above threshold) */
#define PE_SCFLAG_SINGULAR 2u  /* Gram not positive definite even in FP64: C and noise_inv are zero */

/* Per subcarrier: G = H^H H + sigma2 I (nLayer x nLayer Hermitian), Cholesky G = L L^H, Ginv = G^-1,
   W = Ginv H^H (the MMSE matrix), mu_l = 1 - sigma2 * Ginv[l][l], noise_inv_l = mu_l / (1 - mu_l)
   (post-equalization SINR of the unbiased estimate). C = W, or diag(1/mu) W with PE_MMSE_UNBIASED.
   The solve runs in FP32; subcarriers whose 1-norm condition number kappa_1(G) exceeds cond_threshold
   (<= 0 selects 1e4), or whose FP32 Cholesky breaks down, are recomputed in FP64 by the same kernel.
   The reference only names this stage (aerial_bottleneck.cu:51-53: eqMmseCoefCompHighMimoKernel).
   sigma2[nSC] > 0; noise_inv[nSC][nLayer]; sc_flags[nSC] (required). 1 <= nLayer <= 8,
   1 <= nAnt <= 1024. */
pe_status pe_gram_mmse(const pe_c64 *H, const float *sigma2, pe_c64 *C, float *noise_inv, uint8_t *sc_flags,
                       int nSC, int nLayer, int nAnt, unsigned flags, float cond_threshold, void *stream);

/* ---- max-log LLR soft demapper (computePamLlr) ----------------------------------------------------- */

/* Drop-in for llr_v1_baseline(softEst, noiseInv, llrOut, N, nPamBits) (aerial_llr_bench_v2.cu:114-118,
   aerial_llr_v3.cu:84-86): FP16 complex in, FP16 LLR out [N][2*nPamBits], noiseInv per symbol,
   1 <= nPamBits <= 4 (QPSK..256-QAM). FP32 arithmetic, results bit-identical to the reference formula. */
pe_status pe_llr_demap_f16(const pe_c16 *softEst, const float *noiseInv, uint16_t *llrOut, int N,
                           int nPamBits, void *stream);

/* FP32 demapper on the equalizer output: Y_eq[nSC][nLayer][nSym], noise_inv[nSC][nLayer]
   -> llr[nSC][nLayer][nSym][2*nPamBits] float. */
pe_status pe_llr_demap(const pe_c64 *Y_eq, const float *noise_inv, float *llr, int nSC, int nLayer, int nSym,
                       int nPamBits, void *stream);

/* ---- fused coefficient apply + demap: Y_eq stays in registers, only LLRs are written ---------------- */

/* No single reference kernel: the reference runs coef apply and demap as two launches with Y_eq in HBM
   between them (pipeline_overlap.cu:293-301). Same shapes as pe_coef_apply plus noise_inv[nSC][nLayer]. */
pe_status pe_coef_apply_demap(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, float *llr,
                              int nSC, int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec,
                              void *stream);

/* Same as pe_coef_apply_demap, with the LLRs written in the LDPC decoder's wire format: binary16
   llr[nSC][nLayer][nSym][2*nPamBits], interleaved [b0_I, b0_Q, b1_I, ...], clamped to +-65504 and rounded
   to nearest even, as the reference demapper stores them (aerial_llr_bench_v2.cu:117,221-227). Each value
   is the binary16 rounding of the FP32 LLR pe_coef_apply_demap writes for the same inputs and precision.
   llr must be 16-byte aligned (8-byte for nPamBits 2, 4-byte for 1 and 3). */
pe_status pe_coef_apply_demap_f16(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, uint16_t *llr,
                                  int nSC, int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec,
                                  void *stream);

/* ---- descrambling and the rate-recovery layout of the LLRs ------------------------------------------------ */

/* The steps after the reference's demapper output ([b0_I, b0_Q, b1_I, ...] per symbol, aerial_llr_bench_v2.cu:117,
   221-227), which the reference does not implement; fixed by the 3GPP specifications:
   pe_gold_sequence: the TS 38.211 5.2.1 pseudo-random sequence c(0 .. n_bits - 1) for c_init (PUSCH: n_RNTI 2^15
   + n_ID, 6.3.1.1), c(n) = x1(n + 1600) ^ x2(n + 1600), packed: bit g % 32 of c_words[g / 32] = c(g); the unused
   bits of the last word are zero. 1 <= n_bits < 2^30; c_words 4-byte aligned. It depends on c_init only: make it
   once per UE and reuse it every slot. */
pe_status pe_gold_sequence(uint32_t c_init, uint32_t *c_words, int n_bits, void *stream);

/* Where pe_coef_apply_demap_cw puts the LLRs. Codeword k holds the layers l with layer_cw[l] == k (layer mapping of
   TS 38.211 6.3.1.3: its nu_k layers, in increasing l, carry codeword symbols m = nu_k i + v at layer symbol i,
   v = the layer's rank among them), codewords numbered 0 .. nCodewords - 1 with every number used. Resource
   elements are taken frequency first (6.3.1.6): layer symbol i = s nSC + sc for data symbol s and subcarrier sc
   of the nSC-subcarrier allocation. Bit i of codeword symbol m is codeword bit g = m Qm + i (Qm = 2 nPamBits),
   G_k = nSC nSym nu_k Qm bits per codeword. */
typedef struct {
    int layer_cw[8];           /* codeword of each of the nLayer layers */
    void *llr[8];              /* per codeword: G_k LLRs (float, or binary16 with PE_LLR_F16), aligned as the llr
                                  of pe_coef_apply_demap / pe_coef_apply_demap_f16 */
    const uint32_t *c_seq[8];  /* per codeword: c(0 .. G_k - 1) as pe_gold_sequence packs it (4-byte aligned), or
                                  NULL: no descrambling */
    int nCodeBlocks[8];        /* per codeword: 0 = codeword order (LLR of bit g at g); C >= 1 = the codeword's C
                                  code blocks of TS 38.212 5.4.2.1 (E_r = nu Qm floor(G / (nu Qm C)) for
                                  r < C - mod(G / (nu Qm), C), else nu Qm ceil(G / (nu Qm C))), each bit de-interleaved
                                  (5.4.2.2 undone: bit t = i + j Qm of block r at e-index i E_r/Qm + j), blocks
                                  concatenated: the order rate recovery consumes. 1 <= C <= nSC nSym. */
} pe_cw_layout;

/* pe_coef_apply_demap (MFMA precisions) with each LLR negated where c(g) = 1 (descrambling, 6.3.1.1 undone; the
   LLR sign convention is positive = bit 0) and stored at its codeword position as pe_cw_layout says. The values are
   those of pe_coef_apply_demap (or, with PE_LLR_F16, pe_coef_apply_demap_f16) for the same inputs, sign-flipped and
   moved: no other arithmetic. G_k < 2^31. flags: PE_LLR_F16. */
pe_status pe_coef_apply_demap_cw(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, const pe_cw_layout *cw,
                                 int nSC, int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec, unsigned flags,
                                 void *stream);

/* ---- coefficients in the binary16 operand format ---------------------------------------------------- */

/* No reference counterpart; the reference names this as future work ("FP16 native data path: eliminate the
   FP32 -> FP16 conversion", README.md:201). The binary16 MFMA product rounds every coefficient to
   binary16 anyway, so the coefficients can be kept in that form: C16[nSC][nLayer][nAnt] binary16 pairs {re, im}
   and one exponent per subcarrier, C = C16 * 2^c_exp[sc]. c_exp is 0 when the subcarrier's largest |Re|, |Im|
   lies in [2^-9, 2^14), else the frexp exponent of that maximum (the scaled block then peaks in [0.5, 1)).
   The fused apply reads 4 bytes per coefficient instead of 8. Where samples and coefficients both lie in
   [2^-9, 2^14) (neither operand needs range scaling), the results are bit-identical to pe_coef_apply /
   pe_coef_apply_demap with PE_PREC_F16 on the complex64 coefficients; elsewhere they differ only by the
   binary16 rounding of scaled versus unscaled values. Shapes as for PE_PREC_F16; C, C16 and Y_rx 16-byte
   aligned. */
#define PE_LLR_F16 1u  /* LLRs written as binary16 (the wire format of pe_coef_apply_demap_f16) */

pe_status pe_coef_to_c16(const pe_c64 *C, pe_c16 *C16, int *c_exp, int nSC, int nLayer, int nAnt, void *stream);
/* pe_gram_mmse writing the coefficients in this format (half the coefficient bytes written and read); the
   values are those pe_coef_to_c16 makes of pe_gram_mmse's output, bit for bit. nAnt % 4 == 0, nAnt <= 128. */
pe_status pe_gram_mmse_c16(const pe_c64 *H, const float *sigma2, pe_c16 *C16, int *c_exp, float *noise_inv,
                           uint8_t *sc_flags, int nSC, int nLayer, int nAnt, unsigned flags, float cond_threshold,
                           void *stream);
pe_status pe_coef_apply_c16(const pe_c16 *C16, const int *c_exp, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC,
                            int nLayer, int nSym, int nAnt, void *stream);
/* llr: float [nSC][nLayer][nSym][2*nPamBits], or binary16 with PE_LLR_F16 */
pe_status pe_coef_apply_demap_c16(const pe_c16 *C16, const int *c_exp, const pe_c64 *Y_rx, const float *noise_inv,
                                  void *llr, int nSC, int nLayer, int nSym, int nAnt, int nPamBits,
                                  unsigned flags, void *stream);

/* ---- MMSE coefficients and the fused apply + demap in one pass ----------------------------------------- */

/* No single reference kernel: the reference's chain is the coefficient compute it names
   (eqMmseCoefCompHighMimoKernel, aerial_bottleneck.cu:51-53), then coefficient apply and demap as separate
   launches with the coefficients and Y_eq in HBM between them (pipeline_overlap.cu:293-301). Here the
   coefficients of each subcarrier go from the MMSE solve straight into the MFMA operands on chip and are not
   written to HBM. The LLRs (and noise_inv, sc_flags) are bit-identical to pe_gram_mmse(H, sigma2, C, noise_inv,
   sc_flags, ..., flags & PE_MMSE_UNBIASED, cond_threshold) followed by pe_coef_apply_demap (or, with
   PE_EQ_LLR_F16, pe_coef_apply_demap_f16) at the same precision.
   C [nSC][nLayer][nAnt] is a workspace whose contents afterwards are unspecified, except with PE_EQ_WRITE_C
   (then C equals pe_gram_mmse's output). The on-chip path serves 32 <= nAnt <= 64 with nLayer >= 4, where it
   measured faster; other shapes run the two kernels with C in HBM (same results). Shapes as
   pe_coef_apply_demap with prec PE_PREC_F16, PE_PREC_BF16 or their split forms (MFMA only); H, Y_rx and llr
   16-byte aligned, C 8-byte aligned. flags: PE_MMSE_UNBIASED, and: */
#define PE_EQ_WRITE_C 2u  /* write every coefficient to C as well */
#define PE_EQ_LLR_F16 4u  /* LLRs in binary16 (the wire format of pe_coef_apply_demap_f16) instead of FP32 */
pe_status pe_mmse_apply_demap(const pe_c64 *H, const float *sigma2, const pe_c64 *Y_rx, pe_c64 *C, float *noise_inv,
                              uint8_t *sc_flags, void *llr, int nSC, int nLayer, int nSym, int nAnt, int nPamBits,
                              pe_prec prec, unsigned flags, float cond_threshold, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under hipGraph capture,
   tight serving loops, harnesses that forbid hipMalloc): pe_mmse_apply_demap without the C workspace argument.
   Its workspace is a fixed array in the library's device memory, PE_NOALLOC_SCRATCH_BYTES per device, set aside
   when the library's code is loaded on that device and held until the process ends; the call itself allocates
   nothing (no hipMalloc, no allocating library routine). nSC * nLayer * nAnt * 8 bytes must fit the array (one
   cell of 273 PRB = 3276 subcarriers at 8 layers x 128 antennas does), else PE_ERR_CAPACITY. Results (llr,
   noise_inv, sc_flags) are bit-identical to pe_mmse_apply_demap with the same arguments. flags:
   PE_MMSE_UNBIASED and PE_EQ_LLR_F16 (PE_EQ_WRITE_C has no C to write: PE_ERR_INVALID). Every call on one device
   shares that device's array: calls whose work overlaps on the device (different streams, or another thread)
   race on it and must be serialized by the caller; calls ordered on one stream are safe. No state carries over
   between calls (each call writes every workspace byte it reads). */
#define PE_NOALLOC_SCRATCH_BYTES (32u << 20)
pe_status pe_mmse_apply_demap_noalloc(const pe_c64 *H, const float *sigma2, const pe_c64 *Y_rx, float *noise_inv,
                                      uint8_t *sc_flags, void *llr, int nSC, int nLayer, int nSym, int nAnt,
                                      int nPamBits, pe_prec prec, unsigned flags, float cond_threshold, void *stream);

/* ---- SRS filter multiply (channel estimation upstream of the equalizer) ---------------------------- */

#define PE_SRS_MFMA 1u    /* MFMA with FP32 accumulation: one binary16 rounding per output (K % 4 == 0,
                             SRS 16-byte aligned) instead of the reference's binary16 chain */
#define PE_SRS_CONJ_W 2u  /* multiply by conj(W[sc][k]) (the operation as srs_wmma_bench.cu:6 writes it); without
                             it W is not conjugated, as the reference's complex_conjmul computes (:43-47) */

/* Drop-in for srs_filter_v1_scalar<K, N_PORTS, FOCC_LENGTH>(H_est, SRS, W, FOCC, nAnt, nBlocks)
   (srs_wmma_bench.cu:57-90; the WMMA variant srs_filter_v2_wmma :102-231 computes the same sum):
     H_est[b][ant][sc][port] = sum_{k<K} (W[sc][k] * FOCC[port][k % foccLen]) * SRS[b][ant][k]
   SRS [nBlocks][nAnt][K], W [K][K], FOCC [nPorts][foccLen], H_est [nBlocks][nAnt][K][nPorts], binary16
   complex. flags = 0: bit-identical to the reference's binary16 arithmetic (each product and sum of W*FOCC
   rounded, then two binary16 fused multiply-adds per part and k, in k order). 1 <= K <= 48,
   1 <= nPorts <= 8, 1 <= foccLen <= 48, 1 <= nAnt <= 4096; H_est 16-byte aligned for nPorts 4 and 8
   (8-byte for 2). */
pe_status pe_srs_filter(pe_c16 *H_est, const pe_c16 *SRS, const pe_c16 *W, const pe_c16 *FOCC, int nAnt,
                        int nBlocks, int K, int nPorts, int foccLen, unsigned flags, void *stream);

/* ---- DMRS frequency interpolation (channel estimation upstream of the equalizer) ------------------- */

/* No reference kernel: the reference names the GEMM [64 x 96] @ [96 x 4] of the channel-estimation frequency
   interpolation, batched over OFDM symbols and UEs (README.md:200), without code. One filter for the batch:
     H_f[b][m][n] = sum_{k < nIn} W[m][k] * X[b][k][n]
   W [nOut][nIn], X [nBatch][nIn][nCol] (estimates on the pilot tones), H_f [nBatch][nOut][nCol], complex64.
   prec PE_PREC_F32: per output one fused complex multiply-add chain in k order (cma() of
   aerial_coef_wmma.cu:60-64), bit-identical to the oracle; needs (nOut * (nIn | 1) + nIn * nCol) * 8 <= 64 KB.
   PE_PREC_BF16: MFMA on bf16-rounded W and X with FP32 accumulation; PE_PREC_BF16_SPLIT: hi + lo bf16 operands
   (Ahi*Bhi + Ahi*Blo + Alo*Bhi). Other precisions: PE_ERR_UNSUPPORTED. 1 <= nOut <= 256, 1 <= nIn <= 128,
   1 <= nCol <= 16; W, X, H_f 8-byte aligned. */
pe_status pe_freq_interp(const pe_c64 *W, const pe_c64 *X, pe_c64 *H_f, int nBatch, int nOut, int nIn, int nCol,
                         pe_prec prec, void *stream);

/* No reference kernel: the chain from the DMRS estimates to LLRs, with the frequency interpolation
   (README.md:199-200 names it as the step upstream of the equalizer, whose fusion with the MMSE removes the round
   trip of H through HBM) fused with the MMSE solve. 4 layers; blocks of 64 output subcarriers:
     X [nBlk][nAnt][nIn][4]   estimates on the nIn pilot tones of block blk, antenna ant (item blk * nAnt + ant of
                              pe_freq_interp), W [64][nIn] the filter,
     H [64 nBlk][nAnt][4]     written: H[64 blk + m][ant][n] = the value pe_freq_interp(W, X, ..., nOut = 64,
                              nCol = 4, interp_prec) writes at [blk * nAnt + ant][m][n], in the equalizer's layout,
   then as pe_mmse_apply_demap(H, sigma2, Y_rx, C, noise_inv, sc_flags, llr, 64 nBlk, 4, nSym, nAnt, nPamBits,
   prec, flags, cond_threshold): C (its workspace, [64 nBlk][4][nAnt]), noise_inv, sc_flags and llr are
   bit-identical to that call on this H (the Gram matrices are summed in the same order). interp_prec
   PE_PREC_BF16 or PE_PREC_BF16_SPLIT; prec as pe_mmse_apply_demap (MFMA precisions); 32 <= nAnt <= 64,
   nAnt % 4 == 0; 1 <= nIn <= 128; H, Y_rx and llr 16-byte aligned, W, X and C 8-byte aligned. */
pe_status pe_interp_mmse_apply_demap(const pe_c64 *W, const pe_c64 *X, const float *sigma2, const pe_c64 *Y_rx,
                                     pe_c64 *H, pe_c64 *C, float *noise_inv, uint8_t *sc_flags, void *llr, int nBlk,
                                     int nIn, int nSym, int nAnt, int nPamBits, pe_prec interp_prec, pe_prec prec,
                                     unsigned flags, float cond_threshold, void *stream);

/* ---- fronthaul input: block floating point U-plane data -> Y_rx ----------------------------------- */

/* No reference kernel: the reference cites the O-RAN WG4 CUS-plane specification (README.md:214) for the
   receiver's input. Block floating point: in holds nSym * nAnt rows (row sym * nAnt + ant starts at
   (sym * nAnt + ant) * row_stride bytes), each nPrb compressed PRBs of 1 + 3 * iqWidth bytes: a udCompParam byte
   whose bits 3..0 are the exponent e, then 24 mantissas I0 Q0 I1 Q1 ... I11 Q11 of iqWidth bits, two's
   complement, most significant bit first. Output Y_rx[nPrb * 12][nSym][nAnt] (the coefficient-apply layout,
   subcarrier = 12 prb + re) = scale * mantissa * 2^e, one rounding (the multiplication by scale).
   1 <= iqWidth <= 16, 1 <= nAnt <= 4096, row_stride >= nPrb * (1 + 3 * iqWidth). */
pe_status pe_bfp_decompress(const uint8_t *in, pe_c64 *Y_rx, int nPrb, int nSym, int nAnt, int iqWidth,
                            size_t row_stride, float scale, void *stream);

/* ---- HBM streaming probe ---------------------------------------------------------------------------- */

/* No reference counterpart: measures the device's streaming ceiling, the yardstick for the bandwidth-bound
   kernels above. read_only = 0: copies `bytes` from src to dst (read + write traffic); read_only != 0: reads
   `bytes` of src and may write to dst[0..grid) (4-byte words). 16-byte nontemporal accesses over `grid`
   workgroups of 256 lanes; bytes % 16 == 0, src and dst 16-byte aligned. */
pe_status pe_hbm_probe(const void *src, void *dst, size_t bytes, int read_only, int grid, void *stream);

/* ---- multi-cell slot: one plan per set of uneven cells, run by tile kernels ------------------------ */

typedef struct {
    /* stage 1 (optional, skipped when H == NULL): MMSE coefficients into C / noise_inv */
    const pe_c64 *H;          /* [nSC][nAnt][nLayer] */
    const float *sigma2;      /* [nSC] */
    /* stage 2: fused apply + demap */
    pe_c64 *C;                /* [nSC][nLayer][nAnt]; written by stage 1, read by stage 2 */
    float *noise_inv;         /* [nSC][nLayer] */
    const pe_c64 *Y_rx;       /* [nSC][nSym][nAnt] */
    float *llr;               /* [nSC][nLayer][nSym][2*nPamBits] */
    int nSC, nLayer, nSym, nAnt, nPamBits;
    int prec;                 /* pe_prec of stage 2: the MFMA precisions (PE_PREC_F16 .. PE_PREC_BF16_SPLIT) */
} pe_cell_desc;

typedef struct pe_slot_plan pe_slot_plan;

/* Builds the tile lists of a slot (cells with H for the MMSE stage, cells of each precision for the fused
   apply/demap, ordered by per-subcarrier cost and cut into subcarrier tiles) and uploads descriptors and
   tiles. Allocates its device memory (descriptors, tiles, the MMSE flag counter, per-subcarrier MMSE flags)
   once, here. flags: PE_MMSE_UNBIASED for the MMSE stage, PE_SLOT_SOLVE_RECORDS (below).
   Shapes per cell: nSC >= 1, 1 <= nLayer <= 8, 1 <= nSym <= 16, 4 <= nAnt <= 128 with nAnt % 4 == 0.
   Replaces the reference's one-launch-per-cell-per-stream protocol (multicell_contention.cu:194-283,
   multicell_evidence.cu:497-566). */
#define PE_SLOT_SOLVE_RECORDS 2u /* stage 1 keeps each subcarrier's compact solve (Ginv and the output scales,
                                   nLayer (nLayer + 2) floats) at the start of its block of C instead of the
                                   coefficients, and stage 2 forms the coefficients on chip from H and that record,
                                   as pe_mmse_apply_demap does. Applies to cells with H, an MFMA precision other
                                   than PE_PREC_F16_C16, 32 <= nAnt <= 64 and nLayer >= 4 (where it measured
                                   faster); LLRs and noise_inv are unchanged bit for bit, C holds coefficients only
                                   where the solve ran in FP64. */
pe_status pe_slot_plan_create(const pe_cell_desc *cells, int nCells, unsigned flags, pe_slot_plan **plan);
/* Rewrites the per-cell pointers of an existing plan (same shapes) for the next slot, ordered on `stream`. */
pe_status pe_slot_plan_update(pe_slot_plan *plan, const pe_cell_desc *cells, int nCells, void *stream);
/* Enqueues the slot on `stream`: the MMSE kernel (FP32, with the flagged subcarriers solved in FP64 by the
   same workgroups) and one fused apply/demap kernel per precision, one workgroup per tile. Graph-capturable:
   no memset node is recorded. Each cell's results are bit-identical to the cell run alone through
   pe_gram_mmse and pe_coef_apply_demap. */
pe_status pe_slot_plan_launch(pe_slot_plan *plan, void *stream);
/* Captures pe_slot_plan_launch into a hipGraph on first use and replays it on `stream`. */
pe_status pe_slot_plan_graph_launch(pe_slot_plan *plan, void *stream);
void pe_slot_plan_destroy(pe_slot_plan *plan);
/* Tiles and largest kernel grid (workgroups) of the plan (for reports). */
pe_status pe_slot_plan_info(const pe_slot_plan *plan, int *nTiles, int *nWorkgroups);

/* ---- serving loop --------------------------------------------------------------------------------- */

typedef struct {
    /* Cell population. Per cell: shapes, stage-2 precision (MFMA only) and the device-resident job inputs
       H, sigma2, Y_rx, reused by every job of that cell. C / noise_inv / llr are ignored: each job's buffers
       come from a device arena of hbm_budget_bytes. */
    const pe_cell_desc *cells;
    int n_cells;
    int n_jobs;                /* jobs to serve, each for a uniformly drawn cell */
    double arrival_rate_hz;    /* Poisson arrival rate over all cells; <= 0: every job arrives at t = 0 */
    double deadline_s;         /* per-job deadline after arrival, e.g. 0.5e-3 (one 30 kHz-SCS slot) */
    size_t hbm_budget_bytes;   /* arena for in-flight job buffers (C, noise_inv and, without job_llr, LLRs) */
    int max_batch_jobs;        /* jobs per batch launch */
    int max_inflight;          /* batches in flight, one stream each, 1..8 */
    int shed_late;             /* 0: serve every job; 1: drop jobs whose deadline passed before launch; 2:
                                  deadline-aware admission, drop jobs whose predicted completion (age + predicted
                                  batch latency: online latency fits by GPU occupancy plus a margin calibrated so
                                  that about 0.5 % of batches hold a late job) is past the deadline. Dropped jobs count as
                                  misses. */
    int use_graph;             /* nonzero: each stream replays one captured slot graph; zero: eager launches,
                                  with the coefficients formed on chip from solve records (as with
                                  PE_SLOT_SOLVE_RECORDS) where the cell shape allows it */
    unsigned seed;             /* arrival process and cell draws */
    float *const *job_llr;     /* optional [n_jobs]: LLR destination of each job (sized for its cell; aligned
                                  as the llr of pe_coef_apply_demap) */
    /* Optional per-job inputs (each array NULL or [n_jobs]); a job's slot then carries its own data instead of
       its cell's, as a real slot does: */
    const int *job_cells;               /* cell of each job, 0 .. n_cells - 1 (NULL: drawn uniformly from seed) */
    const pe_c64 *const *job_H;         /* channel estimate of each job, the shape of its cell's H */
    const float *const *job_sigma2;     /* noise variance per subcarrier of each job */
    const pe_c64 *const *job_Y;         /* samples of each job, the shape of its cell's Y_rx */
} pe_serve_config;

typedef struct {
    int n_jobs, n_done, n_missed, n_shed, n_batches, max_inflight_seen;
    double wall_s;             /* first arrival window to last completion */
    double p50_s, p99_s, max_s, mean_s;   /* latency (completion - arrival) of served jobs; nearest rank */
    double miss_rate;          /* (served late + shed) / n_jobs */
    double jobs_per_s, subcarriers_per_s, mean_batch_jobs, offered_rate_hz;
    double iso_job_s;          /* latency of one job alone on an idle GPU (warm-up, median of 5) */
    double contention_index;   /* mean latency of the first 5 batches / iso_job_s: the online contention
                                  indicator of pmu_contention.cu:230-245 */
    double model_s_per_sc;     /* final per-subcarrier slope of the batch-latency model (shed_late = 2) */
    int n_late;                /* served jobs that completed after their deadline (n_missed = n_late + n_shed) */
    double model_margin_s;     /* final admission margin of the latency model (shed_late = 2), calibrated from
                                  observed lateness */
} pe_serve_report;

/* Runs the serving loop to completion on the current device (blocking). Allocates its arena, streams,
   events and executors, and frees them before returning. Per job (optional arrays of n_jobs): latency in
   seconds (-1 when shed) and the cell index. The reference has no serving loop; its closest protocol is the
   N-cell deadline check of multicell_evidence.cu:497-566. */
pe_status pe_serve_run(const pe_serve_config *cfg, pe_serve_report *rep, double *job_latency_s, int *job_cell);

#ifdef __cplusplus
}
#endif
#endif /* PUSCH_EQ_H */
