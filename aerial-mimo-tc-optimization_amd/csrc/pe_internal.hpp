// pe_internal.hpp -- host-side launchers shared between the kernel translation units and the C ABI layer.
// Arguments are validated by the caller (pe_api.cpp); launchers only pick an instantiation and launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace pe {

// Tile parameters chosen per shape (defaults, overridable by the tuned table).
struct CoefTile {
    int f32_kc = 8;         // FP32 kernel: antennas per LDS chunk (4, 8 or 16)
    int mfma_waves = 1;     // MFMA kernel: waves (subcarriers) per workgroup (1, 2, 4)
    int mfma_spw = 1;       // MFMA kernel: subcarriers per one-wave workgroup (1 = no prefetch loop)
};
CoefTile coef_tile_for(int nLayer, int nSym, int nAnt, int prec);

// MMSE: lanes per subcarrier (4, 8 or 16) wherever antennas are summed. It fixes the Gram summation order, so
// it depends on the shape (nLayer, nAnt) only, never on the batch size: every path (single-cell and
// multi-cell) solving a subcarrier of that shape produces the same bits.
int mmse_lanes_for(int nLayer, int nAnt);
// MMSE FP32 pass launch shape for a batch of nSC subcarriers (changes no result): a workgroup of `waves`
// waves solves a tile of `tile` subcarriers (tile <= 64 waves).
struct MmseTile {
    int tile = 4;
    int waves = 1;
    int keep = 0;  // 1: antenna rows stay in registers from the Gram to the coefficients
};
MmseTile mmse_tile_for(int nLayer, int nAnt, int nSC);
// Tile of the one-pass chain's solve kernel (Gram and algebra only, no coefficient phase): a "solve" line of the
// table, else for a single cell (nSC <= 4096) the coefficient kernel's tile and for batches 64 subcarriers on 4
// waves, so that the one-lane-per-subcarrier algebra runs on a full wave (measured on 32 cells of 8 layers x 64
// antennas: 47.7 us against 63.3 us with the coefficient kernel's 8 on 2 waves).
MmseTile solve_tile_for(int nLayer, int nAnt, int nSC);
// Multi-cell executor tiles (zero: executor default).
struct SlotTile {
    int mmse_tile = 0;   // subcarriers per MMSE tile
    int mmse_waves = 0;  // waves per MMSE workgroup
    int mmse_keep = 0;   // MMSE keeps antenna rows in registers where the tile allows it
};
SlotTile slot_tile();

// Alignment the kernels' widest accesses need (bytes). C, Y_rx, C16 and H with an even layer count are read
// 16 bytes at a time; an LLR buffer is written 2 nPamBits values per symbol in the widest stores that fit.
inline bool is_aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }
inline size_t llr_align(int nPamBits, bool f16) {
    if (f16) return nPamBits == 4 ? 16 : nPamBits == 2 ? 8 : 4;
    return (nPamBits == 2 || nPamBits == 4) ? 16 : 8;
}
inline size_t h_align(int nLayer) { return nLayer % 2 == 0 ? 16 : 8; }

// Compute units of the current device (cached per process).
int device_cus();
// Sets the thread-local message returned by pe_last_error().
void set_error(const char *msg);

// FP32 antenna-serial coefficient apply.
hipError_t launch_coef_apply_f32(const float2 *C, const float2 *Y, float2 *O, int nSC, int L, int S, int A,
                                 hipStream_t st);
// Gram product G = M @ H (FP32 antenna-serial chain).
hipError_t launch_gram_f32(const float2 *M, const float2 *H, float2 *G, int nSC, int L, int A, hipStream_t st);
// MFMA coefficient apply; llr == nullptr -> writes Y_eq, else fused demap into llr (Y_eq unused).
// llr_f16 != 0: llr is binary16 storage (the LLR wire format), same element layout.
hipError_t launch_coef_apply_mfma(const float2 *C, const float2 *Y, float2 *O, const float *noise_inv,
                                  float *llr, int nSC, int L, int S, int A, int nPamBits, int prec,
                                  hipStream_t st, int llr_f16 = 0);
// FP32 fused apply + demap with the antenna-serial chain (the PE_PREC_F32 fused path).
hipError_t launch_coef_apply_demap_f32(const float2 *C, const float2 *Y, const float *noise_inv, float *llr,
                                       int nSC, int L, int S, int A, int nPamBits, hipStream_t st,
                                       int llr_f16 = 0);
// Demappers.
hipError_t launch_llr_f16(const uint32_t *soft, const float *noise_inv, uint16_t *out, int N, int nPamBits,
                          hipStream_t st);
hipError_t launch_llr_f32(const float2 *Yeq, const float *noise_inv, float *llr, int nSC, int L, int S,
                          int nPamBits, hipStream_t st);
// SRS filter multiply (binary16 pairs as uint32): the binary16 chain, or MFMA with FP32 accumulation.
hipError_t launch_srs_filter(uint32_t *H, const uint32_t *SRS, const uint32_t *W, const uint32_t *FOCC, int nAnt,
                             int nBlocks, int K, int nPorts, int foccLen, bool conj, bool mfma, hipStream_t st);
// Binary16 coefficient blocks (C16 pairs + per-subcarrier exponent): conversion from complex64, and the
// binary16 MFMA apply from them (llr == nullptr -> Y_eq into O; llr_f16 -> binary16 LLRs).
hipError_t launch_coef_to_c16(const float2 *C, uint32_t *C16, int *cexp, int nSC, int L, int A, hipStream_t st);
hipError_t launch_coef_apply_c16(const uint32_t *C16, const int *cexp, const float2 *Y, float2 *O,
                                 const float *noise_inv, void *llr, int nSC, int L, int S, int A, int nPamBits,
                                 int llr_f16, hipStream_t st);
// DMRS frequency interpolation H_f[b][m][n] = sum_k W[m][k] X[b][k][n]; prec 0 (FP32 chain), 2 (bf16 MFMA),
// 4 (bf16 split MFMA). 1 <= K <= 128, 1 <= NC <= 16; FP32 needs freq_interp_f32_lds_bytes() <= 64 KB.
size_t freq_interp_f32_lds_bytes(int M, int K, int NC);
hipError_t launch_freq_interp(const float2 *W, const float2 *X, float2 *Hf, int nBatch, int M, int K, int NC,
                              int prec, hipStream_t st);
// Fronthaul block floating point rows -> Y_rx [sc][sym][ant] (iqWidth 1..16).
hipError_t launch_bfp_decompress(const uint8_t *in, float2 *Y, int nPrb, int nSym, int nAnt, int iqWidth,
                                 size_t row_stride, float scale, hipStream_t st);
// Streaming probes (HBM ceiling): read_only -> reads src, else copies src to dst.
hipError_t launch_hbm_probe(const void *src, void *dst, size_t bytes, int read_only, int grid, hipStream_t st);
// MMSE in one launch: FP32 tile pass; the subcarriers it flags are solved in FP64 by the same workgroup.
hipError_t launch_gram_mmse(const float2 *H, const float *sigma2, float2 *C, float *noise_inv, uint8_t *flags,
                            int nSC, int L, int A, bool unbiased, float cond_threshold, hipStream_t st);
// The same with the coefficients written in the binary16 operand format (A <= 128).
hipError_t launch_gram_mmse_c16(const float2 *H, const float *sigma2, uint32_t *C16, int *cexp, float *noise_inv,
                                uint8_t *flags, int nSC, int L, int A, bool unbiased, float cond_threshold,
                                hipStream_t st);
// MMSE + fused MFMA apply + demap without the coefficients in HBM (prec 1..4; C written for the subcarriers solved
// in FP64, or for all with write_c; llr_f16 -> binary16 LLRs).
hipError_t launch_mmse_apply_demap(const float2 *H, const float *sigma2, const float2 *Y, float2 *C,
                                   float *noise_inv, uint8_t *flags, void *llr, int nSC, int L, int S, int A,
                                   int nPamBits, int prec, bool unbiased, float thr, int llr_f16, int write_c,
                                   hipStream_t st);
// The apply of the one-pass chain alone: coefficients formed from the solve records in C (FP64-solved subcarriers:
// their blocks in C) and the antenna rows of H (4 <= L <= 8, 32 <= A <= 64).
hipError_t launch_apply_records(const float2 *H, const float2 *Y, float2 *C, const float *noise_inv,
                                const uint8_t *flags, float *llr, int nSC, int L, int S, int A, int nPamBits, int prec,
                                int llr_f16, int write_c, hipStream_t st);
// DMRS frequency interpolation (64 x K filter W, 4 layers) fused with the MMSE solve, then the records apply: X
// [nBlk][A][K][4] -> H [64 nBlk][A][4], records in C, noise_inv, flags, LLRs (interp_prec 2: bf16, 4: bf16 split).
hipError_t launch_interp_mmse_apply_demap(const float2 *W, const float2 *X, const float *sigma2, const float2 *Y,
                                          float2 *H, float2 *C, float *noise_inv, uint8_t *flags, void *llr, int nBlk,
                                          int K, int S, int A, int nPamBits, int interp_prec, int prec, bool unbiased,
                                          float thr, int llr_f16, int write_c, hipStream_t st);
// Codeword output of the fused demapper (descrambling and the rate-recovery layout). Record of layer l, 32 bytes,
// eight per launch: its codeword's LLR array and scrambling sequence, the codeword's layer count nu and the layer's
// rank v in it, and the code-block split (blocks 0 .. r0 - 1 of e_lo bits, the rest of e_lo + nu Qm bits).
struct CwLayer {
    unsigned long long out;   // the codeword's LLR array (float or binary16)
    unsigned long long cseq;  // c(0 ..) of the codeword, bit g % 32 of word g / 32; 0: no descrambling
    uint32_t e_lo;            // 0: codeword order (no de-interleaving)
    uint32_t nu_v;            // nu | v << 8
    uint32_t r0;
    uint32_t pad;
};
// TS 38.211 5.2.1 Gold sequence c(0 .. n_bits - 1) for c_init, packed (bit g % 32 of words[g / 32]).
hipError_t launch_gold_sequence(uint32_t c_init, uint32_t *words, int n_bits, hipStream_t st);
// Fused MFMA apply + demap (prec 1..4) with the LLRs descrambled and laid out per codeword.
hipError_t launch_coef_apply_demap_cw(const float2 *C, const float2 *Y, const float *noise_inv,
                                      const CwLayer (&layers)[8], int nSC, int L, int S, int A, int nPamBits, int prec,
                                      int llr_f16, hipStream_t st);
// Device address of the file-scope scratch array that pe_mmse_apply_demap_noalloc uses as its coefficient
// workspace (PE_NOALLOC_SCRATCH_BYTES), on the current device: looked up once per device with
// hipGetSymbolAddress and kept; no allocation.
hipError_t noalloc_scratch(void **ptr);

}  // namespace pe
