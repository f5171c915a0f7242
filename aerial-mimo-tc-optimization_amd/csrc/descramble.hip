// descramble.hip -- the LLRs in the order the LDPC decoder's rate recovery takes them: descrambled with the
// TS 38.211 5.2.1 Gold sequence and laid out per codeword (TS 38.211 6.3.1), optionally bit de-interleaved per code
// block (TS 38.212 5.4.2.2). The reference's demapper stops at the per-symbol order [b0_I, b0_Q, ...]
// (aerial_llr_bench_v2.cu:117,221-227).
//
//   gold_kernel: c(n) = x1(n + 1600) ^ x2(n + 1600) packed 32 bits per word. A thread writes a run of words: it
//     jumps both 31-bit registers to the run's start with precomputed powers T^(2^k) of their one-step matrices
//     (GF(2), built at compile time), then advances a word at a time: 28 bits of the next word follow from the
//     previous word by shifts (x1(n) = x1(n-28) ^ x1(n-31), x2(n) = x2(n-28) ^ x2(n-29) ^ x2(n-30) ^ x2(n-31)),
//     the last 4 from the 28 just made. The sequence depends only on c_init (n_RNTI 2^15 + n_ID for PUSCH), so a
//     receiver makes it once per UE and reuses it every slot.
//   apply_cw_kernel: the fused MFMA apply + demap (apply_device.hpp, one subcarrier per wave) with the codeword
//     epilogue: each LLR is negated where its codeword bit of c is set and stored at its codeword position. Same
//     LLR bytes as the per-symbol layout, plus one bit of c per LLR read.
#include "apply_device.hpp"
#include "pe_internal.hpp"

namespace pe {

namespace {

using namespace mfma;

// ---- Gold sequence ---------------------------------------------------------------------------------------------
// Register state: bit i = x(p + i), i < 31. One step: s' = s >> 1 | parity(s & taps) << 30.
constexpr uint32_t kTaps1 = 0x9u;  // x1(n + 31) = x1(n + 3) ^ x1(n)
constexpr uint32_t kTaps2 = 0xFu;  // x2(n + 31) = x2(n + 3) ^ x2(n + 2) ^ x2(n + 1) ^ x2(n)
constexpr int kNc = 1600;          // = 50 words

struct Jump {
    uint32_t m[31][31];  // m[k][i]: column i of T^(2^k)
};

constexpr uint32_t parity32(uint32_t x) {
    x ^= x >> 16;
    x ^= x >> 8;
    x ^= x >> 4;
    x ^= x >> 2;
    x ^= x >> 1;
    return x & 1u;
}
constexpr uint32_t step_state(uint32_t s, uint32_t taps) { return (s >> 1) | (parity32(s & taps) << 30); }
constexpr uint32_t apply_cols(const uint32_t (&cols)[31], uint32_t s) {
    uint32_t r = 0;
    for (int i = 0; i < 31; ++i)
        if ((s >> i) & 1u) r ^= cols[i];
    return r;
}
constexpr Jump make_jump(uint32_t taps) {
    Jump j{};
    for (int i = 0; i < 31; ++i) j.m[0][i] = step_state(1u << i, taps);
    for (int k = 1; k < 31; ++k)
        for (int i = 0; i < 31; ++i) j.m[k][i] = apply_cols(j.m[k - 1], j.m[k - 1][i]);
    return j;
}

__constant__ Jump kJump1 = make_jump(kTaps1);
__constant__ Jump kJump2 = make_jump(kTaps2);

// state at position p of the register started from s0 (p < 2^31)
__device__ __forceinline__ uint32_t jump_state(const Jump &J, uint32_t s0, uint32_t p) {
    uint32_t s = s0;
    for (int k = 0; p; ++k, p >>= 1) {
        if (!(p & 1u)) continue;
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < 31; ++i) r ^= ((s >> i) & 1u) ? J.m[k][i] : 0u;
        s = r;
    }
    return s;
}

// The word after prev (x(32k .. 32k + 31) from x(32k - 32 .. 32k - 1)).
template <bool X2>
__device__ __forceinline__ uint32_t next_word(uint32_t prev) {
    uint32_t w = X2 ? ((prev >> 4) ^ (prev >> 3) ^ (prev >> 2) ^ (prev >> 1)) : ((prev >> 4) ^ (prev >> 1));
    unsigned long long u = ((unsigned long long)(w & 0x0FFFFFFFu) << 32) | prev;
#pragma unroll
    for (int j = 28; j < 32; ++j) {
        const int b = 32 + j;
        unsigned long long bit = (u >> (b - 28)) ^ (u >> (b - 31));
        if (X2) bit ^= (u >> (b - 29)) ^ (u >> (b - 30));
        u |= (bit & 1ull) << b;
    }
    return (uint32_t)(u >> 32);
}

constexpr int kGoldWordsPerThread = 16;

__global__ __launch_bounds__(256) void gold_kernel(uint32_t c_init, uint32_t *__restrict__ out, int n_words,
                                                   int n_bits) {
    const int q0 = (blockIdx.x * blockDim.x + threadIdx.x) * kGoldWordsPerThread;
    if (q0 >= n_words) return;
    // registers at x position 32 (q0 + 49): the word before c word q0 (x offset Nc = 50 words)
    const uint32_t p = 32u * (uint32_t)(q0 + kNc / 32 - 1);
    uint32_t s1 = jump_state(kJump1, 1u, p);
    uint32_t s2 = jump_state(kJump2, c_init & 0x7FFFFFFFu, p);
    uint32_t prev1 = s1 | (parity32(s1 & kTaps1) << 31);
    uint32_t prev2 = s2 | (parity32(s2 & kTaps2) << 31);
    const int q1 = min(q0 + kGoldWordsPerThread, n_words);
    for (int q = q0; q < q1; ++q) {
        prev1 = next_word<false>(prev1);
        prev2 = next_word<true>(prev2);
        uint32_t w = prev1 ^ prev2;
        const int rest = n_bits - 32 * q;  // bits of this word inside the sequence
        if (rest < 32) w &= (1u << rest) - 1u;
        out[q] = w;
    }
}

// ---- fused apply + demap into codewords --------------------------------------------------------------------------
struct CwArgs {
    CwLayer layer[8];
};
static_assert(sizeof(CwArgs) == 256, "one 4-byte word per lane of a wave");

constexpr int kCwWaves = 4;

template <int KS, int PREC>
__global__ __launch_bounds__(64 * kCwWaves) void apply_cw_kernel(const float2 *__restrict__ C,
                                                                 const float2 *__restrict__ Y,
                                                                 const float *__restrict__ noise_inv, CwArgs cw,
                                                                 int nSC, int L, int S, int A, int npb, int f16) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_all[kCwWaves * LdsLayout<KS>::bytes(PREC)];
    __shared__ __attribute__((aligned(16))) CwLayer rec_all[kCwWaves][8];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sc = (long)blockIdx.x * kCwWaves + wave;
    if (sc >= nSC) return;  // uniform per wave; no workgroup barrier follows
    unsigned char *lds = lds_all + wave * LdsLayout<KS>::bytes(PREC);
    // the layer records into this wave's LDS (lane i copies word i); read there with per-lane layer indices
    reinterpret_cast<uint32_t *>(rec_all[wave])[lane] = reinterpret_cast<const uint32_t *>(&cw)[lane];
    if (A < 16 * KS) zero_operands<KS, PREC>(lds, lane);
    lds_order_wave();
    const int nC = L * (A >> 1), nY = S * (A >> 1);
    const int rc = (nC + 63) >> 6, ry = (nY + 63) >> 6;
    RawSC<KS> ra;
    load_raw<KS>(ra, C, Y, (size_t)sc, L, S, A, lane, nC, nY, rc, ry);
    process_sc<KS, PREC, -1, false, false, true>(ra, (size_t)sc, lds, nullptr, noise_inv, nullptr, L, S, A, lane, nC,
                                                 nY, rc, ry, npb, f16, 0, nullptr, rec_all[wave], nSC);
}

template <int KS>
hipError_t launch_cw_k(const float2 *C, const float2 *Y, const float *ni, const CwArgs &cw, int nSC, int L, int S,
                       int A, int npb, int prec, int f16, hipStream_t st) {
    const dim3 g((unsigned)((nSC + kCwWaves - 1) / kCwWaves)), b(64 * kCwWaves);
    switch (prec) {
    case 1: hipLaunchKernelGGL((apply_cw_kernel<KS, 1>), g, b, 0, st, C, Y, ni, cw, nSC, L, S, A, npb, f16); break;
    case 2: hipLaunchKernelGGL((apply_cw_kernel<KS, 2>), g, b, 0, st, C, Y, ni, cw, nSC, L, S, A, npb, f16); break;
    case 3: hipLaunchKernelGGL((apply_cw_kernel<KS, 3>), g, b, 0, st, C, Y, ni, cw, nSC, L, S, A, npb, f16); break;
    case 4: hipLaunchKernelGGL((apply_cw_kernel<KS, 4>), g, b, 0, st, C, Y, ni, cw, nSC, L, S, A, npb, f16); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gold_sequence(uint32_t c_init, uint32_t *words, int n_bits, hipStream_t st) {
    const int n_words = (n_bits + 31) / 32;
    const int threads = (n_words + kGoldWordsPerThread - 1) / kGoldWordsPerThread;
    hipLaunchKernelGGL(gold_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, c_init, words, n_words,
                       n_bits);
    return hipGetLastError();
}

hipError_t launch_coef_apply_demap_cw(const float2 *C, const float2 *Y, const float *noise_inv,
                                      const CwLayer (&layers)[8], int nSC, int L, int S, int A, int nPamBits, int prec,
                                      int llr_f16, hipStream_t st) {
    CwArgs cw;
    for (int l = 0; l < 8; ++l) cw.layer[l] = layers[l];
    if (A <= 16) return launch_cw_k<1>(C, Y, noise_inv, cw, nSC, L, S, A, nPamBits, prec, llr_f16, st);
    if (A <= 32) return launch_cw_k<2>(C, Y, noise_inv, cw, nSC, L, S, A, nPamBits, prec, llr_f16, st);
    if (A <= 64) return launch_cw_k<4>(C, Y, noise_inv, cw, nSC, L, S, A, nPamBits, prec, llr_f16, st);
    if (A <= 128) return launch_cw_k<8>(C, Y, noise_inv, cw, nSC, L, S, A, nPamBits, prec, llr_f16, st);
    return hipErrorInvalidValue;
}

}  // namespace pe
