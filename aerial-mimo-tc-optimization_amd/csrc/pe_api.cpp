// pe_api.cpp -- the C ABI (include/pusch_eq.h): argument validation, error reporting, tile-table lookup and
// dispatch to the kernel launchers. The reference has no library layer: each benchmark main() launches its
// kernels directly and exits on the first CUDA error (aerial_coef_wmma.cu:57-58); here every failure is a
// status code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "pusch_eq.h"
#include "pe_internal.hpp"
#include "tune_parse.hpp"

namespace {

thread_local std::string g_err;

pe_status fail(pe_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

pe_status ok() {
    g_err.clear();
    return PE_OK;
}

pe_status from_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return ok();
    return fail(PE_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

bool valid_prec(int p) { return p >= PE_PREC_F32 && p <= PE_PREC_BF16_SPLIT; }

// ---- tile table -----------------------------------------------------------------------------------------
// Format and parsing in tune_parse.hpp. Loaded from PE_TUNE_TABLE (pe_init), pe_tune_load(), or else once from
// tune_table.txt beside the library.
std::mutex g_table_mu;
pe::tune::Table g_tune;    // guarded by g_table_mu
bool g_table_set = false;  // a table was loaded or cleared explicitly

int load_table_file(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    pe::tune::Table t;
    const int n = pe::tune::parse(f, t);
    fclose(f);
    std::lock_guard<std::mutex> lk(g_table_mu);
    g_tune = std::move(t);
    g_table_set = true;
    return n;
}

// tune_table.txt beside this library, read once when no table was set explicitly
void load_default_table() {
    static std::once_flag once;
    std::call_once(once, [] {
        {
            std::lock_guard<std::mutex> lk(g_table_mu);
            if (g_table_set) return;
        }
        Dl_info info;
        if (!dladdr(reinterpret_cast<void *>(&load_default_table), &info) || !info.dli_fname) return;
        std::string path(info.dli_fname);
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tune_table.txt";
        (void)load_table_file(path.c_str());
    });
}

bool check_ptrs(const void *a, const void *b, const void *c) { return a && b && c; }

// Alignment contract of an entry point: each buffer aligned to the widest access the kernels make to it
// (pe::is_aligned, pe::llr_align, pe::h_align). Misaligned pointers are rejected before anything is launched.
struct AlignReq {
    const void *p;
    size_t bytes;
    const char *name;
};
pe_status check_align(std::initializer_list<AlignReq> reqs) {
    for (const AlignReq &r : reqs)
        if (r.p && !pe::is_aligned(r.p, r.bytes))
            return fail(PE_ERR_INVALID, "%s must be %zu-byte aligned", r.name, r.bytes);
    return PE_OK;
}

}  // namespace

namespace pe {
void set_error(const char *msg) { g_err = msg ? msg : ""; }

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            return 256;
    }
    return cus;
}

CoefTile coef_tile_for(int L, int S, int A, int prec) {
    load_default_table();
    std::lock_guard<std::mutex> lk(g_table_mu);
    for (const tune::Coef &e : g_tune.coef) {
        if ((e.L < 0 || e.L == L) && (e.S < 0 || e.S == S) && (e.A < 0 || e.A == A) &&
            (e.prec < 0 || e.prec == prec)) {
            CoefTile t;
            t.f32_kc = e.f32_kc;
            t.mfma_waves = e.mfma_waves;
            t.mfma_spw = e.mfma_spw;
            return t;
        }
    }
    return CoefTile();
}

int mmse_lanes_for(int L, int A) {
    load_default_table();
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        for (const tune::Mmse &e : g_tune.mmse)
            if ((e.L < 0 || e.L == L) && (e.A < 0 || e.A == A)) return e.lanes;
    }
    return A >= 64 ? 16 : A >= 32 ? 8 : 4;
}

MmseTile mmse_tile_for(int L, int A, int nSC) {
    load_default_table();
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        for (const tune::Mmse &e : g_tune.mmse)
            if ((e.L < 0 || e.L == L) && (e.A < 0 || e.A == A) && (e.nsc_max < 0 || nSC <= e.nsc_max) &&
                e.tile > 0) {
                MmseTile t;
                t.tile = e.tile;
                t.waves = e.waves;
                t.keep = e.keep;
                return t;
            }
    }
    // one Gram round per wave; for large batches four waves share the algebra of their subcarriers
    MmseTile t;
    t.waves = nSC > device_cus() * 64 ? 4 : 1;
    t.tile = t.waves * 64 / mmse_lanes_for(L, A);
    return t;
}

MmseTile solve_tile_for(int L, int A, int nSC) {
    load_default_table();
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        for (const tune::Solve &e : g_tune.solve)
            if ((e.L < 0 || e.L == L) && (e.A < 0 || e.A == A) && (e.nsc_max < 0 || nSC <= e.nsc_max)) {
                MmseTile t;
                t.tile = e.tile;
                t.waves = e.waves;
                return t;
            }
    }
    if (nSC <= 4096) {  // a single cell: the coefficient kernel's tile (latency: many small workgroups)
        const MmseTile m = mmse_tile_for(L, A, nSC);
        MmseTile t;
        t.tile = m.tile;
        t.waves = m.waves;
        return t;
    }
    MmseTile t;
    t.tile = 64;
    t.waves = 4;
    return t;
}

SlotTile slot_tile() {
    load_default_table();
    std::lock_guard<std::mutex> lk(g_table_mu);
    SlotTile t;
    t.mmse_tile = g_tune.slot.mmse_tile;
    t.mmse_waves = g_tune.slot.mmse_waves;
    t.mmse_keep = g_tune.slot.mmse_keep;
    return t;
}
}  // namespace pe

extern "C" {

const char *pe_version(void) { return "pusch_eq 0.1 (gfx950)"; }

const char *pe_last_error(void) { return g_err.c_str(); }

const char *pe_status_string(pe_status s) {
    switch (s) {
    case PE_OK: return "ok";
    case PE_ERR_INVALID: return "invalid argument";
    case PE_ERR_UNSUPPORTED: return "unsupported";
    case PE_ERR_DEVICE: return "device error";
    case PE_ERR_NOMEM: return "out of memory";
    case PE_ERR_CAPACITY: return "capacity exceeded";
    }
    return "unknown status";
}

pe_status pe_init(int device) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return from_hip(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return from_hip(e, "hipGetDeviceProperties");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PE_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950 only", device,
                    prop.gcnArchName);
    const char *path = getenv("PE_TUNE_TABLE");
    if (path && *path && load_table_file(path) < 0)
        return fail(PE_ERR_INVALID, "cannot read tile table %s", path);
    void *scratch = nullptr;  // set up the scratch address of pe_mmse_apply_demap_noalloc for this device
    e = pe::noalloc_scratch(&scratch);
    if (e != hipSuccess) return from_hip(e, "scratch address");
    return ok();
}

int pe_tune_load(const char *path) {
    if (!path) {
        std::lock_guard<std::mutex> lk(g_table_mu);
        g_tune = pe::tune::Table();
        g_table_set = true;
        return 0;
    }
    return load_table_file(path);
}

static pe_status check_coef_shape(int nSC, int L, int S, int A, int prec) {
    if (nSC < 0) return fail(PE_ERR_INVALID, "nSC = %d < 0", nSC);
    if (L < 1 || L > 8) return fail(PE_ERR_INVALID, "nLayer = %d outside [1, 8]", L);
    if (S < 1) return fail(PE_ERR_INVALID, "nSym = %d < 1", S);
    if (S > 16) return fail(PE_ERR_UNSUPPORTED, "nSym = %d > 16", S);
    if (A < 4 || A % 4 != 0) return fail(PE_ERR_INVALID, "nAnt = %d must be a positive multiple of 4", A);
    if (!valid_prec(prec)) return fail(PE_ERR_INVALID, "unknown precision %d", prec);
    if (prec == PE_PREC_F32 && A > 256) return fail(PE_ERR_UNSUPPORTED, "nAnt = %d > 256", A);
    if (prec != PE_PREC_F32 && A > 128) return fail(PE_ERR_UNSUPPORTED, "nAnt = %d > 128 on MFMA", A);
    return PE_OK;
}

pe_status pe_coef_apply(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC, int nLayer, int nSym,
                        int nAnt, pe_prec prec, void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (nSC == 0) return ok();
    if (!check_ptrs(C, Y_rx, Y_eq)) return fail(PE_ERR_INVALID, "null pointer");
    s = check_align({{C, 16, "C"}, {Y_rx, 16, "Y_rx"}, {Y_eq, 8, "Y_eq"}});
    if (s != PE_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    const float2 *c = reinterpret_cast<const float2 *>(C);
    const float2 *y = reinterpret_cast<const float2 *>(Y_rx);
    float2 *o = reinterpret_cast<float2 *>(Y_eq);
    hipError_t e = (prec == PE_PREC_F32)
                       ? pe::launch_coef_apply_f32(c, y, o, nSC, nLayer, nSym, nAnt, st)
                       : pe::launch_coef_apply_mfma(c, y, o, nullptr, nullptr, nSC, nLayer, nSym, nAnt, 0,
                                                    (int)prec, st);
    return from_hip(e, "coef_apply launch");
}

pe_status pe_coef_apply_v1(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC, int nSym, void *stream) {
    return pe_coef_apply(C, Y_rx, Y_eq, nSC, 8, nSym, 64, PE_PREC_F32, stream);
}

pe_status pe_coef_apply_range(const pe_c64 *C, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC, int scOffset,
                              int nLayer, int nSym, int nAnt, pe_prec prec, void *stream) {
    if (scOffset < 0) return fail(PE_ERR_INVALID, "scOffset = %d < 0", scOffset);
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (nSC == 0) return ok();
    if (!check_ptrs(C, Y_rx, Y_eq)) return fail(PE_ERR_INVALID, "null pointer");
    const size_t off = (size_t)scOffset;
    return pe_coef_apply(C + off * nLayer * nAnt, Y_rx + off * nSym * nAnt, Y_eq + off * nLayer * nSym, nSC,
                         nLayer, nSym, nAnt, prec, stream);
}

pe_status pe_gram(const pe_c64 *M, const pe_c64 *H, pe_c64 *G, int nSC, int nLayer, int nAnt, void *stream) {
    if (nSC < 0) return fail(PE_ERR_INVALID, "nSC = %d < 0", nSC);
    if (nLayer < 1 || nLayer > 8) return fail(PE_ERR_INVALID, "nLayer = %d outside [1, 8]", nLayer);
    if (nAnt < 1 || nAnt > 1024) return fail(PE_ERR_INVALID, "nAnt = %d outside [1, 1024]", nAnt);
    if (nSC == 0) return ok();
    if (!check_ptrs(M, H, G)) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{M, 8, "M"}, {H, 8, "H"}, {G, 8, "G"}});
    if (s != PE_OK) return s;
    hipError_t e = pe::launch_gram_f32(reinterpret_cast<const float2 *>(M), reinterpret_cast<const float2 *>(H),
                                       reinterpret_cast<float2 *>(G), nSC, nLayer, nAnt, (hipStream_t)stream);
    return from_hip(e, "gram launch");
}

pe_status pe_gram_mmse(const pe_c64 *H, const float *sigma2, pe_c64 *C, float *noise_inv, uint8_t *sc_flags,
                       int nSC, int nLayer, int nAnt, unsigned flags, float cond_threshold, void *stream) {
    if (nSC < 0) return fail(PE_ERR_INVALID, "nSC = %d < 0", nSC);
    if (nLayer < 1 || nLayer > 8) return fail(PE_ERR_INVALID, "nLayer = %d outside [1, 8]", nLayer);
    if (nAnt < 1 || nAnt > 1024) return fail(PE_ERR_INVALID, "nAnt = %d outside [1, 1024]", nAnt);
    if (flags & ~PE_MMSE_UNBIASED) return fail(PE_ERR_INVALID, "unknown flags 0x%x", flags);
    if (nSC == 0) return ok();
    if (!check_ptrs(H, sigma2, C) || !noise_inv || !sc_flags) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{H, pe::h_align(nLayer), "H"}, {sigma2, 4, "sigma2"}, {C, 8, "C"},
                               {noise_inv, 4, "noise_inv"}});
    if (s != PE_OK) return s;
    const float thr = cond_threshold > 0.f ? cond_threshold : 1e4f;
    hipError_t e = pe::launch_gram_mmse(reinterpret_cast<const float2 *>(H), sigma2, reinterpret_cast<float2 *>(C),
                                        noise_inv, sc_flags, nSC, nLayer, nAnt, (flags & PE_MMSE_UNBIASED) != 0,
                                        thr, (hipStream_t)stream);
    return from_hip(e, "gram_mmse launch");
}

pe_status pe_gram_mmse_c16(const pe_c64 *H, const float *sigma2, pe_c16 *C16, int *c_exp, float *noise_inv,
                           uint8_t *sc_flags, int nSC, int nLayer, int nAnt, unsigned flags, float cond_threshold,
                           void *stream) {
    if (nSC < 0) return fail(PE_ERR_INVALID, "nSC = %d < 0", nSC);
    if (nLayer < 1 || nLayer > 8) return fail(PE_ERR_INVALID, "nLayer = %d outside [1, 8]", nLayer);
    if (nAnt < 4 || nAnt % 4 != 0) return fail(PE_ERR_INVALID, "nAnt = %d must be a positive multiple of 4", nAnt);
    if (nAnt > 128) return fail(PE_ERR_UNSUPPORTED, "nAnt = %d > 128 for binary16 coefficients", nAnt);
    if (flags & ~PE_MMSE_UNBIASED) return fail(PE_ERR_INVALID, "unknown flags 0x%x", flags);
    if (nSC == 0) return ok();
    if (!check_ptrs(H, sigma2, C16) || !c_exp || !noise_inv || !sc_flags) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{H, pe::h_align(nLayer), "H"}, {sigma2, 4, "sigma2"}, {C16, 4, "C16"}, {c_exp, 4, "c_exp"},
                               {noise_inv, 4, "noise_inv"}});
    if (s != PE_OK) return s;
    const float thr = cond_threshold > 0.f ? cond_threshold : 1e4f;
    hipError_t e = pe::launch_gram_mmse_c16(reinterpret_cast<const float2 *>(H), sigma2,
                                            reinterpret_cast<uint32_t *>(C16), c_exp, noise_inv, sc_flags, nSC, nLayer,
                                            nAnt, (flags & PE_MMSE_UNBIASED) != 0, thr, (hipStream_t)stream);
    return from_hip(e, "gram_mmse_c16 launch");
}

pe_status pe_llr_demap_f16(const pe_c16 *softEst, const float *noiseInv, uint16_t *llrOut, int N, int nPamBits,
                           void *stream) {
    if (N < 0) return fail(PE_ERR_INVALID, "N = %d < 0", N);
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (N == 0) return ok();
    if (!check_ptrs(softEst, noiseInv, llrOut)) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{softEst, 4, "softEst"}, {noiseInv, 4, "noiseInv"},
                               {llrOut, pe::llr_align(nPamBits, true), "llrOut"}});
    if (s != PE_OK) return s;
    hipError_t e = pe::launch_llr_f16(reinterpret_cast<const uint32_t *>(softEst), noiseInv, llrOut, N, nPamBits,
                                      (hipStream_t)stream);
    return from_hip(e, "llr_demap_f16 launch");
}

pe_status pe_llr_demap(const pe_c64 *Y_eq, const float *noise_inv, float *llr, int nSC, int nLayer, int nSym,
                       int nPamBits, void *stream) {
    if (nSC < 0) return fail(PE_ERR_INVALID, "nSC = %d < 0", nSC);
    if (nLayer < 1 || nLayer > 8) return fail(PE_ERR_INVALID, "nLayer = %d outside [1, 8]", nLayer);
    if (nSym < 1 || nSym > 16) return fail(PE_ERR_INVALID, "nSym = %d outside [1, 16]", nSym);
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (nSC == 0) return ok();
    if (!check_ptrs(Y_eq, noise_inv, llr)) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{Y_eq, 8, "Y_eq"}, {noise_inv, 4, "noise_inv"}, {llr, pe::llr_align(nPamBits, false), "llr"}});
    if (s != PE_OK) return s;
    hipError_t e = pe::launch_llr_f32(reinterpret_cast<const float2 *>(Y_eq), noise_inv, llr, nSC, nLayer, nSym,
                                      nPamBits, (hipStream_t)stream);
    return from_hip(e, "llr_demap launch");
}

pe_status pe_coef_apply_demap(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, float *llr, int nSC,
                              int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec, void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (nSC == 0) return ok();
    if (!check_ptrs(C, Y_rx, noise_inv) || !llr) return fail(PE_ERR_INVALID, "null pointer");
    s = check_align({{C, 16, "C"}, {Y_rx, 16, "Y_rx"}, {noise_inv, 4, "noise_inv"},
                     {llr, pe::llr_align(nPamBits, false), "llr"}});
    if (s != PE_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    const float2 *c = reinterpret_cast<const float2 *>(C);
    const float2 *y = reinterpret_cast<const float2 *>(Y_rx);
    hipError_t e = (prec == PE_PREC_F32)
                       ? pe::launch_coef_apply_demap_f32(c, y, noise_inv, llr, nSC, nLayer, nSym, nAnt, nPamBits, st)
                       : pe::launch_coef_apply_mfma(c, y, nullptr, noise_inv, llr, nSC, nLayer, nSym, nAnt,
                                                    nPamBits, (int)prec, st);
    return from_hip(e, "coef_apply_demap launch");
}

pe_status pe_coef_apply_demap_f16(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, uint16_t *llr,
                                  int nSC, int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec,
                                  void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (nSC == 0) return ok();
    if (!check_ptrs(C, Y_rx, noise_inv) || !llr) return fail(PE_ERR_INVALID, "null pointer");
    s = check_align({{C, 16, "C"}, {Y_rx, 16, "Y_rx"}, {noise_inv, 4, "noise_inv"},
                     {llr, pe::llr_align(nPamBits, true), "llr"}});
    if (s != PE_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    const float2 *c = reinterpret_cast<const float2 *>(C);
    const float2 *y = reinterpret_cast<const float2 *>(Y_rx);
    float *out = reinterpret_cast<float *>(llr);  // binary16 storage; the kernels store through uint16_t
    hipError_t e = (prec == PE_PREC_F32)
                       ? pe::launch_coef_apply_demap_f32(c, y, noise_inv, out, nSC, nLayer, nSym, nAnt, nPamBits,
                                                         st, 1)
                       : pe::launch_coef_apply_mfma(c, y, nullptr, noise_inv, out, nSC, nLayer, nSym, nAnt,
                                                    nPamBits, (int)prec, st, 1);
    return from_hip(e, "coef_apply_demap_f16 launch");
}

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

pe_status pe_coef_to_c16(const pe_c64 *C, pe_c16 *C16, int *c_exp, int nSC, int nLayer, int nAnt, void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, 1, nAnt, (int)PE_PREC_F16);
    if (s != PE_OK) return s;
    if (nSC == 0) return ok();
    if (!check_ptrs(C, C16, c_exp)) return fail(PE_ERR_INVALID, "null pointer");
    if (!aligned16(C) || !aligned16(C16)) return fail(PE_ERR_INVALID, "C and C16 must be 16-byte aligned");
    return from_hip(pe::launch_coef_to_c16(reinterpret_cast<const float2 *>(C), reinterpret_cast<uint32_t *>(C16),
                                           c_exp, nSC, nLayer, nAnt, (hipStream_t)stream),
                    "coef_to_c16 launch");
}

static pe_status coef_apply_c16(const pe_c16 *C16, const int *c_exp, const pe_c64 *Y_rx, pe_c64 *Y_eq,
                                const float *noise_inv, void *llr, int llr_f16, int nSC, int nLayer, int nSym,
                                int nAnt, int nPamBits, void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)PE_PREC_F16);
    if (s != PE_OK) return s;
    if (llr && (nPamBits < 1 || nPamBits > 4)) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (nSC == 0) return ok();
    if (!check_ptrs(C16, c_exp, Y_rx) || (llr ? !noise_inv : !Y_eq)) return fail(PE_ERR_INVALID, "null pointer");
    if (!aligned16(C16) || !aligned16(Y_rx)) return fail(PE_ERR_INVALID, "C16 and Y_rx must be 16-byte aligned");
    s = check_align({{c_exp, 4, "c_exp"}, {Y_eq, 8, "Y_eq"}, {noise_inv, 4, "noise_inv"},
                     {llr, pe::llr_align(nPamBits, llr_f16 != 0), "llr"}});
    if (s != PE_OK) return s;
    hipError_t e = pe::launch_coef_apply_c16(reinterpret_cast<const uint32_t *>(C16), c_exp,
                                             reinterpret_cast<const float2 *>(Y_rx), reinterpret_cast<float2 *>(Y_eq),
                                             noise_inv, llr, nSC, nLayer, nSym, nAnt, llr ? nPamBits : 0, llr_f16,
                                             (hipStream_t)stream);
    return from_hip(e, "coef_apply_c16 launch");
}

pe_status pe_coef_apply_c16(const pe_c16 *C16, const int *c_exp, const pe_c64 *Y_rx, pe_c64 *Y_eq, int nSC,
                            int nLayer, int nSym, int nAnt, void *stream) {
    return coef_apply_c16(C16, c_exp, Y_rx, Y_eq, nullptr, nullptr, 0, nSC, nLayer, nSym, nAnt, 0, stream);
}

pe_status pe_coef_apply_demap_c16(const pe_c16 *C16, const int *c_exp, const pe_c64 *Y_rx, const float *noise_inv,
                                  void *llr, int nSC, int nLayer, int nSym, int nAnt, int nPamBits,
                                  unsigned flags, void *stream) {
    if (flags & ~PE_LLR_F16) return fail(PE_ERR_INVALID, "unknown flag bits 0x%x", flags);
    if (!llr && nSC > 0) return fail(PE_ERR_INVALID, "null pointer");
    return coef_apply_c16(C16, c_exp, Y_rx, nullptr, noise_inv, llr, (flags & PE_LLR_F16) ? 1 : 0, nSC, nLayer,
                          nSym, nAnt, nPamBits, stream);
}

// Arguments of the one-pass chain other than its workspace C (shared by pe_mmse_apply_demap and the variant
// without allocation).
static pe_status check_mmse_apply(const pe_c64 *H, const float *sigma2, const pe_c64 *Y_rx, const float *noise_inv,
                                  const uint8_t *sc_flags, const void *llr, int nSC, int nLayer, int nSym, int nAnt,
                                  int nPamBits, pe_prec prec, unsigned flags, unsigned allowed_flags) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (prec == PE_PREC_F32)
        return fail(PE_ERR_UNSUPPORTED, "precision %d: the fused MMSE + apply runs the MFMA product", (int)prec);
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (flags & ~allowed_flags) return fail(PE_ERR_INVALID, "unknown flag bits 0x%x", flags & ~allowed_flags);
    if (nSC == 0) return PE_OK;
    if (!check_ptrs(H, sigma2, Y_rx) || !noise_inv || !sc_flags || !llr) return fail(PE_ERR_INVALID, "null pointer");
    if (!aligned16(H) || !aligned16(Y_rx) || !aligned16(llr))
        return fail(PE_ERR_INVALID, "H, Y_rx and llr must be 16-byte aligned");
    return check_align({{sigma2, 4, "sigma2"}, {noise_inv, 4, "noise_inv"}});
}

pe_status pe_mmse_apply_demap(const pe_c64 *H, const float *sigma2, const pe_c64 *Y_rx, pe_c64 *C, float *noise_inv,
                              uint8_t *sc_flags, void *llr, int nSC, int nLayer, int nSym, int nAnt, int nPamBits,
                              pe_prec prec, unsigned flags, float cond_threshold, void *stream) {
    pe_status s = check_mmse_apply(H, sigma2, Y_rx, noise_inv, sc_flags, llr, nSC, nLayer, nSym, nAnt, nPamBits, prec,
                                   flags, PE_MMSE_UNBIASED | PE_EQ_WRITE_C | PE_EQ_LLR_F16);
    if (s != PE_OK) return s;
    if (nSC == 0) return ok();
    if (!C) return fail(PE_ERR_INVALID, "null pointer");
    if (reinterpret_cast<uintptr_t>(C) & 7u) return fail(PE_ERR_INVALID, "C must be 8-byte aligned");
    const float thr = cond_threshold > 0.f ? cond_threshold : 1e4f;
    hipError_t e = pe::launch_mmse_apply_demap(
        reinterpret_cast<const float2 *>(H), sigma2, reinterpret_cast<const float2 *>(Y_rx), reinterpret_cast<float2 *>(C),
        noise_inv, sc_flags, llr, nSC, nLayer, nSym, nAnt, nPamBits, (int)prec, (flags & PE_MMSE_UNBIASED) != 0, thr,
        (flags & PE_EQ_LLR_F16) ? 1 : 0, (flags & PE_EQ_WRITE_C) ? 1 : 0, (hipStream_t)stream);
    return from_hip(e, "mmse_apply_demap launch");
}

pe_status pe_mmse_apply_demap_noalloc(const pe_c64 *H, const float *sigma2, const pe_c64 *Y_rx, float *noise_inv,
                                      uint8_t *sc_flags, void *llr, int nSC, int nLayer, int nSym, int nAnt,
                                      int nPamBits, pe_prec prec, unsigned flags, float cond_threshold, void *stream) {
    pe_status s = check_mmse_apply(H, sigma2, Y_rx, noise_inv, sc_flags, llr, nSC, nLayer, nSym, nAnt, nPamBits, prec,
                                   flags, PE_MMSE_UNBIASED | PE_EQ_LLR_F16);
    if (s != PE_OK) return s;
    const size_t ws = (size_t)nSC * nLayer * nAnt * sizeof(pe_c64);
    if (ws > PE_NOALLOC_SCRATCH_BYTES)
        return fail(PE_ERR_CAPACITY, "workspace of %zu bytes (nSC * nLayer * nAnt * 8) exceeds the %u-byte scratch", ws,
                    (unsigned)PE_NOALLOC_SCRATCH_BYTES);
    if (nSC == 0) return ok();
    void *scratch = nullptr;  // the array's address on the current device (no allocation)
    hipError_t e = pe::noalloc_scratch(&scratch);
    if (e != hipSuccess) return from_hip(e, "scratch address");
    const float thr = cond_threshold > 0.f ? cond_threshold : 1e4f;
    e = pe::launch_mmse_apply_demap(reinterpret_cast<const float2 *>(H), sigma2, reinterpret_cast<const float2 *>(Y_rx),
                                    reinterpret_cast<float2 *>(scratch), noise_inv, sc_flags, llr, nSC, nLayer, nSym,
                                    nAnt, nPamBits, (int)prec, (flags & PE_MMSE_UNBIASED) != 0, thr,
                                    (flags & PE_EQ_LLR_F16) ? 1 : 0, 0, (hipStream_t)stream);
    return from_hip(e, "mmse_apply_demap_noalloc launch");
}

pe_status pe_gold_sequence(uint32_t c_init, uint32_t *c_words, int n_bits, void *stream) {
    if (n_bits < 1 || n_bits >= (1 << 30)) return fail(PE_ERR_INVALID, "n_bits = %d outside [1, 2^30)", n_bits);
    if (!c_words) return fail(PE_ERR_INVALID, "null pointer");
    pe_status s = check_align({{c_words, 4, "c_words"}});
    if (s != PE_OK) return s;
    return from_hip(pe::launch_gold_sequence(c_init, c_words, n_bits, (hipStream_t)stream), "gold_sequence launch");
}

pe_status pe_coef_apply_demap_cw(const pe_c64 *C, const pe_c64 *Y_rx, const float *noise_inv, const pe_cw_layout *cw,
                                 int nSC, int nLayer, int nSym, int nAnt, int nPamBits, pe_prec prec, unsigned flags,
                                 void *stream) {
    pe_status s = check_coef_shape(nSC, nLayer, nSym, nAnt, (int)prec);
    if (s != PE_OK) return s;
    if (prec == PE_PREC_F32) return fail(PE_ERR_UNSUPPORTED, "the codeword layout runs with the MFMA precisions");
    if (nPamBits < 1 || nPamBits > 4) return fail(PE_ERR_INVALID, "nPamBits = %d outside [1, 4]", nPamBits);
    if (flags & ~PE_LLR_F16) return fail(PE_ERR_INVALID, "unknown flag bits 0x%x", flags);
    if (!cw) return fail(PE_ERR_INVALID, "null codeword layout");
    const bool f16 = (flags & PE_LLR_F16) != 0;
    const int qm = 2 * nPamBits;
    int nu[8] = {0}, rank[8] = {0}, ncw = 0;
    for (int l = 0; l < nLayer; ++l) {
        const int k = cw->layer_cw[l];
        if (k < 0 || k >= nLayer) return fail(PE_ERR_INVALID, "layer_cw[%d] = %d outside [0, nLayer)", l, k);
        rank[l] = nu[k]++;
        ncw = k + 1 > ncw ? k + 1 : ncw;
    }
    for (int k = 0; k < ncw; ++k)
        if (nu[k] == 0) return fail(PE_ERR_INVALID, "codeword %d has no layer (codewords must be numbered 0 .. n-1)", k);
    if (nSC == 0) return ok();
    if (!check_ptrs(C, Y_rx, noise_inv)) return fail(PE_ERR_INVALID, "null pointer");
    s = check_align({{C, 16, "C"}, {Y_rx, 16, "Y_rx"}, {noise_inv, 4, "noise_inv"}});
    if (s != PE_OK) return s;
    const long long per_layer = (long long)nSC * nSym;  // symbols per layer
    pe::CwLayer layers[8] = {};
    for (int k = 0; k < ncw; ++k) {
        const long long G = per_layer * nu[k] * qm;
        if (G >= (1ll << 31)) return fail(PE_ERR_UNSUPPORTED, "codeword %d: %lld bits >= 2^31", k, G);
        if (!cw->llr[k]) return fail(PE_ERR_INVALID, "codeword %d: null llr", k);
        if (!pe::is_aligned(cw->llr[k], pe::llr_align(nPamBits, f16)))
            return fail(PE_ERR_INVALID, "codeword %d: llr must be %zu-byte aligned", k, pe::llr_align(nPamBits, f16));
        if (cw->c_seq[k] && !pe::is_aligned(cw->c_seq[k], 4))
            return fail(PE_ERR_INVALID, "codeword %d: c_seq must be 4-byte aligned", k);
        const int C_blocks = cw->nCodeBlocks[k];
        if (C_blocks < 0 || C_blocks > per_layer)
            return fail(PE_ERR_INVALID, "codeword %d: nCodeBlocks = %d outside [0, nSC nSym = %lld]", k, C_blocks,
                        per_layer);
    }
    for (int l = 0; l < nLayer; ++l) {
        const int k = cw->layer_cw[l];
        pe::CwLayer &r = layers[l];
        r.out = (unsigned long long)(uintptr_t)cw->llr[k];
        r.cseq = (unsigned long long)(uintptr_t)cw->c_seq[k];
        r.nu_v = (uint32_t)nu[k] | ((uint32_t)rank[l] << 8);
        const int C_blocks = cw->nCodeBlocks[k];
        if (C_blocks > 0) {  // 38.212 5.4.2.1: r < C - mod(N, C) get floor(N / C) symbols per layer, the rest one more
            const long long N = per_layer;
            r.e_lo = (uint32_t)((long long)nu[k] * qm * (N / C_blocks));
            r.r0 = (uint32_t)(C_blocks - N % C_blocks);
        }
    }
    hipError_t e = pe::launch_coef_apply_demap_cw(reinterpret_cast<const float2 *>(C),
                                                  reinterpret_cast<const float2 *>(Y_rx), noise_inv, layers, nSC,
                                                  nLayer, nSym, nAnt, nPamBits, (int)prec, f16 ? 1 : 0,
                                                  (hipStream_t)stream);
    return from_hip(e, "coef_apply_demap_cw launch");
}

pe_status pe_bfp_decompress(const uint8_t *in, pe_c64 *Y_rx, int nPrb, int nSym, int nAnt, int iqWidth,
                            size_t row_stride, float scale, void *stream) {
    if (nPrb < 0) return fail(PE_ERR_INVALID, "nPrb = %d < 0", nPrb);
    if (nSym < 1 || nSym > 65535) return fail(PE_ERR_INVALID, "nSym = %d outside [1, 65535]", nSym);
    if (nAnt < 1 || nAnt > 4096) return fail(PE_ERR_INVALID, "nAnt = %d outside [1, 4096]", nAnt);
    if (iqWidth < 1 || iqWidth > 16) return fail(PE_ERR_INVALID, "iqWidth = %d outside [1, 16]", iqWidth);
    const size_t bpb = 1 + 3 * (size_t)iqWidth;
    if (row_stride < (size_t)nPrb * bpb)
        return fail(PE_ERR_INVALID, "row_stride = %zu < nPrb * (1 + 3 iqWidth) = %zu", row_stride, (size_t)nPrb * bpb);
    if (nPrb == 0) return ok();
    if (!in || !Y_rx) return fail(PE_ERR_INVALID, "null pointer");
    return from_hip(pe::launch_bfp_decompress(in, reinterpret_cast<float2 *>(Y_rx), nPrb, nSym, nAnt, iqWidth,
                                              row_stride, scale, (hipStream_t)stream),
                    "bfp_decompress launch");
}

pe_status pe_srs_filter(pe_c16 *H_est, const pe_c16 *SRS, const pe_c16 *W, const pe_c16 *FOCC, int nAnt,
                        int nBlocks, int K, int nPorts, int foccLen, unsigned flags, void *stream) {
    if (nBlocks < 0) return fail(PE_ERR_INVALID, "nBlocks = %d < 0", nBlocks);
    if (nAnt < 1 || nAnt > 4096) return fail(PE_ERR_INVALID, "nAnt = %d outside [1, 4096]", nAnt);
    if (K < 1 || K > 48) return fail(PE_ERR_INVALID, "K = %d outside [1, 48]", K);
    if (nPorts < 1 || nPorts > 8) return fail(PE_ERR_INVALID, "nPorts = %d outside [1, 8]", nPorts);
    if (foccLen < 1 || foccLen > 48) return fail(PE_ERR_INVALID, "foccLen = %d outside [1, 48]", foccLen);
    if (flags & ~(PE_SRS_MFMA | PE_SRS_CONJ_W)) return fail(PE_ERR_INVALID, "unknown flag bits 0x%x", flags);
    const bool mfma = (flags & PE_SRS_MFMA) != 0;
    if (mfma && K % 4 != 0) return fail(PE_ERR_UNSUPPORTED, "K = %d: PE_SRS_MFMA needs K %% 4 == 0", K);
    if (mfma && (reinterpret_cast<uintptr_t>(SRS) & 15u))
        return fail(PE_ERR_INVALID, "SRS must be 16-byte aligned with PE_SRS_MFMA");
    if (nBlocks == 0) return ok();
    if (!check_ptrs(H_est, SRS, W) || !FOCC) return fail(PE_ERR_INVALID, "null pointer");
    hipError_t e = pe::launch_srs_filter(reinterpret_cast<uint32_t *>(H_est), reinterpret_cast<const uint32_t *>(SRS),
                                         reinterpret_cast<const uint32_t *>(W),
                                         reinterpret_cast<const uint32_t *>(FOCC), nAnt, nBlocks, K, nPorts, foccLen,
                                         (flags & PE_SRS_CONJ_W) != 0, mfma, (hipStream_t)stream);
    return from_hip(e, "srs_filter launch");
}

pe_status pe_freq_interp(const pe_c64 *W, const pe_c64 *X, pe_c64 *H_f, int nBatch, int nOut, int nIn, int nCol,
                         pe_prec prec, void *stream) {
    if (nBatch < 0) return fail(PE_ERR_INVALID, "nBatch = %d < 0", nBatch);
    if (nOut < 1 || nOut > 256) return fail(PE_ERR_INVALID, "nOut = %d outside [1, 256]", nOut);
    if (nIn < 1 || nIn > 128) return fail(PE_ERR_INVALID, "nIn = %d outside [1, 128]", nIn);
    if (nCol < 1 || nCol > 16) return fail(PE_ERR_INVALID, "nCol = %d outside [1, 16]", nCol);
    if (!valid_prec(prec)) return fail(PE_ERR_INVALID, "unknown precision %d", prec);
    if (prec != PE_PREC_F32 && prec != PE_PREC_BF16 && prec != PE_PREC_BF16_SPLIT)
        return fail(PE_ERR_UNSUPPORTED, "precision %d: frequency interpolation runs F32, BF16 or BF16_SPLIT", prec);
    if (prec == PE_PREC_F32 && pe::freq_interp_f32_lds_bytes(nOut, nIn, nCol) > 64 * 1024)
        return fail(PE_ERR_UNSUPPORTED, "PE_PREC_F32 keeps W and one X block in 64 KB of LDS: nOut %d x nIn %d too large",
                    nOut, nIn);
    if (nBatch == 0) return ok();
    if (!check_ptrs(W, X, H_f)) return fail(PE_ERR_INVALID, "null pointer");
    if ((reinterpret_cast<uintptr_t>(W) | reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(H_f)) & 7u)
        return fail(PE_ERR_INVALID, "W, X and H_f must be 8-byte aligned");
    hipError_t e = pe::launch_freq_interp(reinterpret_cast<const float2 *>(W), reinterpret_cast<const float2 *>(X),
                                          reinterpret_cast<float2 *>(H_f), nBatch, nOut, nIn, nCol, (int)prec,
                                          (hipStream_t)stream);
    return from_hip(e, "freq_interp launch");
}

pe_status pe_interp_mmse_apply_demap(const pe_c64 *W, const pe_c64 *X, const float *sigma2, const pe_c64 *Y_rx,
                                     pe_c64 *H, pe_c64 *C, float *noise_inv, uint8_t *sc_flags, void *llr, int nBlk,
                                     int nIn, int nSym, int nAnt, int nPamBits, pe_prec interp_prec, pe_prec prec,
                                     unsigned flags, float cond_threshold, void *stream) {
    if (nBlk < 0) return fail(PE_ERR_INVALID, "nBlk = %d < 0", nBlk);
    if (nBlk > (1 << 24)) return fail(PE_ERR_UNSUPPORTED, "nBlk = %d > 2^24", nBlk);
    if (nIn < 1 || nIn > 128) return fail(PE_ERR_INVALID, "nIn = %d outside [1, 128]", nIn);
    if (interp_prec != PE_PREC_BF16 && interp_prec != PE_PREC_BF16_SPLIT)
        return fail(PE_ERR_UNSUPPORTED, "interpolation precision %d: BF16 or BF16_SPLIT", (int)interp_prec);
    if (nAnt < 32 || nAnt > 64 || nAnt % 4 != 0)
        return fail(PE_ERR_UNSUPPORTED, "nAnt = %d: the fused chain takes 32 to 64 antennas, a multiple of 4", nAnt);
    const int nSC = 64 * nBlk;
    pe_status s = check_mmse_apply(H, sigma2, Y_rx, noise_inv, sc_flags, llr, nSC, 4, nSym, nAnt, nPamBits, prec,
                                   flags, PE_MMSE_UNBIASED | PE_EQ_WRITE_C | PE_EQ_LLR_F16);
    if (s != PE_OK) return s;
    if (nBlk == 0) return ok();
    if (!check_ptrs(W, X, C)) return fail(PE_ERR_INVALID, "null pointer");
    if ((reinterpret_cast<uintptr_t>(W) | reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(C)) & 7u)
        return fail(PE_ERR_INVALID, "W, X and C must be 8-byte aligned");
    const float thr = cond_threshold > 0.f ? cond_threshold : 1e4f;
    hipError_t e = pe::launch_interp_mmse_apply_demap(
        reinterpret_cast<const float2 *>(W), reinterpret_cast<const float2 *>(X), sigma2,
        reinterpret_cast<const float2 *>(Y_rx), reinterpret_cast<float2 *>(H), reinterpret_cast<float2 *>(C), noise_inv,
        sc_flags, llr, nBlk, nIn, nSym, nAnt, nPamBits, (int)interp_prec, (int)prec, (flags & PE_MMSE_UNBIASED) != 0,
        thr, (flags & PE_EQ_LLR_F16) ? 1 : 0, (flags & PE_EQ_WRITE_C) ? 1 : 0, (hipStream_t)stream);
    return from_hip(e, "interp_mmse_apply_demap launch");
}

pe_status pe_hbm_probe(const void *src, void *dst, size_t bytes, int read_only, int grid, void *stream) {
    if (bytes % 16 != 0) return fail(PE_ERR_INVALID, "bytes = %zu is not a multiple of 16", bytes);
    if (grid < 1 || grid > (1 << 20)) return fail(PE_ERR_INVALID, "grid = %d outside [1, 2^20]", grid);
    if (bytes == 0) return ok();
    if (!src || !dst) return fail(PE_ERR_INVALID, "null pointer");
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15u)
        return fail(PE_ERR_INVALID, "src and dst must be 16-byte aligned");
    return from_hip(pe::launch_hbm_probe(src, dst, bytes, read_only, grid, (hipStream_t)stream), "hbm_probe launch");
}

}  // extern "C"
