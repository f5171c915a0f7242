"""CPU oracle for the PUSCH equalization path -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker (or the timed CPU
baseline). The product (libpusch_eq.so and its binding) never imports this module.

Two parts:
  * ctypes wrappers of libpe_oracle.so (pe_oracle.c): the C restatement of the reference arithmetic and
    the glibc rand() input streams of the reference benchmarks;
  * numpy metrics and synthetic-input generators (restated from the reference where it has them).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PE_ORACLE_LIB names another build of pe_oracle.c (the AddressSanitizer build of the CPU checks)
LIB_PATH = os.environ.get("PE_ORACLE_LIB") or os.path.join(_HERE, "libpe_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: run `make -C oracle` (or __graft_entry__.build())" % LIB_PATH)
        h = ctypes.CDLL(LIB_PATH)
        P, I, U, F, D, Z = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_double,
                            ctypes.c_size_t)
        sig = {
            "po_glibc_uniform": ([U, F, I, P, Z], None),
            "po_llr_bench_inputs": ([U, I, P, P], None),
            "po_f32_to_f16_array": ([P, P, Z], None),
            "po_f16_to_f32_array": ([P, P, Z], None),
            "po_coef_apply_v1_f32": ([P, P, P, I, I, I, I], None),
            "po_coef_apply_v1_f32_nofma": ([P, P, P, I, I, I, I], None),
            "po_coef_apply_f64": ([P, P, P, I, I, I, I], None),
            "po_coef_apply_f64d": ([P, P, P, I, I, I, I], None),
            "po_coef_apply_v1_f32_omp": ([P, P, P, I, I, I, I, I], I),
            "po_gram_f32": ([P, P, P, I, I, I], None),
            "po_gram_f64": ([P, P, P, I, I, I], None),
            "po_mmse_f64": ([P, P, P, P, P, P, I, I, I, I], I),
            "po_llr_ref_f64": ([F, F, F, I, P], None),
            "po_llr_f32": ([F, F, F, I, P], None),
            "po_llr_demap_f16": ([P, P, P, I, I], None),
            "po_llr_demap_ref_f64": ([P, P, P, I, I], None),
            "po_llr_demap_f32": ([P, P, P, I, I, I, I], None),
            "po_llr_demap_f64": ([P, P, P, I, I, I, I], None),
            "po_llr_bruteforce_f64": ([D, D, I, P], None),
            "po_llr_lut": ([I], D),
            "po_coef_apply_demap_f32_omp": ([P, P, P, P, I, I, I, I, I, I], I),
            "po_omp_max_threads": ([], I),
            "po_half_ops": ([P, P, P, P, P, P, P, Z], None),
            "po_srs_filter_v1": ([P, P, P, P, I, I, I, I, I, I], None),
            "po_srs_filter_f64": ([P, P, P, P, I, I, I, I, I, I, I], None),
            "po_srs_bench_inputs": ([U, I, I, I, I, I, P, P, P], None),
            "po_bfp_decompress": ([P, P, I, I, I, I, Z, F], None),
            "po_bfp_compress": ([P, P, I, I, I, I, Z], None),
            "po_freq_interp_f32": ([P, P, P, I, I, I, I], None),
            "po_freq_interp_f32_omp": ([P, P, P, I, I, I, I, I], I),
            "po_freq_interp_f64": ([P, P, P, I, I, I, I], None),
            "po_gold_sequence": ([U, I, P], None),
            "po_cw_block_sizes": ([I, I, I, I, P], I),
            "po_cw_layout": ([P, I, I, I, I, P, I, P, P, P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return a


# ---------------------------------------------------------------------------------------------------------
# reference input streams (glibc rand)
# ---------------------------------------------------------------------------------------------------------
def glibc_uniform(seed, n_floats, scale=None):
    """srand(seed); ((float)rand()/RAND_MAX - 0.5f) [* scale] for n_floats values
    (aerial_coef_wmma.cu:383-387 without scale, coef_rigor.cu:101-104 with scale 1.4)."""
    out = np.empty(n_floats, dtype=np.float32)
    lib().po_glibc_uniform(seed, float(scale or 1.0), 1 if scale is not None else 0, _p(out), n_floats)
    return out


def reference_coef_inputs(seed, nSC, L, S, A, scale=None):
    """C then Y from one glibc stream, re before im per element: returns complex64 C [nSC,L,A], Y [nSC,S,A]."""
    nc = nSC * L * A
    ny = nSC * S * A
    v = glibc_uniform(seed, 2 * (nc + ny), scale)
    C = v[: 2 * nc].view(np.complex64).reshape(nSC, L, A)
    Y = v[2 * nc:].view(np.complex64).reshape(nSC, S, A)
    return C.copy(), Y.copy()


def llr_bench_inputs(seed, N):
    """aerial_llr_bench_v2.cu:289-304: binary16 soft symbols [N,2] (uint16 bits) and noiseInv [N]."""
    soft = np.empty((N, 2), dtype=np.uint16)
    ni = np.empty(N, dtype=np.float32)
    lib().po_llr_bench_inputs(seed, N, _p(soft), _p(ni))
    return soft, ni


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().po_f32_to_f16_array(_p(x), _p(out), x.size)
    return out


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().po_f16_to_f32_array(_p(h), _p(out), h.size)
    return out


# ---------------------------------------------------------------------------------------------------------
# arithmetic
# ---------------------------------------------------------------------------------------------------------
def coef_apply_v1_f32(C, Y, contracted=True):
    C, Y = _c64(C), _c64(Y)
    nSC, L, A = C.shape
    S = Y.shape[1]
    O = np.empty((nSC, L, S), dtype=np.complex64)
    fn = lib().po_coef_apply_v1_f32 if contracted else lib().po_coef_apply_v1_f32_nofma
    fn(_p(C), _p(Y), _p(O), nSC, L, S, A)
    return O


def coef_apply_f64(C, Y):
    """FP64 product; C may be complex64 or complex128."""
    Y = _c64(Y)
    nSC, L, A = C.shape
    S = Y.shape[1]
    O = np.empty((nSC, L, S), dtype=np.complex128)
    if C.dtype == np.complex128:
        C = np.ascontiguousarray(C)
        lib().po_coef_apply_f64d(_p(C), _p(Y), _p(O), nSC, L, S, A)
    else:
        C = _c64(C)
        lib().po_coef_apply_f64(_p(C), _p(Y), _p(O), nSC, L, S, A)
    return O


def coef_apply_v1_f32_omp(C, Y, nthreads=0, out=None):
    C, Y = _c64(C), _c64(Y)
    nSC, L, A = C.shape
    S = Y.shape[1]
    O = out if out is not None else np.empty((nSC, L, S), dtype=np.complex64)
    used = lib().po_coef_apply_v1_f32_omp(_p(C), _p(Y), _p(O), nSC, L, S, A, nthreads)
    return O, used


def coef_apply_demap_f32_omp(C, Y, noise_inv, q, nthreads=0, out=None):
    C, Y = _c64(C), _c64(Y)
    ni = np.ascontiguousarray(noise_inv, dtype=np.float32)
    nSC, L, A = C.shape
    S = Y.shape[1]
    O = out if out is not None else np.empty((nSC, L, S, 2 * q), dtype=np.float32)
    used = lib().po_coef_apply_demap_f32_omp(_p(C), _p(Y), _p(ni), _p(O), nSC, L, S, A, q, nthreads)
    return O, used


def gram_f32(M, H):
    M, H = _c64(M), _c64(H)
    nSC, L, A = M.shape
    G = np.empty((nSC, L, L), dtype=np.complex64)
    lib().po_gram_f32(_p(M), _p(H), _p(G), nSC, L, A)
    return G


def gram_f64(M, H):
    M, H = _c64(M), _c64(H)
    nSC, L, A = M.shape
    G = np.empty((nSC, L, L), dtype=np.complex128)
    lib().po_gram_f64(_p(M), _p(H), _p(G), nSC, L, A)
    return G


def mmse_f64(H, sigma2, unbiased=True):
    """Returns C [nSC,L,A] complex128, noise_inv [nSC,L], mu [nSC,L], cond [nSC], n_singular."""
    H = _c64(H)
    s2 = np.ascontiguousarray(sigma2, dtype=np.float32)
    nSC, A, L = H.shape
    C = np.empty((nSC, L, A), dtype=np.complex128)
    ni = np.empty((nSC, L), dtype=np.float64)
    mu = np.empty((nSC, L), dtype=np.float64)
    cond = np.empty(nSC, dtype=np.float64)
    bad = lib().po_mmse_f64(_p(H), _p(s2), _p(C), _p(ni), _p(mu), _p(cond), nSC, L, A, 1 if unbiased else 0)
    return C, ni, mu, cond, bad


def llr_demap_f16(soft_bits, noise_inv, q):
    """llr_v1_baseline end to end: binary16 bits in, binary16 bits out [N, 2q]."""
    soft_bits = np.ascontiguousarray(soft_bits, dtype=np.uint16)
    ni = np.ascontiguousarray(noise_inv, dtype=np.float32)
    N = soft_bits.shape[0]
    out = np.empty((N, 2 * q), dtype=np.uint16)
    lib().po_llr_demap_f16(_p(soft_bits), _p(ni), _p(out), N, q)
    return out


def llr_demap_ref_f64(soft_bits, noise_inv, q):
    soft_bits = np.ascontiguousarray(soft_bits, dtype=np.uint16)
    ni = np.ascontiguousarray(noise_inv, dtype=np.float32)
    N = soft_bits.shape[0]
    out = np.empty((N, 2 * q), dtype=np.float32)
    lib().po_llr_demap_ref_f64(_p(soft_bits), _p(ni), _p(out), N, q)
    return out


def llr_demap_f32(Yeq, noise_inv, q):
    Yeq = _c64(Yeq)
    ni = np.ascontiguousarray(noise_inv, dtype=np.float32)
    nSC, L, S = Yeq.shape
    out = np.empty((nSC, L, S, 2 * q), dtype=np.float32)
    lib().po_llr_demap_f32(_p(Yeq), _p(ni), _p(out), nSC, L, S, q)
    return out


def llr_demap_f64(Yeq, noise_inv, q):
    Yeq = np.ascontiguousarray(Yeq, dtype=np.complex128)
    ni = np.ascontiguousarray(noise_inv, dtype=np.float64)
    nSC, L, S = Yeq.shape
    out = np.empty((nSC, L, S, 2 * q), dtype=np.float64)
    lib().po_llr_demap_f64(_p(Yeq), _p(ni), _p(out), nSC, L, S, q)
    return out


def llr_bruteforce(x, noise, q):
    out = np.empty(q, dtype=np.float64)
    lib().po_llr_bruteforce_f64(float(x), float(noise), q, _p(out))
    return out


def llr_lut():
    return np.array([lib().po_llr_lut(i) for i in range(10)])


# ---------------------------------------------------------------------------------------------------------
# metrics (aerial_coef_wmma.cu:340-361, coef_rigor.cu:90-91,132-138)
# ---------------------------------------------------------------------------------------------------------
def cosine(ref, test):
    """Cosine over R^{2N} of interleaved re/im (check(), aerial_coef_wmma.cu:340-354)."""
    a = np.asarray(ref).astype(np.complex128).ravel()
    b = np.asarray(test).astype(np.complex128).ravel()
    dot = np.sum(a.real * b.real + a.imag * b.imag)
    na = np.sum(np.abs(a) ** 2)
    nb = np.sum(np.abs(b) ** 2)
    return float(dot / (np.sqrt(na) * np.sqrt(nb) + 1e-30))


def max_abs_err(ref, test):
    """Max complex |ref - test| (aerial_coef_wmma.cu:349-350)."""
    d = np.asarray(ref).astype(np.complex128) - np.asarray(test).astype(np.complex128)
    return float(np.max(np.abs(d))) if d.size else 0.0


def evm(ref, test):
    """sqrt(sum|d|^2 / sum|ref|^2) (coef_rigor.cu:132-138); returns (ratio, dB = 20 log10 ratio)."""
    a = np.asarray(ref).astype(np.complex128)
    d = a - np.asarray(test).astype(np.complex128)
    num = float(np.sum(np.abs(d) ** 2))
    den = float(np.sum(np.abs(a) ** 2))
    r = np.sqrt(num / den) if den > 0 else 0.0
    return r, (20.0 * np.log10(r) if r > 0 else -np.inf)


def max_rel_err(ref, test, floor=1e-9):
    """Per-element relative error, elements with |ref|^2 <= floor skipped (coef_rigor.cu:135)."""
    a = np.asarray(ref).astype(np.complex128).ravel()
    d = a - np.asarray(test).astype(np.complex128).ravel()
    r2 = np.abs(a) ** 2
    m = r2 > floor
    return float(np.max(np.sqrt(np.abs(d[m]) ** 2 / r2[m]))) if np.any(m) else 0.0


def norm_rel_err(ref, test):
    """Norm-wise ||ref - test||_inf / ||ref||_inf over complex magnitudes."""
    a = np.asarray(ref).astype(np.complex128)
    d = a - np.asarray(test).astype(np.complex128)
    den = float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(d))) / den if den > 0 else 0.0


def percentile_nearest_rank(v, p):
    """sorted[min(n-1, ceil(p/100 * n) - 1)] (pct(), coef_rigor.cu:90-91)."""
    s = np.sort(np.asarray(v, dtype=np.float64))
    n = s.size
    i = min(n - 1, int(np.ceil(p / 100.0 * n)) - 1)
    return float(s[max(i, 0)])


def median_upper(v):
    """t[n/2] after sorting (median_ms(), aerial_coef_wmma.cu:357-361)."""
    s = np.sort(np.asarray(v, dtype=np.float64))
    return float(s[s.size // 2])


# ---------------------------------------------------------------------------------------------------------
# synthetic inputs of the benchmark configurations (numpy Generator; seeds recorded by callers)
# ---------------------------------------------------------------------------------------------------------
def cn(rng, shape, var=1.0):
    s = np.sqrt(var / 2.0)
    return (rng.standard_normal(shape) * s + 1j * rng.standard_normal(shape) * s).astype(np.complex64)


def rayleigh_h(rng, nSC, A, L):
    """H [nSC, A, L] with i.i.d. CN(0, 1/A) entries."""
    return cn(rng, (nSC, A, L), 1.0 / A)


def qam_map(bits, q):
    """TS 38.211 section 5.1 mapping. bits [..., 2q] (b0..b_{2q-1}) -> complex symbols, unit average power.
    Axis I uses b0, b2, ...; axis Q uses b1, b3, ...; t_{q-1} = 1 - 2c_{q-1}, t_i = (1 - 2c_i)(2^{q-1-i} - t_{i+1})."""
    bits = np.asarray(bits)
    norm = np.sqrt(2.0 * (4.0 ** q - 1.0) / 3.0)
    out = []
    for iq in range(2):
        c = bits[..., iq::2].astype(np.float64)
        t = 1.0 - 2.0 * c[..., q - 1]
        for i in range(q - 2, -1, -1):
            t = (1.0 - 2.0 * c[..., i]) * (2.0 ** (q - 1 - i) - t)
        out.append(t / norm)
    return (out[0] + 1j * out[1]).astype(np.complex128)


def chain_inputs(rng, nSC, A, L, S, q, snr_db):
    """Single-cell chain: random QAM symbols x [nSC, L, S] through Rayleigh H [nSC, A, L] plus AWGN.
    Y[sc, s, a] = sum_l H[sc, a, l] x[sc, l, s] + n,  n ~ CN(0, sigma2), sigma2 = 10^(-snr/10).
    Returns dict with H, Y (complex64), bits [nSC, L, S, 2q], x, sigma2 [nSC] float32."""
    sigma2 = 10.0 ** (-snr_db / 10.0)
    H = rayleigh_h(rng, nSC, A, L)
    bits = rng.integers(0, 2, size=(nSC, L, S, 2 * q), dtype=np.uint8)
    x = qam_map(bits, q)
    Y = np.einsum("nal,nls->nsa", H.astype(np.complex128), x)
    Y = (Y + cn(rng, Y.shape, sigma2)).astype(np.complex64)
    return {"H": H, "Y": Y, "bits": bits, "x": x, "sigma2": np.full(nSC, sigma2, dtype=np.float32)}


# ---------------------------------------------------------------------------------------------------------
# binary16 arithmetic and the SRS filter multiply (srs_wmma_bench.cu)
# ---------------------------------------------------------------------------------------------------------
def half_ops(a, b, c):
    """Element-wise binary16 mul, add, sub of a, b and fma(a, b, c), each rounded once to nearest even.
    Inputs and outputs are uint16 bit patterns."""
    a, b, c = (np.ascontiguousarray(x, dtype=np.uint16) for x in (a, b, c))
    outs = [np.empty_like(a) for _ in range(4)]
    lib().po_half_ops(_p(a), _p(b), _p(c), *[_p(o) for o in outs], a.size)
    return dict(zip(("mul", "add", "sub", "fma"), outs))


def _srs_dims(SRS, W, FOCC):
    nBlocks, nAnt, K, two = SRS.shape
    nPorts, foccLen, _ = FOCC.shape
    assert two == 2 and W.shape == (K, K, 2)
    return nBlocks, nAnt, K, nPorts, foccLen


def srs_filter_v1(SRS, W, FOCC, conj_w=False):
    """srs_filter_v1_scalar (srs_wmma_bench.cu:57-90) on binary16 bit patterns: SRS [nBlocks, nAnt, K, 2],
    W [K, K, 2], FOCC [nPorts, foccLen, 2] -> H_est [nBlocks, nAnt, K, nPorts, 2] (uint16)."""
    SRS, W, FOCC = (np.ascontiguousarray(x, dtype=np.uint16) for x in (SRS, W, FOCC))
    nBlocks, nAnt, K, nPorts, foccLen = _srs_dims(SRS, W, FOCC)
    out = np.empty((nBlocks, nAnt, K, nPorts, 2), dtype=np.uint16)
    lib().po_srs_filter_v1(_p(out), _p(SRS), _p(W), _p(FOCC), nAnt, nBlocks, K, nPorts, foccLen, int(conj_w))
    return out


def srs_filter_f64(SRS, W, FOCC, conj_w=False, prod_f16=False):
    """The same sum in double from the binary16 inputs -> complex128 [nBlocks, nAnt, K, nPorts]."""
    SRS, W, FOCC = (np.ascontiguousarray(x, dtype=np.uint16) for x in (SRS, W, FOCC))
    nBlocks, nAnt, K, nPorts, foccLen = _srs_dims(SRS, W, FOCC)
    out = np.empty((nBlocks, nAnt, K, nPorts, 2), dtype=np.float64)
    lib().po_srs_filter_f64(_p(out), _p(SRS), _p(W), _p(FOCC), nAnt, nBlocks, K, nPorts, foccLen, int(conj_w),
                            int(prod_f16))
    return out[..., 0] + 1j * out[..., 1]


def srs_bench_inputs(seed, nAnt, nBlocks, K, nPorts, foccLen):
    """srs_wmma_bench.cu:262-268 inputs: (SRS, W, FOCC) as binary16 bit patterns."""
    SRS = np.empty((nBlocks, nAnt, K, 2), dtype=np.uint16)
    W = np.empty((K, K, 2), dtype=np.uint16)
    FOCC = np.empty((nPorts, foccLen, 2), dtype=np.uint16)
    lib().po_srs_bench_inputs(seed, nAnt, nBlocks, K, nPorts, foccLen, _p(SRS), _p(W), _p(FOCC))
    return SRS, W, FOCC


def f16_complex(bits):
    """[..., 2] binary16 bit patterns -> complex128."""
    v = np.ascontiguousarray(bits, dtype=np.uint16).view(np.float16).astype(np.float64)
    return v[..., 0] + 1j * v[..., 1]


# ---------------------------------------------------------------------------------------------------------
# fronthaul block floating point (O-RAN U-plane compression format)
# ---------------------------------------------------------------------------------------------------------
def bfp_block_bytes(iq_width):
    return 1 + 3 * iq_width


def bfp_compress(x, iq_width, row_stride=None):
    """int32 samples [nSym, nAnt, nPrb*12, 2] -> packed rows uint8 [nSym*nAnt, row_stride]."""
    x = np.ascontiguousarray(x, dtype=np.int32)
    nSym, nAnt, nSC, _ = x.shape
    nPrb = nSC // 12
    rs = row_stride or nPrb * bfp_block_bytes(iq_width)
    out = np.zeros((nSym * nAnt, rs), dtype=np.uint8)
    lib().po_bfp_compress(_p(x), _p(out), nPrb, nSym, nAnt, iq_width, rs)
    return out


def bfp_decompress(packed, nPrb, nSym, nAnt, iq_width, scale=1.0):
    """packed rows [nSym*nAnt, row_stride] -> complex64 Y_rx [nPrb*12, nSym, nAnt]."""
    packed = np.ascontiguousarray(packed, dtype=np.uint8)
    Y = np.empty((nPrb * 12, nSym, nAnt, 2), dtype=np.float32)
    lib().po_bfp_decompress(_p(packed), _p(Y), nPrb, nSym, nAnt, iq_width, packed.shape[1], float(scale))
    return Y.view(np.complex64)[..., 0]


# ---------------------------------------------------------------------------------------------------------
# DMRS frequency interpolation: H_f[b][m][n] = sum_k W[m][k] X[b][k][n] (README.md:200 names the GEMM)
# ---------------------------------------------------------------------------------------------------------
def _fi_dims(W, X):
    M, K = W.shape
    nBatch, K2, NC = X.shape
    assert K2 == K, (W.shape, X.shape)
    return nBatch, M, K, NC


def freq_interp_f32(W, X):
    """FP32 chain, k ascending, fused cma() per k: W [M, K], X [nBatch, K, NC] complex64 -> [nBatch, M, NC]."""
    W, X = _c64(W), _c64(X)
    nBatch, M, K, NC = _fi_dims(W, X)
    out = np.empty((nBatch, M, NC), dtype=np.complex64)
    lib().po_freq_interp_f32(_p(W), _p(X), _p(out), nBatch, M, K, NC)
    return out


def freq_interp_f32_omp(W, X, out=None, nthreads=0):
    """The FP32 chain with OpenMP over the batch; returns (out, threads used)."""
    W, X = _c64(W), _c64(X)
    nBatch, M, K, NC = _fi_dims(W, X)
    if out is None:
        out = np.empty((nBatch, M, NC), dtype=np.complex64)
    used = lib().po_freq_interp_f32_omp(_p(W), _p(X), _p(out), nBatch, M, K, NC, int(nthreads))
    return out, used


def freq_interp_f64(W, X):
    """FP64 sum of exact products of the operands as given (complex128 accepted): -> complex128."""
    W = np.ascontiguousarray(W, dtype=np.complex128)
    X = np.ascontiguousarray(X, dtype=np.complex128)
    nBatch, M, K, NC = _fi_dims(W, X)
    out = np.empty((nBatch, M, NC), dtype=np.complex128)
    lib().po_freq_interp_f64(_p(W), _p(X), _p(out), nBatch, M, K, NC)
    return out


def round_bf16(a):
    """float32 values rounded to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def round_bf16_complex(z):
    """complex64 -> complex64 with real and imaginary parts rounded to bfloat16."""
    z = np.ascontiguousarray(z, dtype=np.complex64)
    return (round_bf16(z.real) + 1j * round_bf16(z.imag)).astype(np.complex64)


# ---------------------------------------------------------------------------------------------------------
# descrambling and rate-recovery layout (TS 38.211 5.2.1, 6.3.1; TS 38.212 5.4.2)
# ---------------------------------------------------------------------------------------------------------
def gold_sequence(c_init, n_bits):
    """c(0 .. n_bits - 1) of TS 38.211 5.2.1 (uint8, one bit per entry)."""
    out = np.empty(n_bits, dtype=np.uint8)
    lib().po_gold_sequence(int(c_init) & 0xFFFFFFFF, int(n_bits), _p(out))
    return out


def cw_block_sizes(G, nu, qm, C):
    """E_r of the C code blocks (TS 38.212 5.4.2.1); None when C is out of range."""
    E = np.empty(max(int(C), 1), dtype=np.int32)
    return None if lib().po_cw_block_sizes(int(G), int(nu), int(qm), int(C), _p(E)) else E[:C].copy()


def cw_layout(llr, layer_cw, c_seqs=None, n_code_blocks=None):
    """The demapper's LLRs [nSC, L, S, 2q] (float32) moved into per-codeword arrays: codeword k holds the layers
    with layer_cw[l] == k; c_seqs[k] (uint8 bits or None) descrambles it; n_code_blocks[k] > 0 de-interleaves it per
    code block. Returns a list of float32 arrays of G_k LLRs."""
    llr = np.ascontiguousarray(llr, dtype=np.float32)
    nSC, L, S, qm = llr.shape
    lc = np.ascontiguousarray(layer_cw, dtype=np.int32)
    ncw = int(lc.max()) + 1
    nu = np.bincount(lc, minlength=ncw)
    ncb = np.ascontiguousarray(n_code_blocks if n_code_blocks is not None else [0] * ncw, dtype=np.int32)
    outs = [np.full(nSC * S * int(nu[k]) * qm, np.nan, dtype=np.float32) for k in range(ncw)]
    cs = None
    keep = []
    if c_seqs is not None:
        arr = []
        for k in range(ncw):
            if c_seqs[k] is None:
                arr.append(None)
            else:
                a = np.ascontiguousarray(c_seqs[k], dtype=np.uint8)
                keep.append(a)
                arr.append(a.ctypes.data)
        cs = (ctypes.c_void_p * ncw)(*arr)
    op = (ctypes.c_void_p * ncw)(*[o.ctypes.data for o in outs])
    lib().po_cw_layout(_p(llr), nSC, L, S, qm // 2, _p(lc), ncw, cs, _p(ncb), op)
    return outs
