"""ctypes binding of libpusch_eq.so, the C ABI declared in include/pusch_eq.h.

Python is the test and benchmark harness here: torch supplies device memory and streams, every numeric
operation runs in the HIP kernels behind the C ABI. Nothing in this module computes a result on the host,
and there is no fallback: if the shared library is missing, `lib()` raises.

Tensor conventions follow the reference kernels (complex64 interleaved re/im):
    C [nSC, nLayer, nAnt], Y_rx [nSC, nSym, nAnt], Y_eq [nSC, nLayer, nSym]      (aerial_coef_wmma.cu:72-74)
    H [nSC, nAnt, nLayer], G [nSC, nLayer, nLayer]                             (aerial_bottleneck.cu:57-59)
    llr [nSC, nLayer, nSym, 2*q] float32, or [N, 2*q] binary16                  (aerial_llr_bench_v2.cu:117)
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpusch_eq.so")

PE_OK, PE_ERR_INVALID, PE_ERR_UNSUPPORTED, PE_ERR_DEVICE, PE_ERR_NOMEM, PE_ERR_CAPACITY = range(6)
PREC_F32, PREC_F16, PREC_BF16, PREC_F16_SPLIT, PREC_BF16_SPLIT, PREC_F16_C16 = range(6)
PREC_NAMES = {"f32": PREC_F32, "f16": PREC_F16, "bf16": PREC_BF16, "f16_split": PREC_F16_SPLIT,
              "bf16_split": PREC_BF16_SPLIT, "f16_c16": PREC_F16_C16}
MMSE_UNBIASED = 1
SCFLAG_FP64 = 1
SCFLAG_SINGULAR = 2

# every function the header declares, with its ctypes signature
_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint
_F = ctypes.c_float
SIGNATURES = {
    "pe_version": ([], ctypes.c_char_p),
    "pe_last_error": ([], ctypes.c_char_p),
    "pe_status_string": ([_I], ctypes.c_char_p),
    "pe_init": ([_I], _I),
    "pe_tune_load": ([ctypes.c_char_p], _I),
    "pe_coef_apply": ([_P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "pe_coef_apply_v1": ([_P, _P, _P, _I, _I, _P], _I),
    "pe_coef_apply_range": ([_P, _P, _P, _I, _I, _I, _I, _I, _I, _P], _I),
    "pe_gram": ([_P, _P, _P, _I, _I, _I, _P], _I),
    "pe_gram_mmse": ([_P, _P, _P, _P, _P, _I, _I, _I, _U, _F, _P], _I),
    "pe_llr_demap_f16": ([_P, _P, _P, _I, _I, _P], _I),
    "pe_llr_demap": ([_P, _P, _P, _I, _I, _I, _I, _P], _I),
    "pe_coef_apply_demap": ([_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P], _I),
    "pe_coef_apply_demap_f16": ([_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P], _I),
    "pe_srs_filter": ([_P, _P, _P, _P, _I, _I, _I, _I, _I, _U, _P], _I),
    "pe_hbm_probe": ([_P, _P, ctypes.c_size_t, _I, _I, _P], _I),
    "pe_bfp_decompress": ([_P, _P, _I, _I, _I, _I, ctypes.c_size_t, _F, _P], _I),
    "pe_freq_interp": ([_P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "pe_interp_mmse_apply_demap": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _U, _F, _P], _I),
    "pe_coef_to_c16": ([_P, _P, _P, _I, _I, _I, _P], _I),
    "pe_gram_mmse_c16": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _U, _F, _P], _I),
    "pe_coef_apply_c16": ([_P, _P, _P, _P, _I, _I, _I, _I, _P], _I),
    "pe_coef_apply_demap_c16": ([_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _U, _P], _I),
    "pe_mmse_apply_demap": ([_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _U, _F, _P], _I),
    "pe_mmse_apply_demap_noalloc": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _U, _F, _P], _I),
    "pe_gold_sequence": ([_U, _P, _I, _P], _I),
    "pe_coef_apply_demap_cw": ([_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _U, _P], _I),
    "pe_slot_plan_create": ([_P, _I, _U, _P], _I),
    "pe_slot_plan_update": ([_P, _P, _I, _P], _I),
    "pe_slot_plan_launch": ([_P, _P], _I),
    "pe_slot_plan_graph_launch": ([_P, _P], _I),
    "pe_slot_plan_destroy": ([_P], None),
    "pe_slot_plan_info": ([_P, _P, _P], _I),
    "pe_serve_run": ([_P, _P, _P, _P], _I),
}


class CellDesc(ctypes.Structure):
    """pe_cell_desc of include/pusch_eq.h."""
    _fields_ = [("H", _P), ("sigma2", _P), ("C", _P), ("noise_inv", _P), ("Y_rx", _P), ("llr", _P),
                ("nSC", _I), ("nLayer", _I), ("nSym", _I), ("nAnt", _I), ("nPamBits", _I), ("prec", _I)]


class PeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("pusch_eq status %d: %s" % (status, msg))
        self.status = status


_lib = None


def lib():
    """Loads libpusch_eq.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime (libamdhip64.so.7). Importing torch first makes this library bind
        # to that same runtime instance, so device pointers and streams from torch are valid here; loading
        # this library first would pull a second runtime from /opt/rocm into the process.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with __graft_entry__.build()" % LIB_PATH)
        h = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def _check(status):
    if status != PE_OK:
        raise PeError(status, lib().pe_last_error().decode())


def _ptr(t):
    """A tensor's device address for a c_void_p argument (plain ints convert; None is NULL)."""
    return t.data_ptr() if t is not None else None


def _need(t, name, dtype, shape=None, min_numel=None):
    """Checks a tensor handed to the library before any launch: its dtype (a torch dtype, or an element size in
    bytes for the binary16-pair buffers), contiguity, and shape (None entries match any size) or minimum number of
    elements. Every wrapper then checks with _same_device that all its tensors are on one GPU, so dtype and shape
    mistakes are reported whatever the device. Raises TypeError / ValueError; returns t. (Written for speed: it
    runs for every tensor of every call.)"""
    if not isinstance(t, _tensor_type()):
        raise TypeError("%s must be a torch tensor, got %s" % (name, type(t).__name__))
    if type(dtype) is int:
        if t.element_size() != dtype:
            raise TypeError("%s: elements of %d bytes expected, got %s" % (name, dtype, t.dtype))
    elif t.dtype != dtype:
        raise TypeError("%s: dtype %s expected, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if shape is not None:
        got = t.shape
        ok = len(got) == len(shape)
        if ok:
            for w, g in zip(shape, got):
                if w is not None and w != g:
                    ok = False
                    break
        if not ok:
            raise ValueError("%s: shape %s expected, got %s" % (name, tuple(shape), tuple(got)))
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError("%s: at least %d elements expected, got %d" % (name, min_numel, t.numel()))
    return t


_TENSOR = None


def _tensor_type():
    global _TENSOR
    if _TENSOR is None:
        import torch
        _TENSOR = torch.Tensor
    return _TENSOR


def _same_device(*ts):
    """All given tensors (None entries skipped) on one GPU."""
    first = None
    for t in ts:
        if t is None:
            continue
        d = t.device
        if first is None:
            first = d
            if d.type != "cuda":
                raise ValueError("tensors must be GPU tensors, got device %s" % d)
        elif d != first:
            raise ValueError("tensors on different devices: %s" % sorted({str(u.device) for u in ts if u is not None}))


def _q(n_pam_bits):
    q = int(n_pam_bits)
    if not 1 <= q <= 4:
        raise ValueError("n_pam_bits = %d outside [1, 4]" % q)
    return q


def _stream(stream):
    """The HIP stream handle of a torch stream (the current stream when None), for a c_void_p argument."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _prec(p):
    return PREC_NAMES[p] if isinstance(p, str) else int(p)


def init(device=0):
    _check(lib().pe_init(device))


def version():
    return lib().pe_version().decode()


def coef_apply(C, Y, out=None, prec="f32", stream=None):
    """Y_eq[sc, l, s] = sum_a C[sc, l, a] * Y[sc, s, a]."""
    import torch
    c64 = torch.complex64
    nSC, L, A = _need(C, "C", c64, (None, None, None)).shape
    S = _need(Y, "Y", c64, (nSC, None, A)).shape[1]
    if out is None:
        out = torch.empty((nSC, L, S), dtype=c64, device=C.device)
    _need(out, "out", c64, (nSC, L, S))
    _same_device(C, Y, out)
    _check(lib().pe_coef_apply(_ptr(C), _ptr(Y), _ptr(out), nSC, L, S, A, _prec(prec), _stream(stream)))
    return out


def coef_apply_v1(C, Y, out, nSC, nSym, stream=None):
    """The reference's exact signature (8 layers, 64 antennas)."""
    import torch
    c64 = torch.complex64
    _need(C, "C", c64, min_numel=nSC * 8 * 64)
    _need(Y, "Y", c64, min_numel=nSC * nSym * 64)
    _need(out, "out", c64, min_numel=nSC * 8 * nSym)
    _same_device(C, Y, out)
    _check(lib().pe_coef_apply_v1(_ptr(C), _ptr(Y), _ptr(out), nSC, nSym, _stream(stream)))
    return out


def coef_apply_range(C, Y, out, sc_offset, n_sc, prec="f32", stream=None):
    """Subcarriers [sc_offset, sc_offset + n_sc) of full-size tensors C [N, L, A], Y [N, S, A], out [N, L, S]."""
    import torch
    c64 = torch.complex64
    N, L, A = _need(C, "C", c64, (None, None, None)).shape
    S = _need(Y, "Y", c64, (N, None, A)).shape[1]
    _need(out, "out", c64, (N, L, S))
    if sc_offset < 0 or n_sc < 0 or sc_offset + n_sc > N:
        raise ValueError("range [%d, %d) outside the %d subcarriers" % (sc_offset, sc_offset + n_sc, N))
    _same_device(C, Y, out)
    _check(lib().pe_coef_apply_range(_ptr(C), _ptr(Y), _ptr(out), n_sc, sc_offset, L, S, A, _prec(prec),
                                     _stream(stream)))
    return out


def gram(M, H, out=None, stream=None):
    """G[sc] = M[sc] @ H[sc] (plain product)."""
    import torch
    c64 = torch.complex64
    nSC, L, A = _need(M, "M", c64, (None, None, None)).shape
    _need(H, "H", c64, (nSC, A, L))
    if out is None:
        out = torch.empty((nSC, L, L), dtype=c64, device=M.device)
    _need(out, "out", c64, (nSC, L, L))
    _same_device(M, H, out)
    _check(lib().pe_gram(_ptr(M), _ptr(H), _ptr(out), nSC, L, A, _stream(stream)))
    return out


def gram_mmse(H, sigma2, unbiased=True, cond_threshold=0.0, C=None, noise_inv=None, flags=None, stream=None):
    """C = (H^H H + sigma2 I)^-1 H^H per subcarrier; returns (C, noise_inv, flags)."""
    import torch
    nSC, A, L = _need(H, "H", torch.complex64, (None, None, None)).shape
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    dev = H.device
    if C is None:
        C = torch.empty((nSC, L, A), dtype=torch.complex64, device=dev)
    if noise_inv is None:
        noise_inv = torch.empty((nSC, L), dtype=torch.float32, device=dev)
    if flags is None:
        flags = torch.empty((nSC,), dtype=torch.uint8, device=dev)
    _need(C, "C", torch.complex64, (nSC, L, A))
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    _need(flags, "flags", torch.uint8, (nSC,))
    _same_device(H, sigma2, C, noise_inv, flags)
    _check(lib().pe_gram_mmse(_ptr(H), _ptr(sigma2), _ptr(C), _ptr(noise_inv), _ptr(flags), nSC, L, A,
                              MMSE_UNBIASED if unbiased else 0, float(cond_threshold), _stream(stream)))
    return C, noise_inv, flags


def llr_demap_f16(soft_half, noise_inv, n_pam_bits, out=None, stream=None):
    """soft_half: [N, 2] binary16 bits (re, im) in any 2-byte dtype; returns [N, 2*q] binary16 LLRs."""
    import torch
    q = _q(n_pam_bits)
    N = _need(soft_half, "soft_half", 2, (None, 2)).shape[0]
    _need(noise_inv, "noise_inv", torch.float32, (N,))
    if out is None:
        out = torch.empty((N, 2 * q), dtype=torch.float16, device=soft_half.device)
    _need(out, "out", 2, (N, 2 * q))
    _same_device(soft_half, noise_inv, out)
    _check(lib().pe_llr_demap_f16(_ptr(soft_half), _ptr(noise_inv), _ptr(out), N, n_pam_bits, _stream(stream)))
    return out


def llr_demap(Yeq, noise_inv, n_pam_bits, out=None, stream=None):
    import torch
    q = _q(n_pam_bits)
    nSC, L, S = _need(Yeq, "Yeq", torch.complex64, (None, None, None)).shape
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=torch.float32, device=Yeq.device)
    _need(out, "out", torch.float32, (nSC, L, S, 2 * q))
    _same_device(Yeq, noise_inv, out)
    _check(lib().pe_llr_demap(_ptr(Yeq), _ptr(noise_inv), _ptr(out), nSC, L, S, n_pam_bits, _stream(stream)))
    return out


def coef_apply_demap(C, Y, noise_inv, n_pam_bits, prec="f16", out=None, stream=None):
    import torch
    q = _q(n_pam_bits)
    nSC, L, A = _need(C, "C", torch.complex64, (None, None, None)).shape
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=torch.float32, device=C.device)
    _need(out, "out", torch.float32, (nSC, L, S, 2 * q))
    _same_device(C, Y, noise_inv, out)
    _check(lib().pe_coef_apply_demap(_ptr(C), _ptr(Y), _ptr(noise_inv), _ptr(out), nSC, L, S, A, n_pam_bits,
                                     _prec(prec), _stream(stream)))
    return out


def coef_apply_demap_f16(C, Y, noise_inv, n_pam_bits, prec="f16", out=None, stream=None):
    """Fused apply + demap writing binary16 LLRs (the decoder wire format); returns a torch.float16 tensor
    [nSC, nLayer, nSym, 2*n_pam_bits] (pe_coef_apply_demap_f16)."""
    import torch
    q = _q(n_pam_bits)
    nSC, L, A = _need(C, "C", torch.complex64, (None, None, None)).shape
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=torch.float16, device=C.device)
    _need(out, "out", 2, (nSC, L, S, 2 * q))
    _same_device(C, Y, noise_inv, out)
    _check(lib().pe_coef_apply_demap_f16(_ptr(C), _ptr(Y), _ptr(noise_inv), _ptr(out), nSC, L, S, A,
                                         n_pam_bits, _prec(prec), _stream(stream)))
    return out


EQ_WRITE_C = 2
EQ_LLR_F16 = 4
SLOT_SOLVE_RECORDS = 2


def mmse_apply_demap(H, sigma2, Y, n_pam_bits, prec="f16", unbiased=True, cond_threshold=0.0, write_c=False,
                     llr_f16=False, C=None, noise_inv=None, flags=None, out=None, stream=None):
    """MMSE coefficients and the fused apply + demap in one pass (pe_mmse_apply_demap): H [nSC, nAnt, nLayer],
    Y [nSC, nSym, nAnt] -> LLRs [nSC, nLayer, nSym, 2*n_pam_bits] (float32, or float16 with llr_f16). C is the
    workspace for subcarriers solved in FP64 (every row with write_c). Returns (llr, C, noise_inv, flags)."""
    import torch
    q = _q(n_pam_bits)
    nSC, A, L = _need(H, "H", torch.complex64, (None, None, None)).shape
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    dev = H.device
    ldt = torch.float16 if llr_f16 else torch.float32
    if C is None:
        C = torch.empty((nSC, L, A), dtype=torch.complex64, device=dev)
    if noise_inv is None:
        noise_inv = torch.empty((nSC, L), dtype=torch.float32, device=dev)
    if flags is None:
        flags = torch.empty((nSC,), dtype=torch.uint8, device=dev)
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=ldt, device=dev)
    _need(C, "C", torch.complex64, (nSC, L, A))
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    _need(flags, "flags", torch.uint8, (nSC,))
    _need(out, "out", ldt, (nSC, L, S, 2 * q))
    _same_device(H, sigma2, Y, C, noise_inv, flags, out)
    f = (MMSE_UNBIASED if unbiased else 0) | (EQ_WRITE_C if write_c else 0) | (EQ_LLR_F16 if llr_f16 else 0)
    _check(lib().pe_mmse_apply_demap(_ptr(H), _ptr(sigma2), _ptr(Y), _ptr(C), _ptr(noise_inv), _ptr(flags),
                                     _ptr(out), nSC, L, S, A, n_pam_bits, _prec(prec), f, float(cond_threshold),
                                     _stream(stream)))
    return out, C, noise_inv, flags


class CwLayout(ctypes.Structure):
    """pe_cw_layout of include/pusch_eq.h."""
    _fields_ = [("layer_cw", _I * 8), ("llr", _P * 8), ("c_seq", _P * 8), ("nCodeBlocks", _I * 8)]


def gold_sequence(c_init, n_bits, out=None, stream=None):
    """TS 38.211 5.2.1 sequence c(0 .. n_bits - 1) for c_init (pe_gold_sequence): int32 tensor of ceil(n_bits / 32)
    words, bit g % 32 of word g / 32 = c(g)."""
    import torch
    nw = (int(n_bits) + 31) // 32
    if out is None:
        out = torch.empty(nw, dtype=torch.int32, device=torch.device("cuda", torch.cuda.current_device()))
    _need(out, "out", 4, min_numel=nw)
    _same_device(out)
    _check(lib().pe_gold_sequence(int(c_init) & 0xFFFFFFFF, _ptr(out), int(n_bits), _stream(stream)))
    return out


def coef_apply_demap_cw(C, Y, noise_inv, n_pam_bits, layer_cw, c_seqs=None, n_code_blocks=None, prec="f16",
                        llr_f16=False, outs=None, stream=None):
    """Fused apply + demap with the LLRs descrambled and laid out per codeword (pe_coef_apply_demap_cw). layer_cw:
    codeword of each layer; c_seqs: per codeword an int32 word tensor from gold_sequence, or None; n_code_blocks:
    per codeword 0 (codeword order) or the number of code blocks to de-interleave. Returns the list of per-codeword
    LLR tensors (float32, or float16 with llr_f16), G_k = nSC nSym nu_k 2q values each."""
    import torch
    q = _q(n_pam_bits)
    nSC, L, A = _need(C, "C", torch.complex64, (None, None, None)).shape
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    if len(layer_cw) != L:
        raise ValueError("layer_cw has %d entries for %d layers" % (len(layer_cw), L))
    ncw = max(layer_cw) + 1
    nu = [list(layer_cw).count(k) for k in range(ncw)]
    ldt = torch.float16 if llr_f16 else torch.float32
    G = [nSC * S * nu[k] * 2 * q for k in range(ncw)]
    if outs is None:
        outs = [torch.empty(G[k], dtype=ldt, device=C.device) for k in range(ncw)]
    for k in range(ncw):
        _need(outs[k], "outs[%d]" % k, ldt, (G[k],))
        if c_seqs is not None and c_seqs[k] is not None:
            _need(c_seqs[k], "c_seqs[%d]" % k, 4, min_numel=(G[k] + 31) // 32)
    _same_device(C, Y, noise_inv, *outs, *[c for c in (c_seqs or []) if c is not None])
    lay = CwLayout()
    for l in range(L):
        lay.layer_cw[l] = int(layer_cw[l])
    for k in range(ncw):
        lay.llr[k] = outs[k].data_ptr()
        lay.c_seq[k] = c_seqs[k].data_ptr() if c_seqs is not None and c_seqs[k] is not None else None
        lay.nCodeBlocks[k] = int(n_code_blocks[k]) if n_code_blocks is not None else 0
    _check(lib().pe_coef_apply_demap_cw(_ptr(C), _ptr(Y), _ptr(noise_inv), ctypes.byref(lay), nSC, L, S, A, q,
                                        _prec(prec), LLR_F16 if llr_f16 else 0, _stream(stream)))
    return outs


NOALLOC_SCRATCH_BYTES = 32 << 20


def mmse_apply_demap_noalloc(H, sigma2, Y, n_pam_bits, prec="f16", unbiased=True, cond_threshold=0.0,
                             llr_f16=False, noise_inv=None, flags=None, out=None, stream=None):
    """pe_mmse_apply_demap_noalloc: the one-pass chain with its workspace in the library's fixed device scratch
    (nSC * nLayer * nAnt * 8 <= NOALLOC_SCRATCH_BYTES); the library allocates nothing. Output tensors are allocated
    here by torch unless given. Returns (llr, noise_inv, flags)."""
    import torch
    q = _q(n_pam_bits)
    nSC, A, L = _need(H, "H", torch.complex64, (None, None, None)).shape
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    dev = H.device
    ldt = torch.float16 if llr_f16 else torch.float32
    if noise_inv is None:
        noise_inv = torch.empty((nSC, L), dtype=torch.float32, device=dev)
    if flags is None:
        flags = torch.empty((nSC,), dtype=torch.uint8, device=dev)
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=ldt, device=dev)
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    _need(flags, "flags", torch.uint8, (nSC,))
    _need(out, "out", ldt, (nSC, L, S, 2 * q))
    _same_device(H, sigma2, Y, noise_inv, flags, out)
    f = (MMSE_UNBIASED if unbiased else 0) | (EQ_LLR_F16 if llr_f16 else 0)
    _check(lib().pe_mmse_apply_demap_noalloc(_ptr(H), _ptr(sigma2), _ptr(Y), _ptr(noise_inv), _ptr(flags), _ptr(out),
                                             nSC, L, S, A, q, _prec(prec), f, float(cond_threshold), _stream(stream)))
    return out, noise_inv, flags


SRS_MFMA = 1
SRS_CONJ_W = 2


def srs_filter(SRS, W, FOCC, mfma=False, conj_w=False, out=None, stream=None):
    """SRS filter multiply (pe_srs_filter). SRS [nBlocks, nAnt, K], W [K, K], FOCC [nPorts, foccLen] as
    binary16 complex pairs: torch int32 tensors whose 32 bits hold {re, im} binary16 (or any 4-byte dtype
    viewed so). Returns H_est [nBlocks, nAnt, K, nPorts] of the same dtype."""
    import torch
    nBlocks, nAnt, K = _need(SRS, "SRS", 4, (None, None, None)).shape
    _need(W, "W", 4, (K, K))
    nPorts, foccLen = _need(FOCC, "FOCC", 4, (None, None)).shape
    if out is None:
        out = torch.empty((nBlocks, nAnt, K, nPorts), dtype=SRS.dtype, device=SRS.device)
    _need(out, "out", 4, (nBlocks, nAnt, K, nPorts))
    _same_device(SRS, W, FOCC, out)
    flags = (SRS_MFMA if mfma else 0) | (SRS_CONJ_W if conj_w else 0)
    _check(lib().pe_srs_filter(_ptr(out), _ptr(SRS), _ptr(W), _ptr(FOCC), nAnt, nBlocks, K, nPorts, foccLen, flags,
                               _stream(stream)))
    return out


LLR_F16 = 1


def coef_to_c16(C, C16=None, c_exp=None, stream=None):
    """complex64 coefficients [nSC, L, A] -> (C16 int32 [nSC, L, A] holding binary16 {re, im} pairs,
    c_exp int32 [nSC]) with C = C16 * 2^c_exp (pe_coef_to_c16)."""
    import torch
    nSC, L, A = _need(C, "C", torch.complex64, (None, None, None)).shape
    if C16 is None:
        C16 = torch.empty((nSC, L, A), dtype=torch.int32, device=C.device)
    if c_exp is None:
        c_exp = torch.empty((nSC,), dtype=torch.int32, device=C.device)
    _need(C16, "C16", 4, (nSC, L, A))
    _need(c_exp, "c_exp", torch.int32, (nSC,))
    _same_device(C, C16, c_exp)
    _check(lib().pe_coef_to_c16(_ptr(C), _ptr(C16), _ptr(c_exp), nSC, L, A, _stream(stream)))
    return C16, c_exp


def gram_mmse_c16(H, sigma2, unbiased=True, cond_threshold=0.0, C16=None, c_exp=None, noise_inv=None, flags=None,
                  stream=None):
    """MMSE coefficients written as binary16 pairs + per-subcarrier exponent (pe_gram_mmse_c16). Returns
    (C16 int32 [nSC, L, A], c_exp int32 [nSC], noise_inv [nSC, L], sc_flags uint8 [nSC])."""
    import torch
    nSC, A, L = _need(H, "H", torch.complex64, (None, None, None)).shape
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    dev = H.device
    if C16 is None:
        C16 = torch.empty((nSC, L, A), dtype=torch.int32, device=dev)
    if c_exp is None:
        c_exp = torch.empty((nSC,), dtype=torch.int32, device=dev)
    if noise_inv is None:
        noise_inv = torch.empty((nSC, L), dtype=torch.float32, device=dev)
    if flags is None:
        flags = torch.empty((nSC,), dtype=torch.uint8, device=dev)
    _need(C16, "C16", 4, (nSC, L, A))
    _need(c_exp, "c_exp", torch.int32, (nSC,))
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    _need(flags, "flags", torch.uint8, (nSC,))
    _same_device(H, sigma2, C16, c_exp, noise_inv, flags)
    _check(lib().pe_gram_mmse_c16(_ptr(H), _ptr(sigma2), _ptr(C16), _ptr(c_exp), _ptr(noise_inv), _ptr(flags), nSC,
                                  L, A, MMSE_UNBIASED if unbiased else 0, float(cond_threshold), _stream(stream)))
    return C16, c_exp, noise_inv, flags


def coef_apply_c16(C16, c_exp, Y, out=None, stream=None):
    import torch
    nSC, L, A = _need(C16, "C16", 4, (None, None, None)).shape
    _need(c_exp, "c_exp", torch.int32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    if out is None:
        out = torch.empty((nSC, L, S), dtype=torch.complex64, device=Y.device)
    _need(out, "out", torch.complex64, (nSC, L, S))
    _same_device(C16, c_exp, Y, out)
    _check(lib().pe_coef_apply_c16(_ptr(C16), _ptr(c_exp), _ptr(Y), _ptr(out), nSC, L, S, A, _stream(stream)))
    return out


def coef_apply_demap_c16(C16, c_exp, Y, noise_inv, n_pam_bits, llr_f16=False, out=None, stream=None):
    """Fused apply + demap from binary16 coefficients (pe_coef_apply_demap_c16): float32 LLRs, or binary16
    (torch.float16) with llr_f16."""
    import torch
    q = _q(n_pam_bits)
    nSC, L, A = _need(C16, "C16", 4, (None, None, None)).shape
    _need(c_exp, "c_exp", torch.int32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    ldt = torch.float16 if llr_f16 else torch.float32
    if out is None:
        out = torch.empty((nSC, L, S, 2 * q), dtype=ldt, device=Y.device)
    _need(out, "out", ldt, (nSC, L, S, 2 * q))
    _same_device(C16, c_exp, Y, noise_inv, out)
    _check(lib().pe_coef_apply_demap_c16(_ptr(C16), _ptr(c_exp), _ptr(Y), _ptr(noise_inv), _ptr(out), nSC, L, S, A,
                                         n_pam_bits, LLR_F16 if llr_f16 else 0, _stream(stream)))
    return out


def bfp_decompress(packed, nPrb, nSym, nAnt, iq_width, scale=1.0, out=None, stream=None):
    """Fronthaul block floating point rows (uint8 tensor [nSym*nAnt, row_stride]) -> complex64 Y_rx
    [nPrb*12, nSym, nAnt] (pe_bfp_decompress)."""
    import torch
    _need(packed, "packed", torch.uint8, (nSym * nAnt, None))
    if out is None:
        out = torch.empty((nPrb * 12, nSym, nAnt), dtype=torch.complex64, device=packed.device)
    _need(out, "out", torch.complex64, (nPrb * 12, nSym, nAnt))
    _same_device(packed, out)
    _check(lib().pe_bfp_decompress(_ptr(packed), _ptr(out), nPrb, nSym, nAnt, iq_width, packed.shape[1],
                                   float(scale), _stream(stream)))
    return out


def freq_interp(W, X, prec="f32", out=None, stream=None):
    """DMRS frequency interpolation (pe_freq_interp): W [nOut, nIn], X [nBatch, nIn, nCol] complex64 ->
    H_f [nBatch, nOut, nCol], H_f[b] = W @ X[b]. prec "f32" (bit-exact chain), "bf16" or "bf16_split" (MFMA)."""
    import torch
    nOut, nIn = _need(W, "W", torch.complex64, (None, None)).shape
    nBatch, _, nCol = _need(X, "X", torch.complex64, (None, nIn, None)).shape
    if out is None:
        out = torch.empty((nBatch, nOut, nCol), dtype=torch.complex64, device=X.device)
    _need(out, "out", torch.complex64, (nBatch, nOut, nCol))
    _same_device(W, X, out)
    _check(lib().pe_freq_interp(_ptr(W), _ptr(X), _ptr(out), nBatch, nOut, nIn, nCol, _prec(prec), _stream(stream)))
    return out


def interp_mmse_apply_demap(W, X, sigma2, Y, n_pam_bits, interp_prec="bf16", prec="f16", unbiased=True,
                            cond_threshold=0.0, write_c=False, llr_f16=False, H=None, C=None, noise_inv=None,
                            flags=None, out=None, stream=None):
    """DMRS estimates to LLRs with the frequency interpolation fused into the MMSE solve
    (pe_interp_mmse_apply_demap): W [64, nIn], X [nBlk, nAnt, nIn, 4], sigma2 [64 nBlk], Y [64 nBlk, nSym, nAnt]
    complex64 -> LLRs [64 nBlk, 4, nSym, 2 n_pam_bits], with H [64 nBlk, nAnt, 4] written (the values
    freq_interp(W, X.reshape(-1, nIn, 4), interp_prec) computes, in the equalizer's layout). The LLRs, noise_inv and
    flags are those of mmse_apply_demap on that H. Returns (llr, H, C, noise_inv, flags)."""
    import torch
    q = _q(n_pam_bits)
    _need(W, "W", torch.complex64, (64, None))
    nIn = W.shape[1]
    nBlk, A = _need(X, "X", torch.complex64, (None, None, nIn, 4)).shape[:2]
    nSC = 64 * nBlk
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    dev = X.device
    ldt = torch.float16 if llr_f16 else torch.float32
    if H is None:
        H = torch.empty((nSC, A, 4), dtype=torch.complex64, device=dev)
    if C is None:
        C = torch.empty((nSC, 4, A), dtype=torch.complex64, device=dev)
    if noise_inv is None:
        noise_inv = torch.empty((nSC, 4), dtype=torch.float32, device=dev)
    if flags is None:
        flags = torch.empty((nSC,), dtype=torch.uint8, device=dev)
    if out is None:
        out = torch.empty((nSC, 4, S, 2 * q), dtype=ldt, device=dev)
    _need(H, "H", torch.complex64, (nSC, A, 4))
    _need(C, "C", torch.complex64, (nSC, 4, A))
    _need(noise_inv, "noise_inv", torch.float32, (nSC, 4))
    _need(flags, "flags", torch.uint8, (nSC,))
    _need(out, "out", ldt, (nSC, 4, S, 2 * q))
    _same_device(W, X, sigma2, Y, H, C, noise_inv, flags, out)
    f = (MMSE_UNBIASED if unbiased else 0) | (EQ_WRITE_C if write_c else 0) | (EQ_LLR_F16 if llr_f16 else 0)
    _check(lib().pe_interp_mmse_apply_demap(_ptr(W), _ptr(X), _ptr(sigma2), _ptr(Y), _ptr(H), _ptr(C),
                                            _ptr(noise_inv), _ptr(flags), _ptr(out), nBlk, nIn, S, A, q,
                                            _prec(interp_prec), _prec(prec), f, float(cond_threshold),
                                            _stream(stream)))
    return out, H, C, noise_inv, flags


def hbm_probe(src, dst, read_only=False, grid=4096, nbytes=None, stream=None):
    """Streaming probe (pe_hbm_probe): copy src -> dst, or read src (dst: at least `grid` words)."""
    n = src.numel() * src.element_size() if nbytes is None else nbytes
    _need(src, "src", src.element_size(), min_numel=-(-n // src.element_size()))
    if read_only:
        _need(dst, "dst", dst.element_size(), min_numel=-(-4 * grid // dst.element_size()))
    else:
        _need(dst, "dst", dst.element_size(), min_numel=-(-n // dst.element_size()))
    _same_device(src, dst)
    _check(lib().pe_hbm_probe(_ptr(src), _ptr(dst), n, 1 if read_only else 0, grid, _stream(stream)))


def cell_desc(C, noise_inv, Y, llr, n_pam_bits, prec="f16", H=None, sigma2=None):
    """Builds a pe_cell_desc from torch tensors (H/sigma2 None: coefficients are inputs). With prec "f16_c16"
    the complex64 C tensor's storage holds the binary16 coefficients and exponents (see c16_view)."""
    import torch
    q = _q(n_pam_bits)
    nSC, L, A = _need(C, "C", torch.complex64, (None, None, None)).shape
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _need(noise_inv, "noise_inv", torch.float32, (nSC, L))
    _need(llr, "llr", torch.float32, (nSC, L, S, 2 * q))
    if (H is None) != (sigma2 is None):
        raise ValueError("H and sigma2 go together")
    if H is not None:
        _need(H, "H", torch.complex64, (nSC, A, L))
        _need(sigma2, "sigma2", torch.float32, (nSC,))
    _same_device(C, Y, noise_inv, llr, H, sigma2)
    return CellDesc(H.data_ptr() if H is not None else None, sigma2.data_ptr() if sigma2 is not None else None,
                    C.data_ptr(), noise_inv.data_ptr(), Y.data_ptr(), llr.data_ptr(), nSC, L, S, A, q, _prec(prec))


def c16_view(C):
    """(C16 int32 [nSC, L, A], c_exp int32 [nSC]) views of a complex64 coefficient buffer [nSC, L, A] used by a
    PE_PREC_F16_C16 cell."""
    nSC, L, A = C.shape
    w = C.view(dtype=__import__("torch").int32).reshape(-1)
    return w[: nSC * L * A].view(nSC, L, A), w[nSC * L * A: nSC * L * A + nSC]


def job_desc(H, sigma2, Y, n_pam_bits, prec="f16"):
    """pe_cell_desc of a serving-loop cell: inputs only (each job's C / noise_inv / LLR come from the arena)."""
    import torch
    q = _q(n_pam_bits)
    nSC, A, L = _need(H, "H", torch.complex64, (None, None, None)).shape
    _need(sigma2, "sigma2", torch.float32, (nSC,))
    S = _need(Y, "Y", torch.complex64, (nSC, None, A)).shape[1]
    _same_device(H, sigma2, Y)
    return CellDesc(H.data_ptr(), sigma2.data_ptr(), None, None, Y.data_ptr(), None, nSC, L, S, A, q, _prec(prec))


class SlotPlan:
    """Owner of a pe_slot_plan: one slot of many cells as three tile kernels."""

    def __init__(self, cells, unbiased=True, solve_records=False):
        arr = (CellDesc * len(cells))(*cells)
        h = ctypes.c_void_p()
        f = (MMSE_UNBIASED if unbiased else 0) | (SLOT_SOLVE_RECORDS if solve_records else 0)
        _check(lib().pe_slot_plan_create(arr, len(cells), f, ctypes.byref(h)))
        self._h = h
        self.n_cells = len(cells)

    def update(self, cells, stream=None):
        arr = (CellDesc * len(cells))(*cells)
        _check(lib().pe_slot_plan_update(self._h, arr, len(cells), _stream(stream)))

    def launch(self, stream=None):
        _check(lib().pe_slot_plan_launch(self._h, _stream(stream)))

    def graph_launch(self, stream=None):
        _check(lib().pe_slot_plan_graph_launch(self._h, _stream(stream)))

    def info(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().pe_slot_plan_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self._h:
            lib().pe_slot_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeConfig(ctypes.Structure):
    """pe_serve_config of include/pusch_eq.h."""
    _fields_ = [("cells", _P), ("n_cells", _I), ("n_jobs", _I), ("arrival_rate_hz", ctypes.c_double),
                ("deadline_s", ctypes.c_double), ("hbm_budget_bytes", ctypes.c_size_t), ("max_batch_jobs", _I),
                ("max_inflight", _I), ("shed_late", _I), ("use_graph", _I), ("seed", _U), ("job_llr", _P),
                ("job_cells", _P), ("job_H", _P), ("job_sigma2", _P), ("job_Y", _P)]


class ServeReport(ctypes.Structure):
    """pe_serve_report of include/pusch_eq.h."""
    _fields_ = [("n_jobs", _I), ("n_done", _I), ("n_missed", _I), ("n_shed", _I), ("n_batches", _I),
                ("max_inflight_seen", _I), ("wall_s", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("max_s", ctypes.c_double), ("mean_s", ctypes.c_double),
                ("miss_rate", ctypes.c_double), ("jobs_per_s", ctypes.c_double),
                ("subcarriers_per_s", ctypes.c_double), ("mean_batch_jobs", ctypes.c_double),
                ("offered_rate_hz", ctypes.c_double), ("iso_job_s", ctypes.c_double),
                ("contention_index", ctypes.c_double), ("model_s_per_sc", ctypes.c_double), ("n_late", _I),
                ("model_margin_s", ctypes.c_double)]


def serve_run(cells, n_jobs, rate_hz, deadline_s=0.5e-3, hbm_budget=1 << 30, max_batch=64, max_inflight=2,
              shed_late=False, use_graph=True, seed=0, job_llr=None, job_cells=None, job_H=None, job_sigma2=None,
              job_Y=None):
    """Runs the slot serving loop (pe_serve_run). cells: list of CellDesc with H, sigma2, Y_rx set.
    job_llr: optional list of n_jobs float32 GPU tensors receiving each job's LLRs (each at least the size of
    its cell's LLRs). job_cells: optional cell index of each job (else drawn from seed); with it, job_H /
    job_sigma2 / job_Y: optional lists of n_jobs GPU tensors, each job's own inputs in its cell's shapes.
    shed_late: False/0 (serve every job), True/1 (drop jobs already late at launch) or 2 (deadline-aware
    admission). Returns (report dict, per-job latencies in seconds (-1 = shed), per-job cell index)."""
    import numpy as np
    import torch
    arr = (CellDesc * len(cells))(*cells)
    keep = []

    def ptr_array(ts, name, dtype, shape_of):
        if ts is None:
            return None
        if len(ts) != n_jobs:
            raise ValueError("%s: %d tensors for %d jobs" % (name, len(ts), n_jobs))
        for j, t in enumerate(ts):
            _need(t, "%s[%d]" % (name, j), dtype, *shape_of(j))
        _same_device(*ts)
        a = (ctypes.c_void_p * n_jobs)(*[t.data_ptr() for t in ts])
        keep.append(a)
        return ctypes.cast(a, _P)

    jc = None
    if job_cells is not None:
        jc = np.ascontiguousarray(job_cells, dtype=np.int32)
        if jc.shape != (n_jobs,) or jc.min() < 0 or jc.max() >= len(cells):
            raise ValueError("job_cells: %d indices in [0, %d) expected" % (n_jobs, len(cells)))
    elif job_H is not None or job_sigma2 is not None or job_Y is not None:
        raise ValueError("per-job inputs need job_cells")

    def cshape(j, f):
        return (f(cells[int(jc[j])]), None) if jc is not None else (None, None)

    llr_len = lambda c: c.nSC * c.nLayer * c.nSym * 2 * c.nPamBits
    max_llr = max(llr_len(c) for c in cells)
    p_llr = ptr_array(job_llr, "job_llr", torch.float32,
                      lambda j: (None, llr_len(cells[int(jc[j])]) if jc is not None else max_llr))
    p_H = ptr_array(job_H, "job_H", torch.complex64, lambda j: cshape(j, lambda c: (c.nSC, c.nAnt, c.nLayer)))
    p_s2 = ptr_array(job_sigma2, "job_sigma2", torch.float32, lambda j: cshape(j, lambda c: (c.nSC,)))
    p_Y = ptr_array(job_Y, "job_Y", torch.complex64, lambda j: cshape(j, lambda c: (c.nSC, c.nSym, c.nAnt)))
    cfg = ServeConfig(ctypes.cast(arr, _P), len(cells), n_jobs, float(rate_hz), float(deadline_s), int(hbm_budget),
                      max_batch, max_inflight, int(shed_late), 1 if use_graph else 0, seed, p_llr,
                      jc.ctypes.data_as(_P) if jc is not None else None, p_H, p_s2, p_Y)
    rep = ServeReport()
    lat = np.zeros(n_jobs, dtype=np.float64)
    cell = np.zeros(n_jobs, dtype=np.int32)
    _check(lib().pe_serve_run(ctypes.byref(cfg), ctypes.byref(rep), lat.ctypes.data_as(_P), cell.ctypes.data_as(_P)))
    return {f[0]: getattr(rep, f[0]) for f in ServeReport._fields_}, lat, cell
