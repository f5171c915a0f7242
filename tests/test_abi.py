"""CPU tests of the C ABI boundary: the library loads, exports every symbol include/pusch_eq.h declares, and
rejects bad arguments before touching a device. No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pusch_eq.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pe_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for required in ("pe_coef_apply", "pe_gram_mmse", "pe_llr_demap_f16", "pe_llr_demap", "pe_coef_apply_demap",
                     "pe_coef_apply_demap_f16", "pe_srs_filter", "pe_coef_to_c16", "pe_coef_apply_demap_c16",
                     "pe_bfp_decompress", "pe_gram_mmse_c16", "pe_mmse_apply_demap", "pe_freq_interp",
                     "pe_gram", "pe_coef_apply_v1"):
        assert required in names


def test_library_exports_every_declared_symbol(pe):
    h = pe.lib()
    missing = [n for n in declared_functions() if not hasattr(h, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol(pe):
    assert set(declared_functions()) <= set(pe.SIGNATURES)


def test_version_and_status_strings(pe):
    assert "gfx950" in pe.version()
    h = pe.lib()
    assert h.pe_status_string(0) == b"ok"
    assert h.pe_status_string(1) == b"invalid argument"


def _call_coef(pe, L=8, S=11, A=64, nsc=4, prec=0, ptr=1):
    p = ctypes.c_void_p(ptr)
    return pe.lib().pe_coef_apply(p, p, p, nsc, L, S, A, prec, None)


@pytest.mark.parametrize("kw,status", [
    (dict(L=0), 1), (dict(L=9), 1), (dict(S=0), 1), (dict(S=17), 2), (dict(A=6), 1), (dict(A=0), 1),
    (dict(A=260), 2), (dict(A=132, prec=1), 2), (dict(prec=7), 1), (dict(nsc=-1), 1), (dict(ptr=0), 1),
])
def test_coef_apply_rejects_bad_arguments_without_a_device(pe, kw, status):
    assert _call_coef(pe, **kw) == status
    assert pe.lib().pe_last_error()  # message set


def test_empty_calls_succeed_without_launching(pe):
    assert _call_coef(pe, nsc=0, ptr=0) == 0
    h = pe.lib()
    n = ctypes.c_void_p(0)
    assert h.pe_llr_demap_f16(n, n, n, 0, 4, None) == 0
    assert h.pe_llr_demap(n, n, n, 0, 8, 11, 4, None) == 0
    assert h.pe_gram_mmse(n, n, n, n, n, 0, 8, 64, 1, 0.0, None) == 0
    assert h.pe_coef_apply_demap(n, n, n, n, 0, 8, 11, 64, 4, 1, None) == 0
    assert h.pe_coef_apply_demap_f16(n, n, n, n, 0, 8, 11, 64, 4, 1, None) == 0


def test_other_entry_points_validate(pe):
    h = pe.lib()
    p = ctypes.c_void_p(1)
    n = ctypes.c_void_p(0)
    assert h.pe_llr_demap_f16(p, p, p, 10, 5, None) == 1
    assert h.pe_llr_demap_f16(p, p, p, -1, 4, None) == 1
    assert h.pe_llr_demap(p, p, p, 4, 8, 11, 0, None) == 1
    assert h.pe_gram_mmse(p, p, p, p, n, 4, 8, 64, 1, 0.0, None) == 1   # flags buffer required
    assert h.pe_gram_mmse(p, p, p, p, p, 4, 9, 64, 1, 0.0, None) == 1
    assert h.pe_gram_mmse(p, p, p, p, p, 4, 8, 64, 4, 0.0, None) == 1   # unknown flag bit
    assert h.pe_gram(p, p, p, 4, 8, 0, None) == 1
    assert h.pe_coef_apply_demap(p, p, p, p, 4, 8, 11, 64, 5, 1, None) == 1
    assert h.pe_coef_apply_demap_f16(p, p, p, p, 4, 8, 11, 64, 5, 1, None) == 1
    assert h.pe_coef_apply_demap_f16(p, p, p, n, 4, 8, 11, 64, 4, 1, None) == 1
    assert h.pe_coef_apply_range(p, p, p, 4, -1, 8, 11, 64, 0, None) == 1
    # SRS filter: (H, SRS, W, FOCC, nAnt, nBlocks, K, nPorts, foccLen, flags, stream)
    assert h.pe_srs_filter(n, n, n, n, 64, 0, 24, 4, 8, 0, None) == 0          # empty
    assert h.pe_srs_filter(p, p, p, p, 64, -1, 24, 4, 8, 0, None) == 1
    assert h.pe_srs_filter(p, p, p, p, 64, 2, 49, 4, 8, 0, None) == 1         # K > 48
    assert h.pe_srs_filter(p, p, p, p, 64, 2, 24, 9, 8, 0, None) == 1         # nPorts > 8
    assert h.pe_srs_filter(p, p, p, p, 64, 2, 24, 4, 0, 0, None) == 1         # foccLen 0
    assert h.pe_srs_filter(p, p, p, p, 64, 2, 24, 4, 8, 4, None) == 1         # unknown flag
    assert h.pe_srs_filter(p, p, p, n, 64, 2, 24, 4, 8, 0, None) == 1         # null FOCC
    assert h.pe_hbm_probe(p, p, 17, 0, 64, None) == 1                         # bytes % 16
    assert h.pe_hbm_probe(p, p, 32, 0, 0, None) == 1                          # grid
    assert h.pe_hbm_probe(n, n, 0, 0, 64, None) == 0                          # empty
    assert h.pe_hbm_probe(p, ctypes.c_void_p(256), 32, 0, 64, None) == 1      # unaligned src
    a16 = ctypes.c_void_p(256)
    # fronthaul: (in, Y_rx, nPrb, nSym, nAnt, iqWidth, row_stride, scale, stream)
    assert h.pe_bfp_decompress(n, n, 0, 11, 64, 9, 0, 1.0, None) == 0
    assert h.pe_bfp_decompress(p, p, 4, 11, 64, 17, 4 * 52, 1.0, None) == 1   # iqWidth
    assert h.pe_bfp_decompress(p, p, 4, 11, 64, 9, 4 * 28 - 1, 1.0, None) == 1  # row too short
    assert h.pe_bfp_decompress(p, n, 4, 11, 64, 9, 4 * 28, 1.0, None) == 1     # null output
    # binary16 coefficients: (C16, c_exp, Y, noise_inv, llr, nSC, L, S, A, q, flags, stream)
    assert h.pe_coef_to_c16(n, n, n, 0, 8, 64, None) == 0
    assert h.pe_coef_to_c16(a16, a16, n, 4, 8, 64, None) == 1              # null c_exp
    assert h.pe_coef_to_c16(a16, p, a16, 4, 8, 64, None) == 1              # unaligned C16
    assert h.pe_coef_to_c16(a16, a16, a16, 4, 8, 256, None) == 2           # nAnt > 128
    assert h.pe_coef_apply_demap_c16(n, n, n, n, n, 0, 8, 11, 64, 4, 0, None) == 0
    assert h.pe_coef_apply_demap_c16(a16, a16, a16, a16, a16, 4, 8, 11, 64, 4, 2, None) == 1  # flag
    assert h.pe_coef_apply_demap_c16(a16, a16, a16, a16, a16, 4, 8, 11, 64, 5, 0, None) == 1  # q
    assert h.pe_coef_apply_demap_c16(a16, a16, a16, n, a16, 4, 8, 11, 64, 4, 0, None) == 1    # noise_inv
    assert h.pe_coef_apply_c16(a16, a16, a16, n, 4, 8, 11, 64, None) == 1                      # Y_eq
    assert h.pe_coef_apply_c16(a16, a16, a16, a16, 4, 8, 11, 62, None) == 1                    # nAnt % 4
    # one-pass MMSE + apply: (H, sigma2, Y, C, noise_inv, flags, llr, nSC, L, S, A, q, prec, flags, thr, stream)
    ma = h.pe_mmse_apply_demap
    assert ma(n, n, n, n, n, n, n, 0, 8, 11, 64, 4, 1, 1, 0.0, None) == 0                  # empty
    assert ma(a16, a16, a16, a16, a16, a16, a16, 4, 8, 11, 64, 4, 0, 1, 0.0, None) == 2    # FP32: no MFMA
    assert ma(a16, a16, a16, a16, a16, a16, a16, 4, 8, 11, 64, 4, 1, 8, 0.0, None) == 1    # unknown flag
    assert ma(a16, a16, a16, a16, a16, a16, a16, 4, 8, 11, 64, 5, 1, 1, 0.0, None) == 1    # q
    assert ma(a16, a16, a16, a16, a16, a16, a16, 4, 8, 11, 132, 4, 1, 1, 0.0, None) == 2   # nAnt > 128
    assert ma(a16, a16, a16, n, a16, a16, a16, 4, 8, 11, 64, 4, 1, 1, 0.0, None) == 1      # C workspace
    assert ma(a16, a16, a16, a16, a16, n, a16, 4, 8, 11, 64, 4, 1, 1, 0.0, None) == 1      # sc_flags
    assert ma(a16, a16, p, a16, a16, a16, a16, 4, 8, 11, 64, 4, 1, 1, 0.0, None) == 1      # unaligned Y
    assert h.pe_srs_filter(a16, a16, a16, a16, 64, 2, 10, 4, 8, 1, None) == 2  # MFMA needs K % 4 == 0
    assert h.pe_srs_filter(a16, p, a16, a16, 64, 2, 12, 4, 8, 1, None) == 1    # MFMA needs aligned SRS
    # frequency interpolation: (W, X, H_f, nBatch, nOut, nIn, nCol, prec, stream)
    fi = h.pe_freq_interp
    assert fi(n, n, n, 0, 64, 96, 4, 0, None) == 0                   # empty batch
    assert fi(a16, a16, a16, 4, 0, 96, 4, 0, None) == 1              # nOut
    assert fi(a16, a16, a16, 4, 64, 129, 4, 2, None) == 1            # nIn > 128
    assert fi(a16, a16, a16, 4, 64, 96, 17, 2, None) == 1            # nCol > 16
    assert fi(a16, a16, a16, 4, 64, 96, 4, 1, None) == 2             # f16: not offered
    assert fi(a16, a16, a16, 4, 256, 128, 4, 0, None) == 2           # FP32: W does not fit LDS
    assert fi(a16, a16, a16, 4, 64, 96, 4, 9, None) == 1             # unknown precision
    assert fi(a16, n, a16, 4, 64, 96, 4, 2, None) == 1               # null X
    assert fi(a16, p, a16, 4, 64, 96, 4, 2, None) == 1               # unaligned X


def test_slot_plan_rejects_bad_cells_without_a_device(pe):
    h = pe.lib()
    out = ctypes.c_void_p()
    one = ctypes.c_void_p(16)
    good = pe.CellDesc(None, None, one, one, one, one, 12, 8, 11, 64, 4, 1)
    bad_shape = pe.CellDesc(None, None, one, one, one, one, 12, 9, 11, 64, 4, 1)
    bad_prec = pe.CellDesc(None, None, one, one, one, one, 12, 8, 11, 64, 4, 0)
    no_llr = pe.CellDesc(None, None, one, one, one, None, 12, 8, 11, 64, 4, 1)
    for cell, status in ((bad_shape, 1), (bad_prec, 2), (no_llr, 1)):
        arr = (pe.CellDesc * 2)(good, cell)
        assert h.pe_slot_plan_create(arr, 2, 1, ctypes.byref(out)) == status
        assert b"cell 1" in h.pe_last_error()
    assert h.pe_slot_plan_create(None, 0, 1, ctypes.byref(out)) == 1
    assert h.pe_slot_plan_launch(None, None) == 1


def test_serve_rejects_bad_config_without_a_device(pe):
    h = pe.lib()
    one = ctypes.c_void_p(16)
    good = pe.CellDesc(one, one, None, None, one, None, 12, 8, 11, 64, 4, 1)
    no_h = pe.CellDesc(None, one, None, None, one, None, 12, 8, 11, 64, 4, 1)
    arr = (pe.CellDesc * 1)(good)
    rep = pe.ServeReport()

    def run(cells=arr, n=1, jobs=10, budget=1 << 30, batch=8, inflight=2, deadline=5e-4, shed=0):
        cfg = pe.ServeConfig(ctypes.cast(cells, ctypes.c_void_p) if cells is not None else None, n, jobs, 1000.0,
                             deadline, budget, batch, inflight, shed, 1, 1, None)
        return h.pe_serve_run(ctypes.byref(cfg), ctypes.byref(rep), None, None)

    assert h.pe_serve_run(None, ctypes.byref(rep), None, None) == 1
    assert run(cells=None, n=0) == 1
    assert run(jobs=0) == 1
    assert run(inflight=9) == 1
    assert run(deadline=0.0) == 1
    assert run(shed=3) == 1 and run(shed=-1) == 1
    assert run(budget=1000) == 5          # PE_ERR_CAPACITY: one job does not fit
    assert run(cells=(pe.CellDesc * 1)(no_h)) == 1


def test_tune_table_parser(pe, tmp_path):
    """pe_tune_load returns the number of coef, mmse and solve entries it read, skips comments and malformed
    lines, and NULL restores the built-in defaults; the committed table parses entirely."""
    h = pe.lib()
    t = tmp_path / "table.txt"
    t.write_text("# comment\n"
                 "coef 8 11 64 1 8 4 1\n"
                 "coef -1 -1 -1 -1 16 2 1\n"
                 "mmse 8 64 4096 16 8 2 1\n"
                 "mmse 8 64 -1 16\n"
                 "solve 8 64 -1 64 4\n"
                 "solve 8 64 -1 64\n"    # too few fields
                 "slot 32 4 0\n"
                 "coef 1 2 3\n"          # too few fields
                 "bogus line\n")
    assert h.pe_tune_load(str(t).encode()) == 5
    assert h.pe_tune_load(str(tmp_path / "missing.txt").encode()) == -1
    assert h.pe_tune_load(None) == 0
    committed = os.path.join(ROOT, "aerial-mimo-tc-optimization_amd", "tune_table.txt")
    lines = [ln for ln in open(committed) if ln.startswith(("coef ", "mmse ", "solve "))]
    assert h.pe_tune_load(committed.encode()) == len(lines)


def _misaligned_cases(pe):
    """(entry point name, call with every pointer 256-byte aligned except the one named, buffer name)."""
    h = pe.lib()
    A = ctypes.c_void_p(256)

    def args(n_ptr, bad, off):
        return [ctypes.c_void_p(256 + off) if i == bad else A for i in range(n_ptr)]

    cases = []
    # (C, Y_rx, Y_eq, nSC, L, S, A, prec): C and Y_rx are read 16 bytes at a time, Y_eq written 8
    for bad, off, name in ((0, 8, "C"), (1, 8, "Y_rx"), (2, 4, "Y_eq")):
        for prec in (0, 1):
            cases.append((lambda b=bad, o=off, p=prec: h.pe_coef_apply(*args(3, b, o), 4, 8, 11, 64, p, None), name))
    cases.append((lambda: h.pe_coef_apply_v1(*args(3, 0, 8), 4, 11, None), "C"))
    # fused apply + demap, float LLRs: 16-byte stores at q = 2, 4, 8-byte at q = 1, 3
    for q, off in ((4, 8), (2, 8), (3, 4), (1, 4)):
        cases.append((lambda q=q, o=off: h.pe_coef_apply_demap(*args(4, 3, o), 4, 8, 11, 64, q, 1, None), "llr"))
    cases.append((lambda: h.pe_coef_apply_demap(*args(4, 2, 2), 4, 8, 11, 64, 4, 1, None), "noise_inv"))
    # binary16 LLRs: 16 / 8 / 4 bytes at q = 4 / 2 / 1, 3
    for q, off in ((4, 8), (2, 4), (1, 2), (3, 2)):
        cases.append((lambda q=q, o=off: h.pe_coef_apply_demap_f16(*args(4, 3, o), 4, 8, 11, 64, q, 1, None),
                      "llr"))
        cases.append((lambda q=q, o=off: h.pe_llr_demap_f16(*args(3, 2, o), 10, q, None), "llrOut"))
    cases.append((lambda: h.pe_llr_demap_f16(*args(3, 0, 2), 10, 4, None), "softEst"))
    cases.append((lambda: h.pe_llr_demap(*args(3, 0, 4), 4, 8, 11, 4, None), "Y_eq"))
    cases.append((lambda: h.pe_llr_demap(*args(3, 2, 8), 4, 8, 11, 4, None), "llr"))
    # MMSE: H rows of an even layer count are read 16 bytes at a time
    cases.append((lambda: h.pe_gram_mmse(*args(5, 0, 8), 4, 8, 64, 1, 0.0, None), "H"))
    cases.append((lambda: h.pe_gram_mmse(*args(5, 2, 4), 4, 8, 64, 1, 0.0, None), "C"))
    cases.append((lambda: h.pe_gram_mmse_c16(*args(6, 0, 8), 4, 8, 64, 1, 0.0, None), "H"))
    cases.append((lambda: h.pe_gram(*args(3, 1, 4), 4, 8, 64, None), "H"))
    cases.append((lambda: h.pe_coef_apply_demap_c16(*args(5, 4, 8), 4, 8, 11, 64, 4, 0, None), "llr"))
    return cases


def test_misaligned_pointers_are_rejected_without_a_device(pe):
    """Every entry point checks that each buffer is aligned for the widest access its kernels make (16-byte loads
    of C, Y_rx and H rows, LLR stores of 2 q values) and returns PE_ERR_INVALID, naming the buffer, before
    anything is launched; the same calls with aligned buffers of an odd layer count pass the check for H."""
    h = pe.lib()
    for call, name in _misaligned_cases(pe):
        assert call() == 1
        msg = h.pe_last_error().decode()
        assert name in msg and "aligned" in msg, msg
    # an odd layer count reads H rows 8 bytes at a time: an 8-byte-aligned H passes the alignment check (the
    # call then fails on the null flags buffer, not on H)
    p8 = ctypes.c_void_p(264)
    A = ctypes.c_void_p(256)
    assert h.pe_gram_mmse(p8, A, A, A, None, 4, 7, 64, 1, 0.0, None) == 1
    assert b"null" in h.pe_last_error()


def test_slot_and_serve_check_alignment_without_a_device(pe):
    h = pe.lib()
    out = ctypes.c_void_p()
    a = ctypes.c_void_p(256)
    b8 = ctypes.c_void_p(264)
    good = pe.CellDesc(None, None, a, a, a, a, 12, 8, 11, 64, 4, 1)
    bad_y = pe.CellDesc(None, None, a, a, b8, a, 12, 8, 11, 64, 4, 1)
    bad_llr = pe.CellDesc(None, None, a, a, a, b8, 12, 8, 11, 64, 4, 1)
    for cell, name in ((bad_y, b"Y_rx"), (bad_llr, b"llr")):
        arr = (pe.CellDesc * 2)(good, cell)
        assert h.pe_slot_plan_create(arr, 2, 1, ctypes.byref(out)) == 1
        assert b"cell 1" in h.pe_last_error() and name in h.pe_last_error()
    # serving: per-job inputs need job_cells, and each must be aligned
    cells = (pe.CellDesc * 1)(pe.CellDesc(a, a, None, None, a, None, 12, 8, 11, 64, 4, 1))
    rep = pe.ServeReport()
    jc = (ctypes.c_int * 2)(0, 0)
    ys = (ctypes.c_void_p * 2)(256, 264)

    def run(job_cells, job_y):
        cfg = pe.ServeConfig(ctypes.cast(cells, ctypes.c_void_p), 1, 2, 0.0, 5e-4, 1 << 30, 8, 1, 0, 0, 1, None,
                             job_cells, None, None, job_y)
        return h.pe_serve_run(ctypes.byref(cfg), ctypes.byref(rep), None, None)

    assert run(None, ctypes.cast(ys, ctypes.c_void_p)) == 1
    assert b"job_cells" in h.pe_last_error()
    assert run(ctypes.cast(jc, ctypes.c_void_p), ctypes.cast(ys, ctypes.c_void_p)) == 1
    assert b"job 1" in h.pe_last_error() and b"job_Y" in h.pe_last_error()
    jc_bad = (ctypes.c_int * 2)(0, 3)
    assert run(ctypes.cast(jc_bad, ctypes.c_void_p), None) == 1
    assert b"out of range" in h.pe_last_error()


def test_binding_rejects_mismatched_tensors_before_any_launch(pe):
    """The binding checks dtype, contiguity and shape of every tensor before it calls the library (these run
    first, so CPU tensors show them here), then that the tensors are on a GPU: nothing reaches a kernel with a
    buffer smaller than the shapes imply or of the wrong element type."""
    import torch
    c64, f32 = torch.complex64, torch.float32
    C = torch.zeros((4, 8, 64), dtype=c64)
    Y = torch.zeros((4, 11, 64), dtype=c64)
    ni = torch.zeros((4, 8), dtype=f32)
    H = torch.zeros((4, 64, 8), dtype=c64)
    s2 = torch.zeros(4, dtype=f32)
    with pytest.raises(TypeError, match="C: dtype"):
        pe.coef_apply(C.to(torch.complex128), Y)
    with pytest.raises(ValueError, match="Y: shape"):
        pe.coef_apply(C, Y[:, :, :32].contiguous())
    with pytest.raises(ValueError, match="contiguous"):
        pe.coef_apply(C, Y.transpose(0, 1))
    with pytest.raises(ValueError, match="GPU tensors"):
        pe.coef_apply(C, Y)
    with pytest.raises(TypeError, match="sigma2: dtype"):
        pe.gram_mmse(H, s2.double())
    with pytest.raises(ValueError, match="H: shape|contiguous"):
        pe.gram_mmse(H[:, ::2], s2)
    with pytest.raises(ValueError, match="noise_inv: shape"):
        pe.coef_apply_demap(C, Y, ni[:, :4].contiguous(), 4)
    with pytest.raises(ValueError, match="out: shape"):
        pe.coef_apply_demap(C, Y, ni, 4, out=torch.zeros((4, 8, 11, 6), dtype=f32))
    with pytest.raises(ValueError, match="n_pam_bits"):
        pe.coef_apply_demap(C, Y, ni, 5)
    with pytest.raises(ValueError, match="Y: shape"):
        pe.mmse_apply_demap(H, s2, Y[:2].contiguous(), 4)
    with pytest.raises(TypeError, match="noise_inv: dtype"):
        pe.llr_demap(torch.zeros((4, 8, 11), dtype=c64), ni.half(), 4)
    with pytest.raises(ValueError, match="out: at least"):
        pe.coef_apply_v1(C, Y, torch.zeros(10, dtype=c64), 4, 11)
    with pytest.raises(ValueError, match="range"):
        pe.coef_apply_range(C, Y, torch.zeros((4, 8, 11), dtype=c64), 3, 2)
    with pytest.raises(ValueError, match="H: shape"):
        pe.cell_desc(C, ni, Y, torch.zeros((4, 8, 11, 8), dtype=f32), 4, H=H[:, :32].contiguous(), sigma2=s2)
    with pytest.raises(TypeError, match="W: elements"):
        pe.srs_filter(torch.zeros((2, 64, 24), dtype=torch.int32), torch.zeros((24, 24), dtype=torch.int16),
                      torch.zeros((4, 8), dtype=torch.int32))


def test_noalloc_entry_point_validates_without_a_device(pe):
    """pe_mmse_apply_demap_noalloc: work beyond the fixed scratch is PE_ERR_CAPACITY, PE_EQ_WRITE_C (no C to
    write) is an unknown flag, both before any device call."""
    h = pe.lib()
    a = ctypes.c_void_p(256)
    n = ctypes.c_void_p(0)
    na = h.pe_mmse_apply_demap_noalloc
    # (H, sigma2, Y, noise_inv, flags, llr, nSC, L, S, A, q, prec, flags, thr, stream)
    assert na(n, n, n, n, n, n, 0, 8, 11, 64, 4, 1, 1, 0.0, None) == 0                 # empty
    assert na(a, a, a, a, a, a, 4097, 8, 11, 128, 4, 1, 1, 0.0, None) == 5             # 33.6 MB > 32 MiB
    assert b"scratch" in h.pe_last_error()
    assert na(a, a, a, a, a, a, 16, 8, 11, 64, 4, 1, 1 | 2, 0.0, None) == 1            # PE_EQ_WRITE_C
    assert na(a, a, a, a, a, a, 16, 8, 11, 64, 4, 0, 1, 0.0, None) == 2                # FP32: no MFMA
    assert na(a, a, ctypes.c_void_p(264), a, a, a, 16, 8, 11, 64, 4, 1, 1, 0.0, None) == 1  # unaligned Y
    assert pe.NOALLOC_SCRATCH_BYTES == 32 << 20


def test_interp_mmse_entry_point_validates_without_a_device(pe):
    """pe_interp_mmse_apply_demap: shapes outside the fused chain and bad pointers are rejected before any device
    call; an empty batch succeeds without launching."""
    h = pe.lib()
    a = ctypes.c_void_p(256)
    n = ctypes.c_void_p(0)
    f = h.pe_interp_mmse_apply_demap
    # (W, X, sigma2, Y, H, C, noise_inv, flags, llr, nBlk, nIn, nSym, nAnt, q, interp_prec, prec, flags, thr, stream)
    ok = (a, a, a, a, a, a, a, a, a, 2, 96, 11, 64, 4, 2, 1, 1, 0.0, None)
    assert f(n, n, n, n, n, n, n, n, n, 0, 96, 11, 64, 4, 2, 1, 1, 0.0, None) == 0   # empty
    assert f(*ok[:9], 2, 129, 11, 64, 4, 2, 1, 1, 0.0, None) == 1                      # nIn > 128
    assert f(*ok[:9], 2, 96, 11, 16, 4, 2, 1, 1, 0.0, None) == 2                       # 16 antennas
    assert f(*ok[:9], 2, 96, 11, 66, 4, 2, 1, 1, 0.0, None) == 2                       # not a multiple of 4
    assert f(*ok[:9], 2, 96, 11, 64, 4, 0, 1, 1, 0.0, None) == 2                       # FP32 interpolation
    assert f(*ok[:9], 2, 96, 11, 64, 4, 2, 0, 1, 0.0, None) == 2                       # FP32 apply
    assert f(*ok[:9], 2, 96, 11, 64, 5, 2, 1, 1, 0.0, None) == 1                       # nPamBits
    assert f(*ok[:9], 2, 96, 11, 64, 4, 2, 1, 8, 0.0, None) == 1                       # unknown flag bit
    assert f(n, *ok[1:], ) == 1                                                       # null W
    assert f(a, a, a, a, ctypes.c_void_p(264), *ok[5:]) == 1                         # H not 16-byte aligned
