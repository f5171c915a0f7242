"""GPU parity tests: every entry point of the C ABI against the oracle on the same seeded inputs.

Bars: bit-exact where the kernel reproduces the reference operation sequence (FP32 antenna-serial coef apply
and Gram product, both demappers, the fused FP32 path); stated EVM / relative-error bounds for the MFMA
precisions and the MMSE solve (tolerances written in each test)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSC, L8, S11, A64 = 1584, 8, 11, 64


def dev_c64(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)


def dev_f32(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def host(t):
    return t.cpu().numpy()


# ------------------------------------------------------------------------------------------------------
# coefficient apply, FP32
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed,scale", [(42, None), (7, 1.4)])
def test_coef_f32_bitexact_on_reference_inputs(po, pe, dev, seed, scale):
    """1584 x 8 x 11 x 64 on the reference's own input streams: GPU FP32 == the scalar chain, bit for bit."""
    C, Y = po.reference_coef_inputs(seed, NSC, L8, S11, A64, scale)
    ref = po.coef_apply_v1_f32(C, Y)
    out = host(pe.coef_apply(dev_c64(C, dev), dev_c64(Y, dev), prec="f32"))
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    # and the stated bar against the uncontracted reading of the same loop
    ref_nofma = po.coef_apply_v1_f32(C, Y, contracted=False)
    assert po.max_rel_err(ref_nofma, out) <= 1e-5 or po.norm_rel_err(ref_nofma, out) <= 1e-6
    assert po.cosine(ref_nofma, out) >= 0.9999999


@pytest.mark.parametrize("L,S,A,nsc", [(1, 1, 4, 3), (2, 4, 8, 13), (3, 7, 12, 5), (4, 11, 16, 40),
                                       (8, 11, 64, 37), (8, 14, 128, 9), (5, 16, 256, 7), (8, 1, 32, 50),
                                       (7, 13, 20, 129)])
def test_coef_f32_bitexact_shapes(po, pe, dev, L, S, A, nsc):
    rng = np.random.default_rng(L * 1000 + S * 10 + A)
    C = po.cn(rng, (nsc, L, A))
    Y = po.cn(rng, (nsc, S, A))
    out = host(pe.coef_apply(dev_c64(C, dev), dev_c64(Y, dev), prec="f32"))
    assert np.array_equal(out, po.coef_apply_v1_f32(C, Y))


def test_coef_v1_signature_and_range(po, pe, dev):
    import torch
    C, Y = po.reference_coef_inputs(42, 132, L8, S11, A64, None)
    ref = po.coef_apply_v1_f32(C, Y)
    dC, dY = dev_c64(C, dev), dev_c64(Y, dev)
    out = torch.zeros((132, 8, 11), dtype=torch.complex64, device=dev)
    pe.coef_apply_v1(dC, dY, out, 132, 11)
    assert np.array_equal(host(out), ref)
    out2 = torch.zeros_like(out)
    pe.coef_apply_range(dC, dY, out2, 12, 60, prec="f32")
    o2 = host(out2)
    assert np.array_equal(o2[12:72], ref[12:72]) and np.all(o2[:12] == 0) and np.all(o2[72:] == 0)


# ------------------------------------------------------------------------------------------------------
# coefficient apply, MFMA precisions
# ------------------------------------------------------------------------------------------------------
# EVM bounds vs the FP64 product on CN(0,1) samples and MMSE coefficients (dB)
MFMA_EVM_DB = {"f16": -60.0, "bf16": -40.0, "f16_split": -100.0, "bf16_split": -90.0}


@pytest.mark.parametrize("prec", ["f16", "bf16", "f16_split", "bf16_split"])
@pytest.mark.parametrize("L,S,A,nsc", [(8, 11, 64, 1584), (4, 14, 32, 100), (2, 11, 16, 77), (8, 7, 128, 50),
                                       (3, 16, 8, 33), (8, 11, 48, 20)])
def test_coef_mfma_evm(po, pe, dev, prec, L, S, A, nsc):
    rng = np.random.default_rng(A + 7 * L + S)
    H = po.rayleigh_h(rng, nsc, A, L)
    C, _, _, _, _ = po.mmse_f64(H, np.full(nsc, 0.01, np.float32), unbiased=False)
    C = C.astype(np.complex64)
    Y = po.cn(rng, (nsc, S, A))
    ref = po.coef_apply_f64(C, Y)
    out = host(pe.coef_apply(dev_c64(C, dev), dev_c64(Y, dev), prec=prec))
    _, db = po.evm(ref, out)
    print("prec=%s L=%d S=%d A=%d EVM=%.1f dB" % (prec, L, S, A, db))
    assert db <= MFMA_EVM_DB[prec]


def test_coef_f16_matches_published_rigor_band(po, pe, dev):
    """srand(7) inputs (coef_rigor.cu:101-104): the reference measured EVM(V2 vs V1) = -70.3 dB for its
    binary16 WMMA path; the binary16 MFMA path rounds the operands the same way."""
    C, Y = po.reference_coef_inputs(7, NSC, L8, S11, A64, 1.4)
    v1 = po.coef_apply_v1_f32(C, Y)
    out = host(pe.coef_apply(dev_c64(C, dev), dev_c64(Y, dev), prec="f16"))
    _, db = po.evm(v1, out)
    print("rigor EVM f16 = %.2f dB" % db)
    published = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                             "reference_published.json")))["coef_rigor_evm_db"]
    assert abs(db - published) < 0.3


def test_coef_mfma_scaling_handles_wide_range(po, pe, dev):
    """Per-row power-of-two scaling: samples far outside binary16 range (1e6) and tiny coefficients (1e-6)
    keep the binary16 EVM."""
    rng = np.random.default_rng(9)
    C = po.cn(rng, (64, 8, 64)) * np.float32(1e-6)
    Y = po.cn(rng, (64, 11, 64)) * np.float32(1e6)
    ref = po.coef_apply_f64(C, Y)
    for prec in ("f16", "f16_split"):
        out = host(pe.coef_apply(dev_c64(C, dev), dev_c64(Y, dev), prec=prec))
        assert np.all(np.isfinite(out))
        assert po.evm(ref, out)[1] <= MFMA_EVM_DB[prec]


# ------------------------------------------------------------------------------------------------------
# demappers
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("q", [1, 2, 3, 4])
def test_llr_f16_bitexact_on_reference_inputs(po, pe, dev, q):
    """aerial_llr_bench_v2.cu inputs (srand(42), N = 139392): binary16 LLRs identical to the formula."""
    import torch
    soft, ni = po.llr_bench_inputs(42, 139392)
    ref = po.llr_demap_f16(soft, ni, q)
    t_soft = torch.from_numpy(soft.view(np.int16)).to(dev)
    out = host(pe.llr_demap_f16(t_soft, dev_f32(ni, dev), q)).view(np.uint16)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("N", [1, 255, 257, 4099])
def test_llr_f16_ragged(po, pe, dev, N):
    import torch
    soft, ni = po.llr_bench_inputs(3, N)
    t_soft = torch.from_numpy(soft.view(np.int16)).to(dev)
    for q in (1, 3):
        out = host(pe.llr_demap_f16(t_soft, dev_f32(ni, dev), q)).view(np.uint16)
        assert np.array_equal(out, po.llr_demap_f16(soft, ni, q))


@pytest.mark.parametrize("q", [1, 2, 3, 4])
@pytest.mark.parametrize("L,S,nsc", [(8, 11, 300), (3, 5, 17)])
def test_llr_f32_bitexact(po, pe, dev, q, L, S, nsc):
    rng = np.random.default_rng(q + L)
    Yeq = po.cn(rng, (nsc, L, S), 0.7)
    ni = rng.uniform(0.5, 400, (nsc, L)).astype(np.float32)
    out = host(pe.llr_demap(dev_c64(Yeq, dev), dev_f32(ni, dev), q))
    assert np.array_equal(out, po.llr_demap_f32(Yeq, ni, q))


# ------------------------------------------------------------------------------------------------------
# fused coef apply + demap
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["f32", "f16", "bf16", "f16_split", "bf16_split"])
@pytest.mark.parametrize("q", [1, 2, 4])
@pytest.mark.parametrize("L,S,A,nsc", [(8, 11, 64, 500), (3, 9, 16, 41)])
def test_fused_equals_apply_then_demap(po, pe, dev, prec, q, L, S, A, nsc):
    """The fused kernel's LLRs equal demap(coef_apply) bit for bit (same product, same demapper)."""
    rng = np.random.default_rng(11 * q + L)
    C = po.cn(rng, (nsc, L, A), 1.0 / A)
    Y = po.cn(rng, (nsc, S, A))
    ni = rng.uniform(1, 100, (nsc, L)).astype(np.float32)
    dC, dY, dni = dev_c64(C, dev), dev_c64(Y, dev), dev_f32(ni, dev)
    fused = host(pe.coef_apply_demap(dC, dY, dni, q, prec=prec))
    two = host(pe.llr_demap(pe.coef_apply(dC, dY, prec=prec), dni, q))
    assert np.array_equal(fused, two)
    if prec == "f32":
        ref, _ = po.coef_apply_demap_f32_omp(C, Y, ni, q, nthreads=1)
        assert np.array_equal(fused, ref)


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16", "f16_split", "bf16_split"])
@pytest.mark.parametrize("q", [1, 2, 3, 4])
@pytest.mark.parametrize("L,S,A,nsc,ni_hi", [(8, 11, 64, 300, 100.0), (3, 9, 16, 41, 1e6), (5, 3, 24, 7, 1e4)])
def test_fused_llr_wire_format_f16(po, pe, dev, prec, q, L, S, A, nsc, ni_hi):
    """Binary16 LLR output = round-to-nearest-even binary16 of the FP32 fused LLRs, bit for bit; the large
    noise_inv case drives LLRs past the binary16 range, where the +-65504 clamp must give finite values."""
    import torch
    rng = np.random.default_rng(5 * q + L)
    C = po.cn(rng, (nsc, L, A), 1.0 / A)
    Y = po.cn(rng, (nsc, S, A))
    ni = rng.uniform(1, ni_hi, (nsc, L)).astype(np.float32)
    dC, dY, dni = dev_c64(C, dev), dev_c64(Y, dev), dev_f32(ni, dev)
    f32 = host(pe.coef_apply_demap(dC, dY, dni, q, prec=prec))
    out = pe.coef_apply_demap_f16(dC, dY, dni, q, prec=prec)
    assert out.dtype == torch.float16 and tuple(out.shape) == (nsc, L, S, 2 * q)
    bits = host(out.view(torch.int16)).view(np.uint16)
    assert np.array_equal(bits, f32.astype(np.float16).view(np.uint16))
    assert np.isfinite(host(out.float())).all()
    if ni_hi >= 1e6:
        assert (np.abs(f32) == 65504.0).any()  # the clamp was exercised
    if prec == "f32":  # against the oracle directly, rounded by numpy (IEEE round-to-nearest-even)
        ref, _ = po.coef_apply_demap_f32_omp(C, Y, ni, q, nthreads=1)
        assert np.array_equal(bits, ref.astype(np.float16).view(np.uint16))


# ------------------------------------------------------------------------------------------------------
# Gram product and MMSE
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("L,A,nsc", [(8, 64, 1584), (4, 32, 55), (1, 4, 9), (6, 100, 21), (3, 200, 37),
                                     (8, 1024, 5), (2, 65, 70), (7, 1, 3), (5, 128, 13)])
def test_gram_bitexact(po, pe, dev, L, A, nsc):
    rng = np.random.default_rng(L + A)
    M = po.cn(rng, (nsc, L, A))
    H = po.cn(rng, (nsc, A, L))
    out = host(pe.gram(dev_c64(M, dev), dev_c64(H, dev)))
    assert np.array_equal(out, po.gram_f32(M, H))


@pytest.mark.parametrize("unbiased", [False, True])
@pytest.mark.parametrize("L,A,nsc,snr", [(8, 64, 1584, 20.0), (4, 32, 300, 10.0), (2, 16, 100, 30.0),
                                         (1, 8, 33, 0.0), (8, 128, 64, 25.0), (5, 20, 40, 15.0)])
def test_mmse_vs_fp64(po, pe, dev, unbiased, L, A, nsc, snr):
    """FP32 solve on well-conditioned Rayleigh channels: per-subcarrier norm-wise relative error of C within
    1e-4 and noise_inv within 1e-3 of the FP64 oracle; no subcarrier needs the FP64 solve."""
    rng = np.random.default_rng(int(snr) + L * 3 + A)
    H = po.rayleigh_h(rng, nsc, A, L)
    s2 = np.full(nsc, 10 ** (-snr / 10), np.float32)
    Cr, nir, _, cond, bad = po.mmse_f64(H, s2, unbiased=unbiased)
    assert bad == 0
    C, ni, fl = pe.gram_mmse(dev_c64(H, dev), dev_f32(s2, dev), unbiased=unbiased)
    C, ni, fl = host(C), host(ni), host(fl)
    well = cond <= 1e4
    assert np.all(fl[well] == 0) and np.all(fl[~well] == pe.SCFLAG_FP64)
    err = np.max(np.abs(C - Cr), axis=(1, 2)) / np.max(np.abs(Cr), axis=(1, 2))
    print("L=%d A=%d max cond %.1f max C rel err %.2e" % (L, A, cond.max(), err.max()))
    assert np.all(err <= 1e-4)
    assert np.max(np.abs(ni - nir) / np.abs(nir)) <= 1e-3


def test_mmse_ill_conditioned_goes_fp64(po, pe, dev):
    rng = np.random.default_rng(12)
    nsc, A, L = 64, 32, 4
    H = po.rayleigh_h(rng, nsc, A, L)
    # every other subcarrier: layer 3 nearly collinear with layer 2 -> kappa far above 1e4
    H[::2, :, 3] = H[::2, :, 2] + 1e-4 * H[::2, :, 3]
    s2 = np.full(nsc, 1e-9, np.float32)
    Cr, nir, _, cond, bad = po.mmse_f64(H, s2, unbiased=False)
    assert bad == 0 and np.all(cond[::2] > 1e4) and np.all(cond[1::2] < 1e4)
    C, ni, fl = pe.gram_mmse(dev_c64(H, dev), dev_f32(s2, dev), unbiased=False)
    C, fl = host(C), host(fl)
    assert np.all(fl[::2] == pe.SCFLAG_FP64) and np.all(fl[1::2] == 0)
    err = np.max(np.abs(C - Cr), axis=(1, 2)) / np.max(np.abs(Cr), axis=(1, 2))
    print("ill-conditioned: max rel err fp64 SCs %.2e, fp32 SCs %.2e" % (err[::2].max(), err[1::2].max()))
    assert np.all(err[::2] <= 1e-5)  # FP64 solve, rounded to complex64
    assert np.all(err[1::2] <= 1e-4)


def test_mmse_singular_is_flagged(po, pe, dev):
    rng = np.random.default_rng(13)
    H = po.rayleigh_h(rng, 16, 16, 4)
    H[:, :, 2] = 0  # dead layer, sigma2 = 0: Gram not positive definite
    C, ni, fl = pe.gram_mmse(dev_c64(H, dev), dev_f32(np.zeros(16, np.float32), dev), unbiased=True)
    assert np.all(host(fl) == pe.SCFLAG_FP64 | pe.SCFLAG_SINGULAR)
    assert np.all(host(C) == 0) and np.all(host(ni) == 0)


# ------------------------------------------------------------------------------------------------------
# HBM probe (the measured-ceiling yardstick)
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("nbytes,grid", [(16, 1), (4096 + 48, 3), (1 << 22, 1024)])
def test_hbm_probe_copy_is_a_copy(pe, dev, nbytes, grid):
    import torch
    a = torch.randint(-2 ** 31, 2 ** 31 - 1, (nbytes // 4,), dtype=torch.int32, device=dev)
    b = torch.zeros_like(a)
    pe.hbm_probe(a, b, grid=grid)
    sink = torch.zeros(grid, dtype=torch.int32, device=dev)
    pe.hbm_probe(a, sink, read_only=True, grid=grid)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


# ------------------------------------------------------------------------------------------------------
# coefficients kept as binary16 pairs + per-subcarrier exponent
# ------------------------------------------------------------------------------------------------------
def c16_reference(C):
    """numpy restatement of the format: e = 0 when max(|Re|, |Im|) of the subcarrier lies in [2^-9, 2^14),
    else frexp's exponent of it; C16 = binary16(C * 2^-e) rounded to nearest even."""
    re, im = C.real.astype(np.float32), C.imag.astype(np.float32)
    m = np.maximum(np.abs(re), np.abs(im)).reshape(C.shape[0], -1).max(1)
    e = np.where((m >= 2.0 ** -9) & (m < 2.0 ** 14), 0, np.frexp(m)[1]).astype(np.int32)
    s = np.ldexp(np.float32(1), -e).astype(np.float32)[:, None, None]
    pairs = np.stack([re * s, im * s], -1).astype(np.float16).view(np.uint16)
    return pairs, e


@pytest.mark.parametrize("L,A,nsc", [(8, 64, 300), (3, 16, 41), (4, 128, 17), (2, 24, 9)])
def test_coef_to_c16_format(po, pe, dev, L, A, nsc):
    rng = np.random.default_rng(L + A)
    C = po.cn(rng, (nsc, L, A), 1.0 / A)
    C[1] *= np.float32(1e-6)
    C[2] *= np.float32(3e5)
    C[3] = 0
    C16, e = pe.coef_to_c16(dev_c64(C, dev))
    pairs, e_ref = c16_reference(C)
    assert np.array_equal(host(e), e_ref)
    assert np.array_equal(host(C16).view(np.uint16).reshape(pairs.shape), pairs)


@pytest.mark.parametrize("q", [1, 2, 4])
@pytest.mark.parametrize("L,S,A,nsc", [(8, 11, 64, 500), (3, 9, 16, 41), (4, 14, 128, 33), (2, 5, 24, 19)])
def test_c16_apply_equals_f16_path_in_range(po, pe, dev, q, L, S, A, nsc):
    """In-range samples and coefficients: the binary16-coefficient path gives the same bits as the binary16
    MFMA path on the complex64 coefficients (Y_eq, FP32 LLRs, binary16 LLRs)."""
    import torch
    rng = np.random.default_rng(3 * q + L)
    H = po.rayleigh_h(rng, nsc, A, L)
    C, _, _, _, _ = po.mmse_f64(H, np.full(nsc, 0.01, np.float32), unbiased=True)
    C = C.astype(np.complex64)
    Y = po.cn(rng, (nsc, S, A))
    ni = rng.uniform(1, 100, (nsc, L)).astype(np.float32)
    dC, dY, dni = dev_c64(C, dev), dev_c64(Y, dev), dev_f32(ni, dev)
    C16, e = pe.coef_to_c16(dC)
    assert int(host(e).max()) == 0  # in range: no coefficient scaling
    assert np.array_equal(host(pe.coef_apply_c16(C16, e, dY)), host(pe.coef_apply(dC, dY, prec="f16")))
    ref = host(pe.coef_apply_demap(dC, dY, dni, q, prec="f16"))
    assert np.array_equal(host(pe.coef_apply_demap_c16(C16, e, dY, dni, q)), ref)
    o16 = pe.coef_apply_demap_c16(C16, e, dY, dni, q, llr_f16=True)
    assert np.array_equal(host(o16.view(torch.int16)).view(np.uint16), ref.astype(np.float16).view(np.uint16))


def test_c16_apply_wide_range_keeps_accuracy(po, pe, dev):
    """Coefficients and samples far outside the binary16 range: scaled blocks keep the binary16 EVM."""
    rng = np.random.default_rng(19)
    C = po.cn(rng, (64, 8, 64)) * np.float32(1e-6)
    C[::3] *= np.float32(1e10)
    Y = po.cn(rng, (64, 11, 64)) * np.float32(1e6)
    ref = po.coef_apply_f64(C, Y)
    C16, e = pe.coef_to_c16(dev_c64(C, dev))
    out = host(pe.coef_apply_c16(C16, e, dev_c64(Y, dev)))
    assert np.all(np.isfinite(out))
    assert po.evm(ref, out)[1] <= MFMA_EVM_DB["f16"]


def _mmse_c16_case(po, kind, rng):
    if kind == "ill":  # every other subcarrier goes through the FP64 solve
        H = po.rayleigh_h(rng, 64, 32, 4)
        H[::2, :, 3] = H[::2, :, 2] + 1e-4 * H[::2, :, 3]
        return H, np.full(64, 1e-9, np.float32)
    if kind == "singular":
        H = po.rayleigh_h(rng, 16, 16, 4)
        H[:, :, 2] = 0
        return H, np.zeros(16, np.float32)
    nsc, A, L = {"l8a64": (1584, 64, 8), "l2a128": (77, 128, 2), "l5a24": (45, 24, 5), "l8a128": (40, 128, 8)}[kind]
    H = po.rayleigh_h(rng, nsc, A, L)
    s2 = rng.uniform(1e-3, 0.3, nsc).astype(np.float32)
    H[3] *= np.float32(1e5)  # a subcarrier whose coefficients (~1e-5) need the block scaling
    return H, s2


@pytest.mark.parametrize("kind", ["l8a64", "l2a128", "l5a24", "l8a128", "ill", "singular"])
def test_gram_mmse_c16_equals_conversion_of_c64(po, pe, dev, kind):
    """pe_gram_mmse_c16 writes exactly what pe_coef_to_c16 makes of pe_gram_mmse's complex64 output (both
    FP32 and FP64 solves, singular subcarriers included); noise_inv and flags unchanged."""
    rng = np.random.default_rng(len(kind))
    H, s2 = _mmse_c16_case(po, kind, rng)
    dH, ds2 = dev_c64(H, dev), dev_f32(s2, dev)
    C, ni, fl = pe.gram_mmse(dH, ds2, unbiased=True)
    C16r, er = pe.coef_to_c16(C)
    C16, e, ni16, fl16 = pe.gram_mmse_c16(dH, ds2, unbiased=True)
    assert np.array_equal(host(fl16), host(fl)) and np.array_equal(host(ni16), host(ni))
    assert np.array_equal(host(e), host(er))
    assert np.array_equal(host(C16), host(C16r))
    if kind == "ill":
        assert (host(fl16) == pe.SCFLAG_FP64).sum() == 32
    if kind.startswith("l"):
        assert host(e)[3] != 0


@pytest.mark.parametrize("snr", [10.0, 30.0])
def test_c16_chain_accuracy(po, pe, dev, snr):
    """MMSE with binary16 coefficients, then the fused 256-QAM demap from them: EVM vs the FP64 chain within
    the binary16 bar and hard decisions equal except near zero."""
    rng = np.random.default_rng(int(snr))
    inp = po.chain_inputs(rng, 1584, 64, 8, 11, 4, snr)
    dH, ds2, dY = dev_c64(inp["H"], dev), dev_f32(inp["sigma2"], dev), dev_c64(inp["Y"], dev)
    C16, e, ni, fl = pe.gram_mmse_c16(dH, ds2, unbiased=True)
    yeq = host(pe.coef_apply_c16(C16, e, dY))
    Cd, nid, _, _, _ = po.mmse_f64(inp["H"], inp["sigma2"], unbiased=True)
    yref = po.coef_apply_f64(Cd, inp["Y"])
    assert po.evm(yref, yeq)[1] <= -60.0
    llr = host(pe.coef_apply_demap_c16(C16, e, dY, ni, 4))
    ref = po.llr_demap_f64(yref, nid, 4)
    eps = 1e-2 * np.median(np.abs(ref))
    sure = np.abs(ref) > eps
    assert np.array_equal(np.sign(llr[sure]), np.sign(ref[sure]))


# ------------------------------------------------------------------------------------------------------
# fronthaul block floating point -> Y_rx
# ------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("w,nPrb,nSym,nAnt,pad", [(9, 132, 11, 64, 0), (14, 25, 3, 16, 7), (16, 7, 2, 130, 1),
                                                  (1, 3, 1, 5, 0), (8, 273, 2, 4, 3), (12, 1, 14, 65, 0)])
def test_bfp_decompress_bitexact(po, pe, dev, w, nPrb, nSym, nAnt, pad):
    """Arbitrary bytes are valid blocks: the GPU decoder equals the oracle's bit for bit (scale 2^-15 and an
    arbitrary scale)."""
    import torch
    rng = np.random.default_rng(w * 100 + nPrb)
    stride = nPrb * po.bfp_block_bytes(w) + pad
    packed = rng.integers(0, 256, (nSym * nAnt, stride), dtype=np.uint8)
    dp = torch.from_numpy(packed).to(dev)
    for scale in (2.0 ** -15, 0.37):
        ref = po.bfp_decompress(packed, nPrb, nSym, nAnt, w, scale)
        out = host(pe.bfp_decompress(dp, nPrb, nSym, nAnt, w, scale))
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_bfp_to_fused_apply(po, pe, dev):
    """Compressed fronthaul samples decoded straight into the fused apply + demap equal the same samples
    given as complex64."""
    import torch
    rng = np.random.default_rng(5)
    nPrb, S, A, L, q = 20, 11, 32, 4, 4
    x = rng.integers(-20000, 20000, (S, A, nPrb * 12, 2)).astype(np.int32)
    packed = po.bfp_compress(x, 9)
    Y_ref = po.bfp_decompress(packed, nPrb, S, A, 9, 2.0 ** -15)
    Y = pe.bfp_decompress(torch.from_numpy(packed).to(dev), nPrb, S, A, 9, 2.0 ** -15)
    H = po.rayleigh_h(rng, nPrb * 12, A, L)
    C, ni, _ = pe.gram_mmse(dev_c64(H, dev), dev_f32(np.full(nPrb * 12, 0.01, np.float32), dev))
    a = host(pe.coef_apply_demap(C, Y, ni, q))
    b = host(pe.coef_apply_demap(C, dev_c64(Y_ref, dev), ni, q))
    assert np.array_equal(a, b)
