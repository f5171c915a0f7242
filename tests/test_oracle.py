"""CPU tests that pin the oracle: against the reference's published outputs, against closed forms and
independent numpy/scipy routines, and against brute force on small inputs. No GPU needed."""
import numpy as np
import pytest

# Published outputs of the reference (A100 runs of the reference programs), kept as a fixture written by
# tests/golden/make_reference_published.py from results/coef_rigor_results.txt:5 (EVM(V2 vs V1) on srand(7),
# U(+-0.7), 1584x8x11x64) and README.md:153-154 (cosine and max abs error on srand(42), U(+-0.5), same shape).
import json
import os

PUBLISHED = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                        "reference_published.json")))
RIGOR_EVM_DB = PUBLISHED["coef_rigor_evm_db"]
README_COS = PUBLISHED["readme_cosine"]
README_MAXABS = PUBLISHED["readme_max_abs"]
NSC, L, S, A = 1584, 8, 11, 64


def f16_round(po, x):
    return po.f16_bits_to_f32(po.f32_to_f16_bits(x))


def fp16_product(po, C, Y):
    """The reference V2 arithmetic up to accumulation order: operands rounded to binary16 (round to nearest
    even, __float2half), products summed at high precision."""
    Ch = (f16_round(po, C.real) + 1j * f16_round(po, C.imag)).astype(np.complex128)
    Yh = (f16_round(po, Y.real) + 1j * f16_round(po, Y.imag)).astype(np.complex128)
    return np.einsum("nla,nsa->nls", Ch, Yh)


def test_llr_lut_is_closed_form(po):
    lut = po.llr_lut()
    closed = np.array([1 / np.sqrt(2), 2 / np.sqrt(10), 1 / np.sqrt(10), 4 / np.sqrt(42), 2 / np.sqrt(42),
                       1 / np.sqrt(42), 8 / np.sqrt(170), 4 / np.sqrt(170), 2 / np.sqrt(170), 1 / np.sqrt(170)])
    assert np.max(np.abs(lut - closed)) < 1e-14
    assert np.array_equal(lut.astype(np.float32), closed.astype(np.float32))


def test_f16_rounding_matches_numpy(po):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10 ** rng.uniform(-9, 5, 20000),
                        np.array([65504, 65519.9, 65520, -65520, 6.1e-5, 5.96e-8, 2.98e-8, 0.0, -0.0],
                                 dtype=np.float32)]).astype(np.float32)
    ours = po.f32_to_f16_bits(x)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    assert np.array_equal(po.f16_bits_to_f32(ref), ref.view(np.float16).astype(np.float32))


def test_glibc_stream_reproduces_published_rigor_evm(po):
    """srand(7), rf() = ((float)rand()/RAND_MAX - 0.5f)*1.4f, all of C then all of Y (coef_rigor.cu:101-104),
    V1 FP32 chain vs binary16 operands -> the published -70.3 dB."""
    C, Y = po.reference_coef_inputs(7, NSC, L, S, A, scale=1.4)
    assert np.abs(C).max() <= 0.7 * np.sqrt(2) + 1e-6
    v1 = po.coef_apply_v1_f32(C, Y)
    v2 = fp16_product(po, C, Y)
    r, db = po.evm(v1, v2)
    assert abs(db - RIGOR_EVM_DB) < 0.1, db
    assert abs(r * 100 - 0.0307) < 0.0005


def test_glibc_stream_reproduces_published_cos_maxabs(po):
    """srand(42), U(+-0.5) (aerial_coef_wmma.cu:383-387) -> cos 0.99999997, max abs 0.001289."""
    C, Y = po.reference_coef_inputs(42, NSC, L, S, A, scale=None)
    v1 = po.coef_apply_v1_f32(C, Y)
    v2 = fp16_product(po, C, Y)
    assert round(po.cosine(v1, v2), 8) == pytest.approx(README_COS, abs=1.1e-8)
    assert po.max_abs_err(v1, v2) == pytest.approx(README_MAXABS, rel=2e-3)


def test_a_wrong_generator_misses_the_published_numbers(po):
    """Negative control: a mistake in the stream (Y before C) or in the product (conjugated C) moves the
    metrics far away from the published ones, so the two tests above can tell them apart."""
    C, Y = po.reference_coef_inputs(42, NSC, L, S, A, scale=None)
    v1 = po.coef_apply_v1_f32(C, Y)
    v1_conj = po.coef_apply_v1_f32(np.conj(C), Y)
    assert po.cosine(v1, v1_conj) < 0.1
    Cs = Y[:, :L, :].copy()  # a shifted stream gives different values
    assert po.max_abs_err(po.coef_apply_v1_f32(Cs, Y), v1) > 0.1


def test_coef_apply_forms_agree_and_match_numpy(po):
    rng = np.random.default_rng(1)
    C = po.cn(rng, (24, 8, 64))
    Y = po.cn(rng, (24, 11, 64))
    ref = np.einsum("nla,nsa->nls", C.astype(np.complex128), Y.astype(np.complex128))
    f64 = po.coef_apply_f64(C, Y)
    assert np.max(np.abs(f64 - ref)) < 1e-12
    for contracted in (True, False):
        f32 = po.coef_apply_v1_f32(C, Y, contracted=contracted)
        assert po.norm_rel_err(f64, f32) < 2e-6
    # the two FP32 readings differ only in rounding
    a = po.coef_apply_v1_f32(C, Y, True)
    b = po.coef_apply_v1_f32(C, Y, False)
    assert not np.array_equal(a, b)


def test_coef_apply_invariants(po):
    rng = np.random.default_rng(2)
    Y = po.cn(rng, (12, 11, 64))
    # C = [I_8 | 0] picks antennas 0..7 of every symbol, transposed, exactly
    C = np.zeros((12, 8, 64), dtype=np.complex64)
    for l in range(8):
        C[:, l, l] = 1.0
    out = po.coef_apply_v1_f32(C, Y)
    assert np.array_equal(out, np.transpose(Y[:, :, :8], (0, 2, 1)))
    assert np.array_equal(po.coef_apply_v1_f32(np.zeros_like(C), Y), np.zeros((12, 8, 11), np.complex64))
    # linearity in C (FP64 form)
    C1, C2 = po.cn(rng, (12, 8, 64)), po.cn(rng, (12, 8, 64))
    lhs = po.coef_apply_f64(C1.astype(np.complex128) * 2 + C2.astype(np.complex128), Y)
    rhs = 2 * po.coef_apply_f64(C1, Y) + po.coef_apply_f64(C2, Y)
    assert np.max(np.abs(lhs - rhs)) < 1e-10


def test_coef_apply_omp_equals_serial(po):
    rng = np.random.default_rng(5)
    C = po.cn(rng, (100, 8, 64))
    Y = po.cn(rng, (100, 11, 64))
    ser = po.coef_apply_v1_f32(C, Y)
    par, used = po.coef_apply_v1_f32_omp(C, Y, nthreads=4)
    assert used == 4 and np.array_equal(ser, par)


def test_gram_matches_matmul(po):
    rng = np.random.default_rng(4)
    M = po.cn(rng, (30, 8, 64))
    H = po.cn(rng, (30, 64, 8))
    ref = np.matmul(M.astype(np.complex128), H.astype(np.complex128))
    g64 = po.gram_f64(M, H)
    assert np.max(np.abs(g64 - ref)) < 1e-12
    g32 = po.gram_f32(M, H)
    assert po.norm_rel_err(ref, g32) < 2e-6


def _mmse_numpy(H, s2):
    Hh = np.conj(np.transpose(H, (0, 2, 1))).astype(np.complex128)
    G = Hh @ H.astype(np.complex128) + s2[:, None, None] * np.eye(H.shape[2])
    return np.linalg.solve(G, Hh), G


def test_mmse_matches_solve_and_identities(po):
    rng = np.random.default_rng(6)
    nsc, a, l = 64, 64, 8
    H = po.rayleigh_h(rng, nsc, a, l)
    s2 = np.full(nsc, 0.01, dtype=np.float32)
    C, ni, mu, cond, bad = po.mmse_f64(H, s2, unbiased=False)
    assert bad == 0
    W, G = _mmse_numpy(H, s2.astype(np.float64))
    assert np.max(np.abs(C - W)) < 1e-10 * np.max(np.abs(W))
    # push-through identity: (H^H H + s2 I)^-1 H^H = H^H (H H^H + s2 I)^-1
    Hd = H.astype(np.complex128)
    Hh = np.conj(np.transpose(Hd, (0, 2, 1)))
    s2v = float(s2[0])  # the float32 value the oracle used
    W2 = Hh @ np.linalg.inv(Hd @ Hh + s2v * np.eye(a))
    assert np.max(np.abs(C - W2)) < 1e-9
    # C H = I - s2 G^-1, mu = diag(C H)
    Gi = np.linalg.inv(G)
    CH = C @ Hd
    assert np.max(np.abs(CH - (np.eye(l) - s2v * Gi))) < 1e-12
    assert np.max(np.abs(mu - np.real(np.diagonal(CH, axis1=1, axis2=2)))) < 1e-10
    assert np.max(np.abs(ni - mu / (1 - mu))) < 1e-6 * np.max(ni)
    assert np.allclose(cond, np.linalg.cond(G, 1), rtol=1e-8)
    Cu, niu, _, _, _ = po.mmse_f64(H, s2, unbiased=True)
    assert np.allclose(Cu, C / mu[:, :, None], rtol=1e-12, atol=1e-14)
    assert np.array_equal(niu, ni)


def test_mmse_zero_forcing_limit_and_singular(po):
    rng = np.random.default_rng(7)
    H = po.rayleigh_h(rng, 16, 32, 4)
    C, _, _, _, bad = po.mmse_f64(H, np.full(16, 1e-12, np.float32), unbiased=False)
    assert bad == 0
    assert np.max(np.abs(C @ H.astype(np.complex128) - np.eye(4))) < 1e-8
    Hs = H.copy()
    Hs[:, :, 3] = 0  # a dead layer with sigma2 = 0: the last pivot is exactly 0, not positive definite
    C, ni, _, cond, bad = po.mmse_f64(Hs, np.zeros(16, np.float32), unbiased=False)
    assert bad == 16 and np.all(C == 0) and np.all(ni == 0) and np.all(np.isinf(cond))


@pytest.mark.parametrize("q", [1, 2, 3, 4])
def test_llr_folding_is_exact_maxlog(po, q):
    """The folding chain of computePamLlr equals brute-force max-log over the 2^q Gray-mapped levels."""
    xs = np.linspace(-1.6, 1.6, 2001)
    noise = 3.7
    for x in xs:
        out = np.empty(2 * q, np.float32)
        import ctypes
        po.lib().po_llr_ref_f64(float(np.float32(x)), float(np.float32(-x)), noise, q,
                                out.ctypes.data_as(ctypes.c_void_p))
        bf_i = po.llr_bruteforce(float(np.float32(x)), noise, q)
        bf_q = po.llr_bruteforce(float(np.float32(-x)), noise, q)
        bf = np.empty(2 * q)
        bf[0::2], bf[1::2] = bf_i, bf_q
        bf = np.clip(bf, -65504, 65504)
        assert np.max(np.abs(out - bf)) <= 1e-5 * max(1.0, np.max(np.abs(bf)))


def _gray_levels(q):
    """TS 38.211 PAM levels of one axis: bits c [2^q, q] (bit i of the code) and amplitudes (unit QAM power)."""
    codes = np.arange(1 << q)
    c = ((codes[:, None] >> np.arange(q)[None, :]) & 1).astype(np.float64)
    t = 1.0 - 2.0 * c[:, q - 1]
    for i in range(q - 2, -1, -1):
        t = (1.0 - 2.0 * c[:, i]) * (2.0 ** (q - 1 - i) - t)
    return c, t / np.sqrt(2.0 * (4.0 ** q - 1.0) / 3.0)


def _maxlog_numpy(x, ni, q, levels=None):
    """Exact max-log per axis, written independently of the oracle's C code: LLR_i = (min over levels with bit
    i = 1 of (x - a)^2 - min over levels with bit i = 0) * noise_inv. x [...], ni broadcastable -> [..., q]."""
    c, lev = _gray_levels(q) if levels is None else levels
    d2 = (x[..., None] - lev) ** 2
    out = [(np.where(c[:, i] == 1, d2, np.inf).min(-1) - np.where(c[:, i] == 0, d2, np.inf).min(-1)) * ni
           for i in range(q)]
    return np.stack(out, -1)


@pytest.mark.parametrize("q", [1, 2, 3, 4])
def test_llr_demap_f64_magnitudes_equal_bruteforce(po, q):
    """po_llr_demap_f64 (the FP64 end of every chain check, and the source of its hard-decision epsilon) over
    the equalizer layout [sc][l][s][2q]: every LLR equals the exact max-log of its axis times noise_inv[sc][l], to
    FP64 rounding (relative to (1 + |y|^2) noise_inv), on a grid of I values through all decision regions, random
    Q values and a distinct noise_inv per (subcarrier, layer). Negative controls: the same brute force with noise
    scaled by 1 + 1e-9, or with one PAM level moved by 1e-9, misses the bar, so a noise-scaling slip or a wrong
    LUT entry in the oracle cannot pass."""
    rng = np.random.default_rng(40 + q)
    nSC, nL, nS = 7, 3, 97
    Y = np.linspace(-1.7, 1.7, nSC * nL * nS).reshape(nSC, nL, nS) + 1j * rng.uniform(-1.7, 1.7, (nSC, nL, nS))
    ni = rng.uniform(0.5, 40.0, (nSC, nL))
    got = po.llr_demap_f64(Y, ni, q)

    def expected(noise, levels=None):
        ref = np.empty_like(got)
        ref[..., 0::2] = _maxlog_numpy(Y.real, noise[:, :, None], q, levels)
        ref[..., 1::2] = _maxlog_numpy(Y.imag, noise[:, :, None], q, levels)
        return np.clip(ref, -65504.0, 65504.0)

    scale = (1.0 + np.abs(Y)[..., None] ** 2) * ni[:, :, None, None]
    bar = 1e-13
    assert np.max(np.abs(got - expected(ni)) / scale) <= bar
    assert np.max(np.abs(got - expected(ni * (1 + 1e-9))) / scale) > bar
    c, lev = _gray_levels(q)
    lev = lev.copy()
    lev[-1] += 1e-9
    assert np.max(np.abs(got - expected(ni, (c, lev))) / scale) > bar


def test_llr_qpsk_closed_form(po):
    d = 1 / np.sqrt(2)
    for x in [-0.9, -0.2, 0.0, 0.3, 1.1]:
        bf = po.llr_bruteforce(x, 2.0, 1)[0]
        assert bf == pytest.approx(4 * d * x * 2.0, abs=1e-12)


@pytest.mark.parametrize("q", [1, 2, 3, 4])
def test_qam_map_and_demap_recover_bits(po, q):
    """Noise-free 38.211 symbols demap to LLRs whose signs give back the bits (positive = bit 0)."""
    rng = np.random.default_rng(10 + q)
    bits = rng.integers(0, 2, size=(2, 3, 5, 2 * q), dtype=np.uint8)
    x = po.qam_map(bits, q)
    assert abs(np.mean(np.abs(po.qam_map(np.array(
        [[(c >> i) & 1 for i in range(2 * q)] for c in range(4 ** q)]), q)) ** 2) - 1) < 1e-12
    llr = po.llr_demap_f64(x, np.ones((2, 3)), q)
    assert np.array_equal((llr < 0).astype(np.uint8), bits)


def test_qam_map_known_points(po):
    assert po.qam_map(np.array([0, 0, 0, 0]), 2) == pytest.approx((1 + 1j) / np.sqrt(10))
    assert po.qam_map(np.array([1, 0, 1, 0]), 2) == pytest.approx((-3 + 1j) / np.sqrt(10))
    assert po.qam_map(np.array([0, 1]), 1) == pytest.approx((1 - 1j) / np.sqrt(2))
    assert po.qam_map(np.zeros(8, np.uint8), 4) == pytest.approx((5 + 5j) / np.sqrt(170))
    assert po.qam_map(np.array([0, 0, 1, 1, 1, 1, 1, 1]), 4) == pytest.approx((15 + 15j) / np.sqrt(170))
    assert po.qam_map(np.array([0, 0, 0, 0, 0, 0]), 3) == pytest.approx((3 + 3j) / np.sqrt(42))


@pytest.mark.parametrize("q", [2, 3, 4])
def test_llr_f16_pipeline_close_to_f64_ref(po, q):
    """The reference's own accuracy check (aerial_llr_bench_v2.cu:379-394): binary16 output vs FP64 ref."""
    soft, ni = po.llr_bench_inputs(42, 139392)
    out = po.f16_bits_to_f32(po.llr_demap_f16(soft, ni, q))
    ref = po.llr_demap_ref_f64(soft, ni, q)
    assert po.cosine(ref, out) > 0.99999
    # FP32 arithmetic then binary16 rounding: within a binary16 ulp of the FP64 result
    assert np.all(np.abs(out - ref) <= np.maximum(np.abs(ref) * 2.0 ** -10, 1e-3))


def test_metrics_helpers(po):
    v = np.array([5, 1, 4, 2, 3, 10, 7, 6, 9, 8], dtype=np.float64)
    assert po.percentile_nearest_rank(v, 50) == 5
    assert po.percentile_nearest_rank(v, 99) == 10
    assert po.percentile_nearest_rank(v, 5) == 1
    assert po.median_upper(v) == 6
    ref = np.array([1 + 1j, 2 - 1j])
    r, db = po.evm(ref, ref * 1.001)
    assert r == pytest.approx(0.001) and db == pytest.approx(-60.0)


# ------------------------------------------------------------------------------------------------------
# binary16 arithmetic and the SRS filter multiply
# ------------------------------------------------------------------------------------------------------
def _finite_halves(rng, n):
    """Random finite binary16 bit patterns, a quarter of them subnormal or zero."""
    h = rng.integers(0, 0x7c00, n, dtype=np.uint32)
    sub = rng.random(n) < 0.25
    h[sub] = rng.integers(0, 0x400, int(sub.sum()), dtype=np.uint32)
    h |= (rng.random(n) < 0.5).astype(np.uint32) << 15
    return h.astype(np.uint16)


def test_half_mul_add_sub_match_numpy(po):
    """numpy evaluates binary16 operations in float32 and rounds once more; float32 carries 24 >= 2*11 + 2
    bits, so that second rounding is exact for +, -, * and numpy is an independent correctly rounded model."""
    rng = np.random.default_rng(0)
    a, b = _finite_halves(rng, 400000), _finite_halves(rng, 400000)
    c = np.zeros_like(a)
    got = po.half_ops(a, b, c)
    fa, fb = a.view(np.float16), b.view(np.float16)
    with np.errstate(over="ignore", invalid="ignore"):
        exp = {"mul": fa * fb, "add": fa + fb, "sub": fa - fb}
    for k, v in exp.items():
        ok = np.isfinite(v)
        assert ok.mean() > 0.9
        assert np.array_equal(got[k][ok], v.view(np.uint16)[ok]), k
        assert np.all((got[k][~ok] & 0x7fff) == 0x7c00), k  # overflow to +-inf


def _nearest_half_exact(x, table_vals, table_bits):
    """Brute force: the binary16 value nearest to the exact rational x, ties to the even significand."""
    from fractions import Fraction
    ax = abs(x)
    i = np.searchsorted(table_vals, float(ax))  # float(ax) is within an ulp of double: check neighbours exactly
    best = None
    for j in range(max(0, i - 2), min(len(table_vals), i + 3)):
        d = abs(Fraction(table_vals[j]) - ax)
        key = (d, table_bits[j] & 1)
        if best is None or key < best[0]:
            best = (key, table_bits[j])
    bits = int(best[1])
    if x < 0:
        bits |= 0x8000
    return bits


def test_half_fma_single_rounding_bruteforce(po):
    """fma(a, b, c) = the binary16 nearest to the exact a*b + c (ties to even), checked by exact rationals
    against a table of every finite binary16, on random triples and on triples built to land on ties."""
    from fractions import Fraction
    pos = np.arange(0, 0x7c00, dtype=np.uint16)
    vals = pos.view(np.float16).astype(np.float64)
    rng = np.random.default_rng(1)
    a, b, c = (_finite_halves(rng, 3000) for _ in range(3))
    # ties: a*b exactly halfway between two binary16 values, c tiny (the rounding direction is c's sign)
    ta = np.array([0x4200, 0x3c00 | 0x155, 0x4600, 0x3e00], dtype=np.uint16)  # 3, ~1.33, 6, 1.5
    tb = np.array([0x3c03, 0x3c01, 0x3c01, 0x3c01], dtype=np.uint16)
    for s in (0x0001, 0x8001, 0x0000, 0x8000):
        a = np.concatenate([a, ta]); b = np.concatenate([b, tb]); c = np.concatenate([c, np.full(4, s, np.uint16)])
    got = po.half_ops(a, b, c)["fma"]
    fa, fb, fc = (x.view(np.float16).astype(np.float64) for x in (a, b, c))
    checked = 0
    for i in range(a.size):
        x = Fraction(float(fa[i])) * Fraction(float(fb[i])) + Fraction(float(fc[i]))
        if abs(x) >= 65520:
            assert got[i] & 0x7fff == 0x7c00
            continue
        if x == 0:
            continue  # signed zeros: checked below
        assert got[i] == _nearest_half_exact(x, vals, pos), (hex(a[i]), hex(b[i]), hex(c[i]))
        checked += 1
    assert checked > 2500
    # signed zeros follow IEEE: (+0)*x + (-0) = +0, (-0)*x + (-0) = -0, x - x = +0
    z = po.half_ops(np.array([0x0000, 0x8000], np.uint16), np.array([0x3c00, 0x3c00], np.uint16),
                    np.array([0x8000, 0x8000], np.uint16))
    assert list(z["fma"]) == [0x0000, 0x8000]
    assert po.half_ops(np.array([0x4500], np.uint16), np.array([0x4500], np.uint16),
                       np.array([0], np.uint16))["sub"][0] == 0


def _srs_case(rng, nBlocks, nAnt, K, P, FL):
    to16 = lambda x: np.stack([x.real, x.imag], -1).astype(np.float16).view(np.uint16)
    S = to16((rng.uniform(-0.7, 0.7, (nBlocks, nAnt, K)) + 1j * rng.uniform(-0.7, 0.7, (nBlocks, nAnt, K))))
    W = to16(rng.uniform(-0.7, 0.7, (K, K)) + 1j * rng.uniform(-0.7, 0.7, (K, K)))
    F = to16(rng.uniform(-0.7, 0.7, (P, FL)) + 1j * rng.uniform(-0.7, 0.7, (P, FL)))
    return S, W, F


def test_srs_oracle_identity_filter_returns_srs(po):
    """W = I and FOCC in {1, -1, j, -j}: every product is exact, so H_est[b][a][sc][p] = FOCC[p][sc % FL] *
    SRS[b][a][sc] bit for bit -- pins the subcarrier, port and FOCC indexing and the output layout."""
    rng = np.random.default_rng(2)
    nB, nA, K, P, FL = 3, 5, 12, 3, 8
    S, _, _ = _srs_case(rng, nB, nA, K, P, FL)
    one = np.float16(1).view(np.uint16)
    W = np.zeros((K, K, 2), np.uint16)
    W[np.arange(K), np.arange(K), 0] = one
    units = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    fv = np.array([units[(3 * p + j) % 4] for p in range(P) for j in range(FL)], np.float16).reshape(P, FL, 2)
    F = fv.view(np.uint16)
    H = po.f16_complex(po.srs_filter_v1(S, W, F))
    s = po.f16_complex(S)
    f = fv[..., 0].astype(np.float64) + 1j * fv[..., 1].astype(np.float64)
    exp = s[:, :, :, None] * f.T[np.arange(K) % FL][None, None, :, :]
    assert np.array_equal(H, exp)
    # a conjugated W changes nothing here (W real); a j on the diagonal turns into -j with conj_w
    W[np.arange(K), np.arange(K), 0] = 0
    W[np.arange(K), np.arange(K), 1] = one
    Hj = po.f16_complex(po.srs_filter_v1(S, W, F))
    Hc = po.f16_complex(po.srs_filter_v1(S, W, F, conj_w=True))
    assert np.array_equal(Hj, 1j * exp) and np.array_equal(Hc, -1j * exp)


@pytest.mark.parametrize("K,P,FL", [(24, 4, 8), (12, 2, 12), (7, 3, 5)])
def test_srs_oracle_close_to_exact_sum(po, K, P, FL):
    """The binary16 chain stays within its rounding bound of the exact sum (built from the same binary16
    taps) and far from the conjugated-W or unrounded-tap alternatives only where those differ."""
    rng = np.random.default_rng(K + P)
    S, W, F = _srs_case(rng, 4, 16, K, P, FL)
    H = po.f16_complex(po.srs_filter_v1(S, W, F))
    X = po.srs_filter_f64(S, W, F, prod_f16=True)
    rel = np.linalg.norm(H - X) / np.linalg.norm(X)
    assert rel < 2e-3
    # per element, against the exact sum: a rounding-error bound (2K roundings of at most 2^-11 relative to
    # the running sums, which stay below sum|s| max|wf|)
    bound = (2 * K + 2) * 2.0 ** -11 * np.abs(po.f16_complex(S)).sum(-1).max() * 1.0
    assert np.abs(H - X).max() < bound
    Xc = po.srs_filter_f64(S, W, F, conj_w=True, prod_f16=True)
    assert np.linalg.norm(H - Xc) / np.linalg.norm(Xc) > 0.5
    Hc = po.f16_complex(po.srs_filter_v1(S, W, F, conj_w=True))
    assert np.linalg.norm(Hc - Xc) / np.linalg.norm(Xc) < 2e-3


def test_srs_bench_inputs_follow_the_stream(po):
    """srand(42) stream order: all SRS values, then W, then FOCC, {re, im} per value, scaled by 1.4."""
    S, W, F = po.srs_bench_inputs(42, 4, 2, 12, 2, 12)
    u = po.glibc_uniform(42, 2 * (S.size // 2 + W.size // 2 + F.size // 2), scale=1.4)
    bits = po.f32_to_f16_bits(u)
    n1, n2 = S.size, S.size + W.size
    assert np.array_equal(S.ravel(), bits[:n1]) and np.array_equal(W.ravel(), bits[n1:n2])
    assert np.array_equal(F.ravel(), bits[n2:])


# ------------------------------------------------------------------------------------------------------
# fronthaul block floating point
# ------------------------------------------------------------------------------------------------------
def _bfp_block(exponent, mantissas, w):
    """One compressed PRB built from a bit string written out here (independent of the oracle's packer)."""
    bits = format(exponent & 0xF, "08b") + "".join(format(m & ((1 << w) - 1), "0%db" % w) for m in mantissas)
    bits += "0" * (-len(bits) % 8)
    return np.array([int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)], dtype=np.uint8)


@pytest.mark.parametrize("w", [1, 8, 9, 14, 16])
def test_bfp_decompress_known_blocks(po, w):
    """Hand-packed PRBs: two's complement extremes, sign extension, exponent, odd block lengths and padded
    rows; subcarrier = 12 prb + re, I then Q."""
    rng = np.random.default_rng(w)
    lo, hi = -(1 << (w - 1)), (1 << (w - 1)) - 1
    nPrb, nSym, nAnt = 3, 2, 2
    bpb = po.bfp_block_bytes(w)
    assert all(len(_bfp_block(0, [0] * 24, w)) == bpb for _ in range(1))
    stride = nPrb * bpb + 5  # padded rows
    packed = np.full((nSym * nAnt, stride), 0xAB, dtype=np.uint8)
    expect = np.zeros((nPrb * 12, nSym, nAnt), dtype=np.complex128)
    for sym in range(nSym):
        for ant in range(nAnt):
            for prb in range(nPrb):
                e = int(rng.integers(0, 16))
                m = rng.integers(lo, hi + 1, 24)
                m[:4] = [lo, hi, -1 if w > 1 else lo, 0]
                packed[sym * nAnt + ant, prb * bpb:(prb + 1) * bpb] = _bfp_block(e | 0xA0, m, w)  # reserved bits set
                expect[prb * 12:(prb + 1) * 12, sym, ant] = (m[0::2] + 1j * m[1::2]) * 2.0 ** e
    Y = po.bfp_decompress(packed, nPrb, nSym, nAnt, w, scale=2.0 ** -15)
    assert np.array_equal(Y.astype(np.complex128), expect * 2.0 ** -15)


@pytest.mark.parametrize("w", [9, 14])
def test_bfp_round_trip(po, w):
    """The oracle encoder picks the smallest exponent that fits; decoding returns the samples rounded toward
    minus infinity at that exponent, exactly when it is 0."""
    rng = np.random.default_rng(40 + w)
    nPrb, nSym, nAnt = 5, 3, 4
    x = rng.integers(-(1 << 15), 1 << 15, (nSym, nAnt, nPrb * 12, 2)).astype(np.int32)
    x[0, 0, :12] //= 1 << 10  # a PRB small enough for exponent 0 at w = 9
    packed = po.bfp_compress(x, w)
    Y = po.bfp_decompress(packed, nPrb, nSym, nAnt, w)
    got = np.stack([Y.real, Y.imag], -1).transpose(1, 2, 0, 3)  # [sym, ant, sc, 2]
    e = packed[:, ::po.bfp_block_bytes(w)][:, :nPrb].reshape(nSym, nAnt, nPrb) & 0xF
    step = np.repeat(2.0 ** e, 12, axis=2)[..., None]
    assert np.all(got <= x) and np.all(x - got < step)
    assert np.all((x // step) * step == got)
    assert e[0, 0, 0] == 0 and np.array_equal(got[0, 0, :12], x[0, 0, :12])
    # minimal: at exponent e - 1 some value of the PRB would not fit
    lo, hi = -(1 << (w - 1)), (1 << (w - 1)) - 1
    xb = x.reshape(nSym, nAnt, nPrb, 24)
    for idx in zip(*np.nonzero(e)):
        m = xb[idx] >> (int(e[idx]) - 1)
        assert m.max() > hi or m.min() < lo


# ------------------------------------------------------------------------------------------------------
# DMRS frequency interpolation (README.md:200 names the GEMM; no reference code)
# ------------------------------------------------------------------------------------------------------
def _fi_case(rng, nB, M, K, NC):
    c = lambda *s: (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(np.complex64)
    return c(M, K), c(nB, K, NC)


@pytest.mark.parametrize("nB,M,K,NC", [(3, 64, 96, 4), (2, 5, 7, 3), (1, 1, 1, 1), (4, 17, 128, 16)])
def test_freq_interp_f64_matches_numpy_matmul(po, nB, M, K, NC):
    """The FP64 oracle against numpy's complex matrix product (a library routine, independent of the loop)."""
    rng = np.random.default_rng(M * 1000 + K)
    W, X = _fi_case(rng, nB, M, K, NC)
    got = po.freq_interp_f64(W, X)
    ref = np.matmul(W.astype(np.complex128)[None], X.astype(np.complex128))
    assert np.allclose(got, ref, rtol=0, atol=1e-12 * K)


@pytest.mark.parametrize("nB,M,K,NC", [(3, 64, 96, 4), (2, 5, 7, 3), (4, 17, 128, 16)])
def test_freq_interp_f32_chain_within_fp32_bound(po, nB, M, K, NC):
    """The FP32 chain is within the forward error bound of 2K fused steps of the exact product, element by
    element; a dropped, conjugated or transposed term would be O(|W| |X|) off."""
    rng = np.random.default_rng(K)
    W, X = _fi_case(rng, nB, M, K, NC)
    got = po.freq_interp_f32(W, X).astype(np.complex128)
    ref = po.freq_interp_f64(W, X)
    mag = np.matmul(np.abs(W.astype(np.complex128))[None], np.abs(X.astype(np.complex128)))
    bound = 2 * K * 2.0 ** -24 * mag * 2
    assert np.all(np.abs(got.real - ref.real) <= bound) and np.all(np.abs(got.imag - ref.imag) <= bound)
    assert np.abs(got - ref).max() > 0  # FP32, not a copy of the FP64 sum


def test_freq_interp_selection_filters_are_exact(po):
    """W = unit-phase selection (W[m][k] = u_m where k = s(m), 0 elsewhere, u_m in {1, i, -1, -i}) picks rows of
    X rotated exactly: catches a wrong sign, a swapped real/imaginary part or a transposed operand."""
    rng = np.random.default_rng(5)
    M, K, NC, nB = 40, 48, 4, 3
    s = rng.integers(0, K, M)
    u = np.array([1, 1j, -1, -1j])[rng.integers(0, 4, M)]
    W = np.zeros((M, K), dtype=np.complex64)
    W[np.arange(M), s] = u
    X = _fi_case(rng, nB, M, K, NC)[1]
    got = po.freq_interp_f32(W, X)
    assert np.array_equal(got, (u[None, :, None] * X[:, s, :]).astype(np.complex64))
    # the omp variant computes the same chain
    omp, used = po.freq_interp_f32_omp(W, X, nthreads=2)
    assert used >= 1 and np.array_equal(omp, got)


def test_round_bf16_matches_torch():
    """The test helper's bfloat16 rounding against torch's (round to nearest even), ties included."""
    import torch
    from pe_oracle import round_bf16
    rng = np.random.default_rng(9)
    v = rng.standard_normal(4096).astype(np.float32)
    ties = (np.arange(64, dtype=np.uint32) << 16 | 0x8000).view(np.float32)  # exactly half-way
    v = np.concatenate([v, ties, -ties, np.float32([0.0, -0.0, 1e-40, 3e38])])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(round_bf16(v).view(np.uint32), ref.view(np.uint32))


def test_published_fixture_is_consistent():
    """The fixture's EVM in percent and in dB describe the same ratio (20 log10), and its shape is the headline's."""
    assert abs(20 * np.log10(PUBLISHED["coef_rigor_evm_pct"] / 100) - RIGOR_EVM_DB) < 0.05
    assert PUBLISHED["coef_rigor_shape"] == [NSC, L, S, A]


# ---------------------------------------------------------------------------------------------------------
# descrambling and rate-recovery layout (TS 38.211 5.2.1, 6.3.1; TS 38.212 5.4.2)
# ---------------------------------------------------------------------------------------------------------
F1 = (1 << 31) | (1 << 3) | 1                          # x1: t^31 + t^3 + 1
F2 = (1 << 31) | (1 << 3) | (1 << 2) | (1 << 1) | 1    # x2: t^31 + t^3 + t^2 + t + 1


def _gf2_mulmod(a, b, f):
    """a * b mod f over GF(2), polynomials as Python ints."""
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if a >> 31 & 1:
            a ^= f
    return r


def _gf2_powmod(k, f):
    r, base = 1, 2  # t^0, t^1
    while k:
        if k & 1:
            r = _gf2_mulmod(r, base, f)
        base = _gf2_mulmod(base, base, f)
        k >>= 1
    return r


def _lfsr_term(k, f, init):
    """x(k) of the linear recurrence with characteristic polynomial f (degree 31) from x(0..30) = bits of init:
    x(k) = sum_i [t^k mod f]_i x(i) over GF(2) (the shift operator t acting on the sequence space)."""
    return bin(_gf2_powmod(k, f) & init).count("1") & 1


@pytest.mark.parametrize("c_init", [0, 1, 0x4000_0000, 0x7FFF_FFFF, (1234 << 15) + 567, 0x2468ACE])
def test_gold_sequence_equals_the_polynomial_residue_route(po, c_init):
    """c(n) = x1(n + 1600) + x2(n + 1600) with each term computed as sum_i [t^(n+1600) mod f]_i x(i) from the
    initial values (x1: x1(0) = 1; x2: bit i of c_init), a route that shares nothing with the oracle's bit-serial
    loop but the definition: pins the taps, Nc = 1600 and the c_init bit order."""
    c = po.gold_sequence(c_init, 300)
    idx = list(range(0, 40)) + [97, 128, 255, 299]
    ref = [_lfsr_term(n + 1600, F1, 1) ^ _lfsr_term(n + 1600, F2, c_init & 0x7FFF_FFFF) for n in idx]
    assert list(c[idx]) == ref


def test_gold_sequence_recurrence_and_linearity(po):
    """c obeys the degree-62 recurrence of f1 f2 (it is the sum of the two register sequences), c_init = 0 leaves
    the x1 m-sequence alone (recurrence of f1), and c is affine in c_init: c(a ^ b) = c(a) ^ c(b) ^ c(0)."""
    n = 4000
    fprod = 0
    for i in range(32):  # carry-less product f1 * f2
        if F2 >> i & 1:
            fprod ^= F1 << i
    taps = [i for i in range(62) if fprod >> i & 1]
    rng = np.random.default_rng(0)
    for c_init in (0, 1, int(rng.integers(0, 2 ** 31)), 0x7FFF_FFFF):
        c = po.gold_sequence(c_init, n).astype(np.int64)
        lhs = c[62:]
        rhs = np.zeros(n - 62, np.int64)
        for t in taps:
            rhs ^= c[t:n - 62 + t]
        assert np.array_equal(lhs, rhs)
    c0 = po.gold_sequence(0, n)
    assert np.array_equal(c0[31:], c0[3:n - 28] ^ c0[:n - 31]) and c0.any()
    for _ in range(5):
        a, b = (int(x) for x in rng.integers(0, 2 ** 31, 2))
        assert np.array_equal(po.gold_sequence(a ^ b, 999), po.gold_sequence(a, 999) ^ po.gold_sequence(b, 999) ^
                              po.gold_sequence(0, 999))


def test_code_block_sizes_known_answer(po):
    """38.212 5.4.2.1 by hand: G = 2112 bits on 2 layers of 256-QAM (N = 132 symbols per layer) in C = 5 blocks:
    floor(132/5) = 26, ceil = 27, mod(132, 5) = 2, so blocks 0..2 get 2*8*26 = 416 bits and blocks 3, 4 get 432."""
    assert list(po.cw_block_sizes(2112, 2, 8, 5)) == [416, 416, 416, 432, 432]
    assert list(po.cw_block_sizes(2112, 2, 8, 4)) == [528] * 4
    assert po.cw_block_sizes(2112, 2, 8, 133) is None and po.cw_block_sizes(2112, 2, 8, 0) is None
    for G, nu, qm, C in ((9504, 4, 6, 7), (288, 1, 2, 144), (1_368_576, 8, 8, 43)):
        E = po.cw_block_sizes(G, nu, qm, C)
        assert E.sum() == G and np.all(E % (nu * qm) == 0) and E.max() - E.min() <= nu * qm
        assert np.all(np.diff(E) >= 0)


def _interleave(e, qm):
    """38.212 5.4.2.2 as written: for j < E/Qm, i < Qm: f(i + j Qm) = e(i E/Qm + j)."""
    E = e.size
    f = np.empty_like(e)
    for j in range(E // qm):
        for i in range(qm):
            f[i + j * qm] = e[i * (E // qm) + j]
    return f


@pytest.mark.parametrize("q,nu,C", [(4, 2, 5), (3, 3, 7), (1, 1, 1), (2, 4, 3)])
def test_deinterleave_inverts_the_bit_interleaver(po, q, nu, C):
    """Per code block, applying the 38.212 interleaver to the oracle's de-interleaved output gives back the
    codeword-order output (the f sequence)."""
    rng = np.random.default_rng(q * 10 + nu)
    nSC, S, qm = 12, 11, 2 * q
    llr = rng.standard_normal((nSC, nu, S, qm)).astype(np.float32)
    f = po.cw_layout(llr, [0] * nu)[0]
    e = po.cw_layout(llr, [0] * nu, n_code_blocks=[C])[0]
    E = po.cw_block_sizes(f.size, nu, qm, C)
    starts = np.concatenate([[0], np.cumsum(E)[:-1]])
    for r in range(C):
        blk = slice(starts[r], starts[r] + E[r])
        assert np.array_equal(_interleave(e[blk], qm), f[blk])


@pytest.mark.parametrize("q,layer_cw", [(4, [0, 0, 0, 0]), (2, [0, 1, 0, 1, 2, 2, 2, 3]), (3, [0]), (1, [1, 0])])
def test_descrambled_layout_recovers_the_codeword_bits(po, q, layer_cw):
    """End to end against a transmitter written from TS 38.211 6.3.1: codeword bits b -> scrambling with c ->
    modulation (Qm consecutive bits per symbol) -> layer mapping (codeword symbol m on layer m mod nu, layer symbol
    m div nu) -> resource elements frequency first (layer symbol i at subcarrier i mod nSC of data symbol
    i div nSC). The oracle's FP64 demapper on the noise-free symbols, then its codeword layout with descrambling,
    returns LLRs whose signs are b (positive = 0) in codeword order, and de-interleaved per block they are the
    de-interleaved b."""
    rng = np.random.default_rng(q + 7 * len(layer_cw))
    nSC, S, qm = 24, 5, 2 * q
    L = len(layer_cw)
    ncw = max(layer_cw) + 1
    nu = np.bincount(layer_cw, minlength=ncw)
    yeq = np.zeros((nSC, L, S), dtype=np.complex128)
    bits, cs = [], []
    for k in range(ncw):
        G = nSC * S * nu[k] * qm
        b = rng.integers(0, 2, G, dtype=np.uint8)
        c = po.gold_sequence((1000 + k) << 15 | 77, G)
        sym = po.qam_map((b ^ c).reshape(-1, qm), q)
        layers = [l for l in range(L) if layer_cw[l] == k]
        for m in range(sym.size):
            i = m // nu[k]
            yeq[i % nSC, layers[m % nu[k]], i // nSC] = sym[m]
        bits.append(b)
        cs.append(c)
    llr = po.llr_demap_f64(yeq, np.ones((nSC, L)), q).astype(np.float32)
    outs = po.cw_layout(llr, layer_cw, cs)
    for k in range(ncw):
        assert np.array_equal((outs[k] < 0).astype(np.uint8), bits[k]), "codeword %d" % k
    outs_e = po.cw_layout(llr, layer_cw, cs, [3] * ncw)
    for k in range(ncw):
        E = po.cw_block_sizes(bits[k].size, nu[k], qm, 3)
        starts = np.concatenate([[0], np.cumsum(E)[:-1]])
        for r in range(3):
            blk = slice(starts[r], starts[r] + E[r])
            assert np.array_equal(_interleave((outs_e[k][blk] < 0).astype(np.uint8), qm), bits[k][blk])
    # without descrambling the signs are the scrambled bits
    raw = po.cw_layout(llr, layer_cw)
    assert np.array_equal((raw[0] < 0).astype(np.uint8), bits[0] ^ cs[0])
