"""GPU parity tests of the DMRS frequency interpolation (pe_freq_interp) against the oracle.

The reference names this GEMM ([64 x 96] @ [96 x 4], batched over OFDM symbols and UEs; README.md:200) without
code, so the oracle's definition H_f[b] = W @ X[b] is the contract. Bars: PE_PREC_F32 is bit-identical to the
oracle's FP32 chain (cma() order, k ascending) on every shape, ragged batches included; the bf16 MFMA paths are
within the FP32 accumulation bound of the exact sum of their rounded operands (bf16) or of the exact operands
(split); unit-phase selection filters on bf16-exact data come out exact on every path (a wrong sign, swapped
parts or a transposed operand fail it); the full bench batch is checked item by item on a sample."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [  # (nBatch, nOut, nIn, nCol)
    (300, 64, 96, 4),    # the reference's shape
    (1001, 64, 96, 1),
    (13, 64, 96, 16),
    (7, 48, 24, 2),
    (1, 1, 1, 1),
    (33, 17, 33, 3),     # 5 items per MFMA tile, 15 of 16 rows
    (5, 256, 12, 16),    # 4 subcarrier groups
    (9, 48, 128, 8),
    (4, 5, 127, 7),
]
MFMA_ONLY = [(17, 100, 96, 4), (4097, 64, 96, 4), (6, 200, 128, 5)]


def _case(rng, nB, M, K, NC):
    c = lambda *s: (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(np.complex64)
    return c(M, K), c(nB, K, NC)


def _run(pe, W, X, prec):
    import torch
    Wd = torch.from_numpy(W).cuda()
    Xd = torch.from_numpy(X).cuda()
    out = pe.freq_interp(Wd, Xd, prec=prec)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _mag(W, X):
    return np.matmul(np.abs(W.astype(np.complex128))[None], np.abs(X.astype(np.complex128)))


def _evm_db(got, ref):
    return 10 * np.log10(np.sum(np.abs(got - ref) ** 2) / np.sum(np.abs(ref) ** 2))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_f32_bit_exact(pe, po, shape):
    nB, M, K, NC = shape
    rng = np.random.default_rng(nB * 7 + M)
    W, X = _case(rng, nB, M, K, NC)
    got = _run(pe, W, X, "f32")
    assert np.array_equal(got.view(np.uint32), po.freq_interp_f32(W, X).view(np.uint32))


@pytest.mark.parametrize("shape", SHAPES + MFMA_ONLY, ids=lambda s: "x".join(map(str, s)))
def test_bf16_within_accumulation_bound(pe, po, shape):
    """bf16 MFMA: the exact sum of the bf16-rounded operands, up to FP32 accumulation of 2 nIn real products."""
    nB, M, K, NC = shape
    rng = np.random.default_rng(nB * 11 + K)
    W, X = _case(rng, nB, M, K, NC)
    got = _run(pe, W, X, "bf16").astype(np.complex128)
    ref = po.freq_interp_f64(po.round_bf16_complex(W), po.round_bf16_complex(X))
    bound = 2 * (2 * K) * 2.0 ** -24 * _mag(W, X) + 1e-30
    assert np.all(np.abs(got.real - ref.real) <= bound)
    assert np.all(np.abs(got.imag - ref.imag) <= bound)
    assert _evm_db(got, po.freq_interp_f64(W, X)) < -45


@pytest.mark.parametrize("shape", SHAPES + MFMA_ONLY, ids=lambda s: "x".join(map(str, s)))
def test_bf16_split_near_fp32(pe, po, shape):
    """hi + lo bf16: within 2^-15 |W| |X| of the exact product of the unrounded operands (the dropped lo*lo
    term and the rounding of the lo parts), plus FP32 accumulation."""
    nB, M, K, NC = shape
    rng = np.random.default_rng(nB * 13 + NC)
    W, X = _case(rng, nB, M, K, NC)
    got = _run(pe, W, X, "bf16_split").astype(np.complex128)
    ref = po.freq_interp_f64(W, X)
    bound = (2.0 ** -15 + 2 * (2 * K) * 2.0 ** -24) * _mag(W, X) + 1e-30
    assert np.all(np.abs(got.real - ref.real) <= bound)
    assert np.all(np.abs(got.imag - ref.imag) <= bound)
    assert _evm_db(got, ref) < -80


@pytest.mark.parametrize("prec", ["f32", "bf16", "bf16_split"])
def test_selection_filter_exact(pe, po, prec):
    """Unit-phase selection on bf16-exact data: every path returns X's rows rotated by 1, i, -1 or -i exactly."""
    rng = np.random.default_rng(21)
    nB, M, K, NC = 29, 70, 96, 4
    s = rng.integers(0, K, M)
    u = np.array([1, 1j, -1, -1j])[rng.integers(0, 4, M)]
    W = np.zeros((M, K), dtype=np.complex64)
    W[np.arange(M), s] = u
    X = po.round_bf16_complex(_case(rng, nB, M, K, NC)[1])
    got = _run(pe, W, X, prec)
    assert np.array_equal(got, (u[None, :, None] * X[:, s, :]).astype(np.complex64))


def test_full_batch_sampled_items(pe, po):
    """The bench batch (57,344 items of 96 x 4 through a 64 x 96 filter): 256 sampled items bit-exact (FP32)
    and within the bf16 bound; every output finite."""
    import torch
    nB, M, K, NC = 14 * 64 * 64, 64, 96, 4
    g = torch.Generator(device="cuda").manual_seed(3)
    Xd = torch.randn((nB, K, NC), dtype=torch.complex64, device="cuda", generator=g)
    Wd = torch.randn((M, K), dtype=torch.complex64, device="cuda", generator=g)
    f32 = pe.freq_interp(Wd, Xd, prec="f32")
    bf = pe.freq_interp(Wd, Xd, prec="bf16")
    torch.cuda.synchronize()
    assert bool(torch.isfinite(torch.view_as_real(f32)).all()) and bool(torch.isfinite(torch.view_as_real(bf)).all())
    idx = np.sort(np.random.default_rng(4).choice(nB, 256, replace=False))
    idx[-1] = nB - 1
    W = Wd.cpu().numpy()
    Xs = Xd[torch.from_numpy(idx).cuda()].cpu().numpy()
    assert np.array_equal(f32[torch.from_numpy(idx).cuda()].cpu().numpy().view(np.uint32),
                          po.freq_interp_f32(W, Xs).view(np.uint32))
    ref = po.freq_interp_f64(po.round_bf16_complex(W), po.round_bf16_complex(Xs))
    got = bf[torch.from_numpy(idx).cuda()].cpu().numpy().astype(np.complex128)
    bound = 2 * (2 * K) * 2.0 ** -24 * _mag(W, Xs)
    assert np.all(np.abs(got.real - ref.real) <= bound) and np.all(np.abs(got.imag - ref.imag) <= bound)


def test_output_untouched_outside_batch(pe):
    """Writes stay inside H_f: a guard region after the batch keeps its values (ragged last MFMA tile)."""
    import torch
    nB, M, K, NC = 7, 64, 96, 3
    W = torch.randn((M, K), dtype=torch.complex64, device="cuda")
    X = torch.randn((nB, K, NC), dtype=torch.complex64, device="cuda")
    for prec in ("f32", "bf16", "bf16_split"):
        buf = torch.full((nB + 2, M, NC), complex(7.0, -7.0), dtype=torch.complex64, device="cuda")
        pe.freq_interp(W, X, prec=prec, out=buf[:nB])
        torch.cuda.synchronize()
        assert bool((buf[nB:] == complex(7.0, -7.0)).all())


@pytest.mark.parametrize("prec", ["f32", "bf16", "bf16_split"])
def test_unaligned_operands(pe, po, prec):
    """W, X and H_f 8-byte but not 16-byte aligned (views one complex value into larger buffers): at 4 columns
    the MFMA kernel then stores element by element; results equal those of aligned operands."""
    import torch
    rng = np.random.default_rng(31)
    nB, M, K, NC = 45, 64, 96, 4
    W, X = _case(rng, nB, M, K, NC)
    Wb = torch.zeros(M * K + 1, dtype=torch.complex64, device="cuda")
    Xb = torch.zeros(nB * K * NC + 1, dtype=torch.complex64, device="cuda")
    Hb = torch.full((nB * M * NC + 2,), complex(5.0, 5.0), dtype=torch.complex64, device="cuda")
    Wu = Wb[1:].view(M, K)
    Xu = Xb[1:].view(nB, K, NC)
    Hu = Hb[1:1 + nB * M * NC].view(nB, M, NC)
    assert Wu.data_ptr() % 16 == 8 and Xu.data_ptr() % 16 == 8 and Hu.data_ptr() % 16 == 8
    Wu.copy_(torch.from_numpy(W))
    Xu.copy_(torch.from_numpy(X))
    pe.freq_interp(Wu, Xu, prec=prec, out=Hu)
    ref = _run(pe, W, X, prec)
    torch.cuda.synchronize()
    assert np.array_equal(Hu.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert complex(Hb[0]) == complex(5.0, 5.0) and complex(Hb[-1]) == complex(5.0, 5.0)
