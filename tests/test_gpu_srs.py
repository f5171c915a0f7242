"""GPU parity tests of the SRS filter multiply (pe_srs_filter) against the oracle.

Bars: the binary16 chain (flags 0) is bit-identical to the oracle's restatement of srs_filter_v1_scalar
(srs_wmma_bench.cu:57-90) on the reference benchmark's own inputs and on ragged shapes; the MFMA variant
(FP32 accumulation, one binary16 rounding) is within one binary16 ulp of the exact sum of the same binary16
taps, element by element, and exact where every product is exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def to_dev(bits, dev):
    """[..., 2] uint16 {re, im} -> int32 tensor [...] on the device (one word per complex value)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(bits, dtype=np.uint16).view(np.int32)[..., 0]).to(dev)


def to_bits(t):
    return t.cpu().numpy().view(np.uint16).reshape(tuple(t.shape) + (2,))


def ordered(bits):
    """binary16 bit patterns -> integers ordered like the values (+0 and -0 both 0)."""
    b = bits.astype(np.int32)
    return np.where(b & 0x8000, -(b & 0x7fff), b & 0x7fff)


def rand_case(rng, nB, nA, K, P, FL):
    to16 = lambda x: np.stack([x.real, x.imag], -1).astype(np.float16).view(np.uint16)
    c = lambda shape: rng.uniform(-0.7, 0.7, shape) + 1j * rng.uniform(-0.7, 0.7, shape)
    return to16(c((nB, nA, K))), to16(c((K, K))), to16(c((P, FL)))


@pytest.mark.parametrize("K,P,FL", [(24, 4, 8), (24, 2, 8), (12, 4, 12), (12, 2, 12)])
def test_srs_f16_bitexact_on_reference_inputs(po, pe, dev, K, P, FL):
    """srs_wmma_bench.cu run_config (:262-268): srand(42), 64 antennas, 66 blocks; every bit equal."""
    S, W, F = po.srs_bench_inputs(42, 64, 66, K, P, FL)
    ref = po.srs_filter_v1(S, W, F)
    out = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev)))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nB,nA,K,P,FL", [(1, 1, 1, 1, 1), (3, 17, 5, 3, 7), (2, 100, 12, 8, 12),
                                          (2, 130, 48, 1, 48), (5, 64, 24, 4, 8), (1, 65, 13, 2, 4)])
@pytest.mark.parametrize("conj", [False, True])
def test_srs_f16_bitexact_ragged(po, pe, dev, nB, nA, K, P, FL, conj):
    rng = np.random.default_rng(nB * 1000 + nA + K)
    S, W, F = rand_case(rng, nB, nA, K, P, FL)
    ref = po.srs_filter_v1(S, W, F, conj_w=conj)
    out = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev), conj_w=conj))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nB,nA,K,P,FL", [(66, 64, 24, 4, 8), (66, 64, 12, 2, 12), (3, 17, 4, 1, 3),
                                          (2, 100, 48, 8, 48), (4, 130, 16, 3, 5), (2, 64, 32, 5, 8)])
@pytest.mark.parametrize("conj", [False, True])
def test_srs_mfma_within_one_ulp_of_exact(po, pe, dev, nB, nA, K, P, FL, conj):
    """FP32 accumulation of exact binary16 products, one rounding: the binary16 rounding of the exact sum
    (built from the same binary16 taps) for almost every element; every element within one binary16 ulp plus
    the FP32 accumulation bound 2K * 2^-23 * sum|terms| (which matters only where the sum cancels); and
    closer to the exact sum than the binary16 chain."""
    rng = np.random.default_rng(7 * K + P)
    S, W, F = rand_case(rng, nB, nA, K, P, FL)
    exact = po.srs_filter_f64(S, W, F, conj_w=conj, prod_f16=True)
    out = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev), mfma=True, conj_w=conj))
    rn = np.stack([exact.real, exact.imag], -1).astype(np.float16).view(np.uint16)
    d = np.abs(ordered(out) - ordered(rn))
    assert (d == 0).mean() > 0.99 and (d <= 1).mean() > 0.9999
    s = po.f16_complex(S)
    wf_max = np.abs(po.f16_complex(W)).max() * np.abs(po.f16_complex(F)).max() * (1 + 2.0 ** -9)
    terms = 2 * np.abs(s).sum(-1) * wf_max  # >= sum of |products| in each real output
    ex = np.stack([exact.real, exact.imag], -1)
    ulp = np.spacing(np.abs(ex).astype(np.float16)).astype(np.float64)
    err = np.abs(out.view(np.float16).astype(np.float64) - ex)
    assert np.all(err <= ulp + 2 * K * 2.0 ** -23 * terms[:, :, None, None, None])
    chain = po.f16_complex(po.srs_filter_v1(S, W, F, conj_w=conj))
    got = po.f16_complex(out)
    assert np.linalg.norm(got - exact) <= np.linalg.norm(chain - exact)


def test_srs_mfma_exact_for_identity_filter(po, pe, dev):
    """W = I, FOCC = 1: one nonzero exact product per output, so H_est = SRS (every port) exactly."""
    rng = np.random.default_rng(3)
    nB, nA, K, P, FL = 4, 70, 24, 4, 8
    S, _, _ = rand_case(rng, nB, nA, K, P, FL)
    one = np.float16(1).view(np.uint16)
    W = np.zeros((K, K, 2), np.uint16)
    W[np.arange(K), np.arange(K), 0] = one
    F = np.zeros((P, FL, 2), np.uint16)
    F[..., 0] = one
    out = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev), mfma=True))
    assert np.array_equal(out, np.repeat(S[:, :, :, None, :], P, axis=3))
    out16 = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev)))
    assert np.array_equal(out16, out)


def test_srs_matches_reference_cosine_check(po, pe, dev):
    """The reference's own acceptance check (cosine of V1 vs V2, srs_wmma_bench.cu:245-255,328): the MFMA
    output against the binary16 chain on the benchmark inputs."""
    S, W, F = po.srs_bench_inputs(42, 64, 66, 24, 4, 8)
    a = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev)))
    b = to_bits(pe.srs_filter(to_dev(S, dev), to_dev(W, dev), to_dev(F, dev), mfma=True))
    assert po.cosine(po.f16_complex(a), po.f16_complex(b)) > 0.99999
