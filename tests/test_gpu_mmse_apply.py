"""pe_mmse_apply_demap: MMSE coefficients and the fused apply + demap in one pass, the coefficients never
written to HBM. Its LLRs, noise_inv and flags must be bit-identical to pe_gram_mmse followed by
pe_coef_apply_demap (pe_coef_apply_demap_f16 for binary16 LLRs) at the same precision, and with PE_EQ_WRITE_C
its coefficient buffer must equal pe_gram_mmse's. Covers every MFMA precision, layer counts 1-8, antenna
classes 4-128, ragged batches, the FP64 fallback (forced by a low condition threshold) and the chain against
the FP64 oracle at the headline shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(dev, nsc, L, S, A, seed):
    import torch
    g = torch.Generator(device=dev).manual_seed(seed)
    H = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g) / A ** 0.5
    s2 = torch.rand((nsc,), device=dev, generator=g) * 0.1 + 0.001
    Y = torch.randn((nsc, S, A), dtype=torch.complex64, device=dev, generator=g)
    return H, s2, Y


def _separate(pe, H, s2, Y, q, prec, thr=0.0, llr_f16=False, unbiased=True):
    C, ni, fl = pe.gram_mmse(H, s2, unbiased=unbiased, cond_threshold=thr)
    f = pe.coef_apply_demap_f16 if llr_f16 else pe.coef_apply_demap
    return f(C, Y, ni, q, prec=prec), C, ni, fl


def _check(pe, dev, nsc, L, S, A, q, prec, thr=0.0, llr_f16=False, write_c=False, unbiased=True, seed=1):
    import torch
    H, s2, Y = _inputs(dev, nsc, L, S, A, seed)
    ref, C, ni, fl = _separate(pe, H, s2, Y, q, prec, thr, llr_f16, unbiased)
    llr, C2, ni2, fl2 = pe.mmse_apply_demap(H, s2, Y, q, prec=prec, unbiased=unbiased, cond_threshold=thr,
                                            write_c=write_c, llr_f16=llr_f16)
    print("nsc=%d L=%d S=%d A=%d q=%d %s thr=%g flagged=%d" % (nsc, L, S, A, q, prec, thr, int((fl != 0).sum())))
    assert llr.dtype == (torch.float16 if llr_f16 else torch.float32)
    assert torch.equal(llr, ref)
    assert torch.equal(ni2, ni)
    assert torch.equal(fl2, fl)
    if write_c:
        assert torch.equal(C2, C)
    return fl


@pytest.mark.parametrize("prec", ["f16", "bf16", "f16_split", "bf16_split"])
def test_headline_shape_every_precision(pe, dev, prec):
    _check(pe, dev, 1584, 8, 11, 64, 4, prec)


@pytest.mark.parametrize("L,S,A,q", [(1, 1, 4, 1), (2, 11, 16, 2), (3, 5, 12, 3), (4, 11, 32, 4), (8, 16, 48, 2),
                                     (2, 11, 128, 4), (8, 11, 128, 4), (5, 7, 100, 3), (6, 14, 64, 4),
                                     (7, 2, 8, 1), (8, 11, 16, 4), (4, 12, 60, 2)])
def test_shapes(pe, dev, L, S, A, q):
    _check(pe, dev, 419, L, S, A, q, "f16", seed=L * 1000 + A)


@pytest.mark.parametrize("nsc", [1, 7, 1001])
def test_ragged_batches(pe, dev, nsc):
    _check(pe, dev, nsc, 8, 11, 64, 4, "bf16_split", seed=nsc)


def test_binary16_llrs_and_written_coefficients(pe, dev):
    _check(pe, dev, 1584, 8, 11, 64, 4, "f16", llr_f16=True, write_c=True)
    _check(pe, dev, 500, 4, 11, 128, 3, "bf16", llr_f16=True, write_c=True)


def test_biased_coefficients(pe, dev):
    _check(pe, dev, 300, 8, 11, 64, 4, "f16_split", unbiased=False, write_c=True)


@pytest.mark.parametrize("L,A,prec,thr", [(8, 64, "f16", 2.0), (4, 32, "bf16_split", 2.0),
                                          (2, 128, "f16_split", 1.01)])
def test_fp64_fallback(pe, dev, L, A, prec, thr):
    """A condition threshold just above 1 sends nearly every subcarrier to the FP64 solve: those are applied by
    the flagged-subcarrier path from the coefficients the FP64 solve wrote."""
    fl = _check(pe, dev, 613, L, 11, A, 4, prec, thr=thr, write_c=True)
    assert int((fl != 0).sum()) > 500


def test_mixed_fallback(pe, dev):
    """Some subcarriers in FP32, some in FP64, in the same tiles, at the default threshold: every third
    subcarrier gets two nearly parallel layers (condition estimate far above 1e4)."""
    import torch
    H, s2, Y = _inputs(dev, 1584, 8, 11, 64, 77)
    s2 = torch.full_like(s2, 1e-6)
    H[::3, :, 1] = H[::3, :, 0] * (1 + 1e-4)
    ref, C, ni, fl = _separate(pe, H, s2, Y, 4, "f16")
    llr, C2, ni2, fl2 = pe.mmse_apply_demap(H, s2, Y, 4, prec="f16", write_c=True)
    n = int((fl != 0).sum())
    print("flagged", n)
    assert 0 < n < 1584
    assert torch.equal(llr, ref) and torch.equal(ni2, ni) and torch.equal(fl2, fl) and torch.equal(C2, C)


def test_chain_against_fp64_oracle(po, pe, dev):
    """Headline chain at 20 dB through the one-pass entry point: hard decisions identical to the FP64 oracle
    chain except where |LLR_fp64| < 1e-2 median |LLR_fp64| (the chain bar)."""
    import torch
    rng = np.random.default_rng(1020)
    inp = po.chain_inputs(rng, 1584, 64, 8, 11, 4, 20.0)
    Cr, nir, _, _, bad = po.mmse_f64(inp["H"], inp["sigma2"], unbiased=True)
    assert bad == 0
    llr_r = po.llr_demap_f64(po.coef_apply_f64(Cr, inp["Y"]), nir, 4)
    H = torch.from_numpy(np.ascontiguousarray(inp["H"], dtype=np.complex64)).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(inp["Y"], dtype=np.complex64)).to(dev)
    s2 = torch.from_numpy(np.ascontiguousarray(inp["sigma2"], dtype=np.float32)).to(dev)
    for prec in ("f16", "f16_split", "bf16_split"):
        llr, _, _, fl = pe.mmse_apply_demap(H, s2, Y, 4, prec=prec)
        llr = llr.cpu().numpy().astype(np.float64)
        eps = 1e-2 * np.median(np.abs(llr_r))
        flips = ((llr < 0) != (llr_r < 0)) & (np.abs(llr_r) >= eps)
        assert np.all(fl.cpu().numpy() == 0)
        assert not flips.any(), (prec, int(flips.sum()))
