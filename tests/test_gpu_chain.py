"""The single-cell chain at the headline shape (1584 subcarriers, 8 layers, 11 symbols, 64 antennas,
256-QAM): MMSE coefficients -> coefficient apply -> max-log LLR on the GPU, against the FP64 chain of the
oracle (FP64 MMSE, FP64 product, FP64 folding demapper).

Bars (stated here, checked below):
  * EVM of the equalized symbols vs FP64 <= -60 dB for every precision except plain bf16;
  * hard decisions (LLR signs) identical to FP64 except where |LLR_fp64| < eps, eps = 1e-2 * median |LLR_fp64|;
  * the fused kernel's LLRs are the demapper applied to the same equalized symbols (bitwise, tested in
    test_gpu_kernels.py), so the EVM of coef_apply is the EVM of the fused path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EVM_BAR_DB = -60.0
EPS_REL = 1e-2


def dev_c64(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)


def dev_f32(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


@pytest.fixture(scope="module", params=[10.0, 20.0, 30.0])
def chain(request, po):
    snr = request.param
    rng = np.random.default_rng(1000 + int(snr))
    inp = po.chain_inputs(rng, 1584, 64, 8, 11, 4, snr)
    Cr, nir, _, _, bad = po.mmse_f64(inp["H"], inp["sigma2"], unbiased=True)
    assert bad == 0
    yeq_r = po.coef_apply_f64(Cr, inp["Y"])
    llr_r = po.llr_demap_f64(yeq_r, nir, 4)
    return snr, inp, yeq_r, llr_r


@pytest.mark.parametrize("prec", ["f32", "f16", "f16_split", "bf16_split", "bf16"])
def test_chain_256qam(po, pe, dev, chain, prec):
    snr, inp, yeq_r, llr_r = chain
    dH, dY, ds2 = dev_c64(inp["H"], dev), dev_c64(inp["Y"], dev), dev_f32(inp["sigma2"], dev)
    C, ni, fl = pe.gram_mmse(dH, ds2, unbiased=True)
    yeq = pe.coef_apply(C, dY, prec=prec).cpu().numpy()
    llr = pe.coef_apply_demap(C, dY, ni, 4, prec=prec).cpu().numpy().astype(np.float64)
    _, db = po.evm(yeq_r, yeq)
    eps = EPS_REL * np.median(np.abs(llr_r))
    flips = ((llr < 0) != (llr_r < 0)) & (np.abs(llr_r) >= eps)
    ber = np.mean((llr_r < 0) != inp["bits"].astype(bool))
    print("snr=%g prec=%s EVM=%.1f dB eps=%.3g flips=%d (fp64 BER vs tx %.2e)" % (snr, prec, db, eps,
                                                                                   flips.sum(), ber))
    assert np.all(fl.cpu().numpy() == 0)
    if prec == "bf16":
        assert db <= -40.0  # reported, below the chain bar by design: use bf16_split for the chain
        return
    assert db <= EVM_BAR_DB
    assert not flips.any()
