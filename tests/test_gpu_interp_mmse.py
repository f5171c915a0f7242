"""pe_interp_mmse_apply_demap: DMRS estimates to LLRs with the frequency interpolation fused into the MMSE solve.
Its H must be bit-identical to pe_freq_interp's output at the same interpolation precision (re-laid out as
H[sc][ant][layer]), and its LLRs, noise_inv, flags and workspace C bit-identical to pe_mmse_apply_demap on that H
(the Gram matrices are summed in the same order). Covers both interpolation precisions, the MFMA apply
precisions, 32-64 antennas (8 and 16 lanes per subcarrier from the tile table, 4 through a table of the test's
own), pilot counts that are and are not multiples of 16, the full 32-cell batch (792 blocks), binary16 LLRs,
PE_EQ_WRITE_C, the FP64 fallback (forced by a low condition threshold), and H against a complex128 product of the
bf16-rounded operands."""
import os

import pytest

pytestmark = pytest.mark.gpu


def _inputs(dev, nblk, A, K, S, seed):
    import torch
    g = torch.Generator(device=dev).manual_seed(seed)
    W = torch.randn((64, K), dtype=torch.complex64, device=dev, generator=g) / K ** 0.5
    X = torch.randn((nblk, A, K, 4), dtype=torch.complex64, device=dev, generator=g) / A ** 0.5
    s2 = torch.rand((64 * nblk,), device=dev, generator=g) * 0.05 + 0.005
    Y = torch.randn((64 * nblk, S, A), dtype=torch.complex64, device=dev, generator=g)
    return W, X, s2, Y


def _check(pe, dev, nblk, A, K, S, q, interp, prec, thr=0.0, llr_f16=False, write_c=False, seed=1):
    import torch
    W, X, s2, Y = _inputs(dev, nblk, A, K, S, seed)
    nsc = 64 * nblk
    print("nblk=%d A=%d K=%d S=%d q=%d interp=%s prec=%s thr=%g f16=%s write_c=%s" % (
        nblk, A, K, S, q, interp, prec, thr, llr_f16, write_c), flush=True)
    C = torch.zeros((nsc, 4, A), dtype=torch.complex64, device=dev)
    llr, H, C, ni, fl = pe.interp_mmse_apply_demap(W, X, s2, Y, q, interp_prec=interp, prec=prec,
                                                   cond_threshold=thr, write_c=write_c, llr_f16=llr_f16, C=C)
    Hf = pe.freq_interp(W, X.reshape(nblk * A, K, 4), prec=interp)  # [blk * A + ant][m][layer]
    Href = Hf.reshape(nblk, A, 64, 4).permute(0, 2, 1, 3).reshape(nsc, A, 4).contiguous()
    assert torch.equal(H, Href)
    C2 = torch.zeros_like(C)
    llr2, C2, ni2, fl2 = pe.mmse_apply_demap(Href, s2, Y, q, prec=prec, cond_threshold=thr, write_c=write_c,
                                             llr_f16=llr_f16, C=C2)
    print("flagged %d" % int((fl2 != 0).sum()), flush=True)
    assert torch.equal(fl, fl2)
    assert torch.equal(ni, ni2)
    assert torch.equal(C, C2)
    assert torch.equal(llr, llr2)
    return H, fl


@pytest.mark.parametrize("nblk,A,K,S,q,interp,prec", [
    (3, 64, 96, 11, 4, "bf16", "f16"),        # the reference's [64 x 96] @ [96 x 4]; 16 lanes
    (5, 32, 96, 11, 2, "bf16", "f16"),        # 16 lanes (table entry for 32 antennas)
    (4, 48, 40, 7, 3, "bf16_split", "bf16"),  # 8 lanes (default rule); 40 pilots: a partial 16-pilot step
    (2, 36, 128, 14, 4, "bf16", "f16_split"),  # 9 tiles of 4 antennas; the largest pilot count
    (1, 64, 17, 1, 1, "bf16_split", "bf16_split"),
    (37, 64, 96, 11, 4, "bf16", "f16"),
    (792, 64, 96, 11, 4, "bf16", "f16"),      # the 32-cell batch: 50,688 subcarriers, 792 workgroups
])
def test_fused_equals_interp_then_mmse_apply(pe, dev, nblk, A, K, S, q, interp, prec):
    _check(pe, dev, nblk, A, K, S, q, interp, prec)


def test_binary16_llrs_and_coefficients_written(pe, dev):
    _check(pe, dev, 6, 64, 96, 11, 4, "bf16", "f16", llr_f16=True, write_c=True)


@pytest.mark.parametrize("A", [32, 64])
def test_fp64_fallback(pe, dev, A):
    """A condition threshold of 1 flags every subcarrier: all are solved in FP64 from H read back."""
    _, fl = _check(pe, dev, 3, A, 96, 11, 4, "bf16", "f16", thr=1.0)
    assert int((fl != 0).sum()) == fl.numel()


def test_four_lanes_per_subcarrier(pe, dev, tmp_path):
    """P = 4 through a table line of the test's own (every entry point reads the same table)."""
    here = os.path.dirname(os.path.abspath(pe.__file__))
    table = os.path.join(here, "tune_table.txt")
    custom = tmp_path / "t.txt"
    custom.write_text("mmse 4 40 -1 4\n" + open(table).read())
    assert pe.lib().pe_tune_load(str(custom).encode()) > 0
    try:
        _check(pe, dev, 3, 40, 96, 11, 4, "bf16", "f16")
    finally:
        assert pe.lib().pe_tune_load(table.encode()) > 0


def test_h_against_complex128_product(pe, dev):
    """H is W @ X of the bf16-rounded operands, accumulated in FP32: within the FP32 accumulation bound of a
    complex128 product of those operands."""
    import torch
    W, X, s2, Y = _inputs(dev, 2, 64, 96, 11, 5)
    _, H, _, _, _ = pe.interp_mmse_apply_demap(W, X, s2, Y, 4)
    rb = lambda t: torch.complex(t.real.to(torch.bfloat16).double(), t.imag.to(torch.bfloat16).double()).cpu()
    Wb, Xb = rb(W), rb(X)
    ref = torch.einsum("mk,bakn->bman", Wb, Xb).reshape(128, 64, 4)
    # each real part is a sum of 2 K exact products (bf16 x bf16 fits FP32) rounded at most 2 K times in FP32
    mag = torch.einsum("mk,bakn->bman", Wb.real.abs() + Wb.imag.abs(), Xb.real.abs() + Xb.imag.abs())
    bound = 2 * 96 * 2.0 ** -24 * mag.reshape(128, 64, 4)
    d = H.cpu().to(torch.complex128) - ref
    assert bool((d.real.abs() <= bound).all()) and bool((d.imag.abs() <= bound).all())
    assert (d.abs().max() / ref.abs().max()).item() < 1e-5
