"""pe_mmse_apply_demap_noalloc (the one-pass chain with its workspace in the library's fixed device scratch,
for callers that cannot allocate) against pe_mmse_apply_demap with a caller workspace: LLRs, noise_inv and
sc_flags bit-identical, on calls in a row, each smaller than the one before. The scratch keeps the previous
call's contents, so a later smaller call that read stale workspace bytes would differ."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(po, dev, seed, nsc, A, L, q, snr, ill_every=0):
    import torch
    rng = np.random.default_rng(seed)
    inp = po.chain_inputs(rng, nsc, A, L, 11, q, snr)
    H = inp["H"].copy()
    s2 = inp["sigma2"].copy()
    if ill_every:  # nearly parallel layers at low noise: these subcarriers are solved in FP64
        H[::ill_every, :, 1] = H[::ill_every, :, 0] * (1 + 1e-4)
        s2[::ill_every] = 1e-6
    t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    return t(H), t(s2, np.float32), t(inp["Y"])


# (seed, nSC, nAnt, nLayer, nPamBits, prec, binary16 LLRs, every n-th subcarrier ill-conditioned)
CALLS = [
    (1, 3276, 128, 8, 4, "f16", False, 0),        # 26.8 MB of workspace: the largest shape, coefficient path
    (2, 1584, 64, 8, 4, "f16", False, 5),         # solve-record path with FP64 subcarriers
    (3, 900, 32, 4, 3, "bf16_split", True, 7),    # records, binary16 LLRs
    (4, 300, 16, 2, 2, "f16_split", False, 0),    # two kernels with C in the scratch
]


def test_noalloc_equals_caller_workspace_on_calls_in_a_row(po, pe, dev):
    import torch
    for i, (seed, nsc, A, L, q, prec, f16, ill) in enumerate(CALLS):
        print("call %d: nSC=%d A=%d L=%d q=%d %s llr_f16=%s" % (i, nsc, A, L, q, prec, f16), flush=True)
        assert nsc * L * A * 8 <= pe.NOALLOC_SCRATCH_BYTES
        H, s2, Y = _inputs(po, dev, seed, nsc, A, L, q, 20.0, ill)
        ref, _, ref_ni, ref_fl = pe.mmse_apply_demap(H, s2, Y, q, prec=prec, llr_f16=f16)
        torch.cuda.synchronize()
        got, ni, fl = pe.mmse_apply_demap_noalloc(H, s2, Y, q, prec=prec, llr_f16=f16)
        torch.cuda.synchronize()
        if ill:
            assert int((ref_fl != 0).sum()) > 0
        assert torch.equal(got, ref), "call %d LLRs" % i
        assert torch.equal(ni, ref_ni), "call %d noise_inv" % i
        assert torch.equal(fl, ref_fl), "call %d flags" % i


def test_noalloc_rejects_work_beyond_the_scratch(pe, dev):
    import torch
    nsc, A, L = 4097, 128, 8  # 33.6 MB > 32 MiB
    H = torch.zeros((nsc, A, L), dtype=torch.complex64, device=dev)
    s2 = torch.ones(nsc, dtype=torch.float32, device=dev)
    Y = torch.zeros((nsc, 11, A), dtype=torch.complex64, device=dev)
    with pytest.raises(pe.PeError) as ei:
        pe.mmse_apply_demap_noalloc(H, s2, Y, 4)
    assert ei.value.status == pe.PE_ERR_CAPACITY
