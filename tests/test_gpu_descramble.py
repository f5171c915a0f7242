"""Descrambling and the rate-recovery layout of the LLRs on the GPU, against the oracle (TS 38.211 5.2.1 and 6.3.1,
TS 38.212 5.4.2, restated in oracle/pe_oracle.c and pinned in tests/test_oracle.py).

Bars (bit-exact; these are bit and index operations):
  * pe_gold_sequence words = the oracle's bit-serial c(n), packed, for several c_init and lengths (partial last
    word zero-filled);
  * pe_coef_apply_demap_cw = the oracle's layout (descrambling sign flips, codeword order or per-block
    de-interleaving) of the LLRs pe_coef_apply_demap / pe_coef_apply_demap_f16 write for the same inputs, bit for
    bit, on single- and multi-codeword layer maps, ragged code blocks and every QAM order;
  * end to end: codeword bits scrambled, modulated, layer-mapped and placed frequency first, sent through a
    Rayleigh channel at 30 dB; MMSE + the codeword apply give back the codeword bits as LLR signs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def packed(bits):
    """uint8 bits -> uint32 words, bit g % 32 of word g / 32."""
    nw = (bits.size + 31) // 32
    b = np.zeros(nw * 32, np.uint8)
    b[:bits.size] = bits
    return np.packbits(b.reshape(nw, 32)[:, ::-1], axis=1).view(">u4").ravel().astype(np.uint32)


@pytest.mark.parametrize("c_init", [0, 1, (17 << 15) + 3, 0x7FFF_FFFF, 0x1234_5678])
@pytest.mark.parametrize("n_bits", [1, 31, 32, 33, 1600, 16 * 32 * 256 + 17, 2_359_296])
def test_gold_sequence_bitexact(po, pe, dev, c_init, n_bits):
    got = pe.gold_sequence(c_init, n_bits).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, packed(po.gold_sequence(c_init, n_bits)))


def _inputs(po, dev, seed, nsc, A, L, S, q):
    import torch
    rng = np.random.default_rng(seed)
    inp = po.chain_inputs(rng, nsc, A, L, S, q, 20.0)
    t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    return t(inp["H"]), t(inp["sigma2"], np.float32), t(inp["Y"])


# (nSC, L, S, A, q, layer_cw, code blocks per codeword, descramble per codeword, prec, binary16 LLRs)
CASES = [
    (1584, 8, 11, 64, 4, [0] * 8, [0], [True], "f16", False),
    (1584, 8, 11, 64, 4, [0] * 8, [9], [True], "f16", True),
    (300, 8, 11, 32, 2, [0, 0, 1, 1, 2, 2, 2, 3], [0, 3, 1, 5], [True, False, True, True], "bf16_split", False),
    (37, 3, 14, 16, 3, [1, 0, 1], [2, 0], [True, True], "f16_split", True),
    (5, 1, 1, 128, 1, [0], [1], [True], "bf16", False),
    (96, 4, 11, 64, 3, [0, 1, 2, 3], [4, 4, 0, 7], [True, True, True, False], "f16", False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "nsc%d-L%d-q%d-%s%s" % (c[0], c[1], c[4], c[8],
                                                                            "-f16llr" if c[9] else ""))
def test_codeword_layout_bitexact(po, pe, dev, case):
    import torch
    nsc, L, S, A, q, layer_cw, ncb, desc, prec, f16 = case
    H, s2, Y = _inputs(po, dev, nsc + L, nsc, A, L, S, q)
    C, ni, _ = pe.gram_mmse(H, s2, unbiased=True)
    ncw = max(layer_cw) + 1
    nu = [layer_cw.count(k) for k in range(ncw)]
    G = [nsc * S * nu[k] * 2 * q for k in range(ncw)]
    cseq_bits = [po.gold_sequence((100 + k) << 15 | 9, G[k]) if desc[k] else None for k in range(ncw)]
    cseq_dev = [pe.gold_sequence((100 + k) << 15 | 9, G[k]) if desc[k] else None for k in range(ncw)]
    outs = pe.coef_apply_demap_cw(C, Y, ni, q, layer_cw, cseq_dev, ncb, prec=prec, llr_f16=f16)
    if f16:  # values of the binary16 fused path; the layout only negates and moves them
        base = pe.coef_apply_demap_f16(C, Y, ni, q, prec=prec).cpu().numpy().astype(np.float32)
    else:
        base = pe.coef_apply_demap(C, Y, ni, q, prec=prec).cpu().numpy()
    ref = po.cw_layout(base, layer_cw, cseq_bits, ncb)
    torch.cuda.synchronize()
    for k in range(ncw):
        got = outs[k].cpu().numpy()
        want = ref[k].astype(np.float16) if f16 else ref[k]
        assert not np.isnan(ref[k]).any()  # every position written exactly once by the oracle
        assert np.array_equal(got.view(np.uint16 if f16 else np.uint32), want.view(np.uint16 if f16 else np.uint32)), \
            "codeword %d" % k


def test_end_to_end_hard_decisions_are_the_codeword_bits(po, pe, dev):
    """Transmitter per TS 38.211 6.3.1 (scrambling, 16-QAM mapping, layer mapping of three codewords over four
    layers, frequency-first resource mapping), Rayleigh channel with 32 antennas at 30 dB; the GPU MMSE and the
    codeword apply give LLRs whose signs are the codeword bits, in codeword order and, de-interleaved per block,
    the de-interleaved bits."""
    import torch
    rng = np.random.default_rng(5)
    nsc, S, A, q = 120, 11, 32, 2
    layer_cw = [0, 1, 1, 2]
    L, ncw, qm = len(layer_cw), 3, 2 * q
    nu = [layer_cw.count(k) for k in range(ncw)]
    x = np.zeros((nsc, L, S), np.complex128)
    bits, c_inits = [], [(4242 + k) << 15 | 1 for k in range(ncw)]
    for k in range(ncw):
        G = nsc * S * nu[k] * qm
        b = rng.integers(0, 2, G, dtype=np.uint8)
        sym = po.qam_map((b ^ po.gold_sequence(c_inits[k], G)).reshape(-1, qm), q)
        layers = [l for l in range(L) if layer_cw[l] == k]
        m = np.arange(sym.size)
        i = m // nu[k]
        for v, l in enumerate(layers):
            sel = (m % nu[k]) == v
            x[i[sel] % nsc, l, i[sel] // nsc] = sym[sel]
        bits.append(b)
    H = po.rayleigh_h(rng, nsc, A, L)
    s2 = 10.0 ** (-30.0 / 10.0)
    Yv = np.einsum("nal,nls->nsa", H.astype(np.complex128), x) + po.cn(rng, (nsc, S, A), s2)
    t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    C, ni, _ = pe.gram_mmse(t(H), t(np.full(nsc, s2), np.float32), unbiased=True)
    cs = [pe.gold_sequence(c_inits[k], bits[k].size) for k in range(ncw)]
    outs = pe.coef_apply_demap_cw(C, t(Yv), ni, q, layer_cw, cs, [0, 0, 0])
    for k in range(ncw):
        assert np.array_equal((outs[k].cpu().numpy() < 0).astype(np.uint8), bits[k]), "codeword %d" % k
    outs_e = pe.coef_apply_demap_cw(C, t(Yv), ni, q, layer_cw, cs, [3, 2, 5])
    for k, nb in zip(range(ncw), (3, 2, 5)):
        E = po.cw_block_sizes(bits[k].size, nu[k], qm, nb)
        starts = np.concatenate([[0], np.cumsum(E)[:-1]])
        got = (outs_e[k].cpu().numpy() < 0).astype(np.uint8)
        for r in range(nb):  # de-interleaved block r: e(i E/Qm + j) = f(i + j Qm)
            f = bits[k][starts[r]:starts[r] + E[r]]
            e = f.reshape(-1, qm).T.ravel()
            assert np.array_equal(got[starts[r]:starts[r] + E[r]], e), "codeword %d block %d" % (k, r)


def test_codeword_layout_rejects_bad_maps(pe, dev):
    import torch
    C = torch.zeros((4, 2, 16), dtype=torch.complex64, device=dev)
    Y = torch.zeros((4, 11, 16), dtype=torch.complex64, device=dev)
    ni = torch.ones((4, 2), dtype=torch.float32, device=dev)
    with pytest.raises(pe.PeError):   # codeword 0 unused
        pe.coef_apply_demap_cw(C, Y, ni, 2, [1, 1])
    with pytest.raises(pe.PeError):   # more code blocks than symbols per layer (4 x 11)
        pe.coef_apply_demap_cw(C, Y, ni, 2, [0, 0], n_code_blocks=[45])
