"""Multi-cell slot plans: uneven cells in one fixed sequence of tile kernels. Every cell's LLRs must be identical,
bit for bit, to the same cell run alone through the single-cell entry points (pe_gram_mmse +
pe_coef_apply_demap), after eager launches, graph replays and pointer updates alike."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_cells(po, dev, seed, shapes):
    import torch
    rng = np.random.default_rng(seed)
    cells = []
    for (prb, A, L, q, prec) in shapes:
        nsc = 12 * prb
        inp = po.chain_inputs(rng, nsc, A, L, 11, q, float(rng.uniform(5, 25)))
        t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        cells.append(dict(H=t(inp["H"]), s2=t(inp["sigma2"], np.float32), Y=t(inp["Y"]), q=q, prec=prec,
                          C=torch.zeros((nsc, L, A), dtype=torch.complex64, device=dev),
                          ni=torch.zeros((nsc, L), dtype=torch.float32, device=dev),
                          llr=torch.zeros((nsc, L, 11, 2 * q), dtype=torch.float32, device=dev)))
    return cells


def alone(pe, c):
    C, ni, _ = pe.gram_mmse(c["H"], c["s2"], unbiased=True)
    return pe.coef_apply_demap(C, c["Y"], ni, c["q"], prec=c["prec"]).cpu().numpy()


SHAPES = [(2, 16, 2, 2, "f16"), (4, 32, 4, 3, "f16"), (9, 64, 8, 4, "f16"), (3, 64, 4, 4, "f16"),
          (2, 48, 8, 2, "f16"), (5, 16, 8, 4, "f16_split"), (1, 128, 8, 4, "f16"), (6, 32, 2, 4, "bf16_split")]


def test_plan_matches_cells_alone(po, pe, dev):
    import torch
    cells = make_cells(po, dev, 5, SHAPES)
    ref = [alone(pe, c) for c in cells]
    descs = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]
    plan = pe.SlotPlan(descs)
    ntiles, nwg = plan.info()
    assert ntiles > 0 and nwg > 0
    plan.launch()
    torch.cuda.synchronize()
    for c, r in zip(cells, ref):
        assert np.array_equal(c["llr"].cpu().numpy(), r)
    # relaunch: the self-resetting queues serve a second launch
    for c in cells:
        c["llr"].zero_()
    plan.launch()
    torch.cuda.synchronize()
    for c, r in zip(cells, ref):
        assert np.array_equal(c["llr"].cpu().numpy(), r)
    plan.close()


def test_plan_graph_replays_and_updates(po, pe, dev):
    import torch
    cells = make_cells(po, dev, 6, SHAPES[:5])
    ref = [alone(pe, c) for c in cells]
    descs = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]
    plan = pe.SlotPlan(descs)
    for rep in range(4):
        for c in cells:
            c["llr"].fill_(float(rep + 1))  # eager work between replays
        plan.graph_launch()
        torch.cuda.synchronize()
        for c, r in zip(cells, ref):
            assert np.array_equal(c["llr"].cpu().numpy(), r), "replay %d" % rep
    # next slot: same shapes, new data and buffers
    cells2 = make_cells(po, dev, 7, SHAPES[:5])
    ref2 = [alone(pe, c) for c in cells2]
    descs2 = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells2]
    plan.update(descs2)
    plan.graph_launch()
    torch.cuda.synchronize()
    for c, r in zip(cells2, ref2):
        assert np.array_equal(c["llr"].cpu().numpy(), r)
    plan.close()


def test_plan_with_given_coefficients(po, pe, dev):
    """Cells without H: stage 1 is skipped and the caller's C / noise_inv are used."""
    import torch
    cells = make_cells(po, dev, 8, SHAPES[:3])
    for c in cells:
        C, ni, _ = pe.gram_mmse(c["H"], c["s2"], unbiased=True)
        c["C"].copy_(C)
        c["ni"].copy_(ni)
    ref = [pe.coef_apply_demap(c["C"], c["Y"], c["ni"], c["q"], prec=c["prec"]).cpu().numpy() for c in cells]
    plan = pe.SlotPlan([pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"]) for c in cells])
    plan.launch()
    torch.cuda.synchronize()
    for c, r in zip(cells, ref):
        assert np.array_equal(c["llr"].cpu().numpy(), r)
    plan.close()


C16_SHAPES = [(2, 16, 2, 2, "f16_c16"), (4, 32, 4, 3, "f16_c16"), (9, 64, 8, 4, "f16_c16"),
              (1, 128, 8, 4, "f16_c16"), (3, 24, 5, 4, "f16_c16"), (2, 128, 2, 1, "f16_c16")]


def alone_c16(pe, c):
    C16, e, ni, _ = pe.gram_mmse_c16(c["H"], c["s2"], unbiased=True)
    return pe.coef_apply_demap_c16(C16, e, c["Y"], ni, c["q"]).cpu().numpy(), C16.cpu().numpy(), e.cpu().numpy()


@pytest.mark.parametrize("graph", [False, True])
def test_plan_binary16_coefficients_match_cells_alone(po, pe, dev, graph):
    """Cells of precision f16_c16: the plan's MMSE writes binary16 coefficients and exponents into each cell's C
    buffer and its apply reads them; coefficients and LLRs equal pe_gram_mmse_c16 + pe_coef_apply_demap_c16 of
    the cell alone, bit for bit."""
    import torch
    cells = make_cells(po, dev, 8, C16_SHAPES)
    ref = [alone_c16(pe, c) for c in cells]
    descs = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]
    plan = pe.SlotPlan(descs)
    for _ in range(2):
        for c in cells:
            c["llr"].zero_()
        plan.graph_launch() if graph else plan.launch()
        torch.cuda.synchronize()
        for c, (llr, C16, e) in zip(cells, ref):
            c16, ce = pe.c16_view(c["C"])
            assert np.array_equal(c16.cpu().numpy(), C16) and np.array_equal(ce.cpu().numpy(), e)
            assert np.array_equal(c["llr"].cpu().numpy(), llr)
    plan.close()


def test_plan_rejects_mixed_coefficient_formats(po, pe, dev):
    cells = make_cells(po, dev, 9, [(2, 16, 2, 2, "f16_c16"), (2, 16, 2, 2, "f16")])
    descs = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]
    with pytest.raises(pe.PeError):
        pe.SlotPlan(descs)


@pytest.mark.parametrize("graph", [False, True])
def test_plan_solve_records_match_cells_alone(po, pe, dev, graph):
    """PE_SLOT_SOLVE_RECORDS: stage 1 keeps compact solve records in C and stage 2 forms the coefficients on chip.
    LLRs and noise_inv equal the cells alone bit for bit, including subcarriers solved in FP64 (every third
    subcarrier of one cell has two nearly parallel layers) and cells the flag leaves on the coefficient path
    (128 antennas; 8 layers on 4 antennas, whose record does not fit the block), over repeated launches and a
    pointer update."""
    import torch
    shapes = SHAPES + [(3, 4, 8, 2, "f16"), (4, 32, 8, 4, "bf16")]
    cells = make_cells(po, dev, 15, shapes)
    bad = cells[2]
    bad["H"][::3, :, 1] = bad["H"][::3, :, 0] * (1 + 1e-4)
    bad["s2"].fill_(1e-6)
    ref, ref_ni = [], []
    for c in cells:
        C, ni, fl = pe.gram_mmse(c["H"], c["s2"], unbiased=True)
        ref.append(pe.coef_apply_demap(C, c["Y"], ni, c["q"], prec=c["prec"]).cpu().numpy())
        ref_ni.append(ni.cpu().numpy())
        if c is bad:
            assert int((fl != 0).sum()) > 0
    descs = [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]
    plan = pe.SlotPlan(descs, solve_records=True)
    for rep in range(3):
        for c in cells:
            c["llr"].fill_(float(rep + 1))
            c["ni"].zero_()
        plan.graph_launch() if graph else plan.launch()
        torch.cuda.synchronize()
        for i, c in enumerate(cells):
            assert np.array_equal(c["llr"].cpu().numpy(), ref[i]), "cell %d rep %d" % (i, rep)
            assert np.array_equal(c["ni"].cpu().numpy(), ref_ni[i]), "cell %d rep %d" % (i, rep)
    cells2 = make_cells(po, dev, 16, shapes)
    ref2 = [alone(pe, c) for c in cells2]
    plan.update([pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells2])
    plan.graph_launch() if graph else plan.launch()
    torch.cuda.synchronize()
    for c, r in zip(cells2, ref2):
        assert np.array_equal(c["llr"].cpu().numpy(), r)
    plan.close()
