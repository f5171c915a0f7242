"""The multi-cell configurations at their real workloads, against the oracle's FP64 chain.

Config 4, one slot of 24 uneven cells: the population of bench.py (bench.draw_shapes: PRB {24,52,106,273},
antennas {16,32,64}, layers {2,4,8}, QAM {16,64,256}, SNR 5-25 dB, cell c seeded 42 + 1000 c as in
multicell_evidence.cu:110-132), 25,800 subcarriers in all. The slot plan runs eagerly, as graph replays and with
solve records; each cell's LLRs and noise_inv must be
  (a) bit-identical to the cell run alone through the single-cell entry points (pe_gram_mmse +
      pe_coef_apply_demap), and
  (b) against the FP64 chain (FP64 MMSE, FP64 product, FP64 folding demapper): coefficients within 1e-4
      (per-subcarrier norm-wise relative error, FP32-pass subcarriers), apply EVM <= -60 dB for the binary16
      and split products (same coefficients, vs the FP64 product), and hard decisions identical except where
      |LLR_fp64| < eps = 1e-2 median |LLR_fp64|.

Config 5, the serving loop over 64 such cells (seed 42 + 1000 c) at about half of its measured capacity with a
0.5 ms deadline (the deadline check of multicell_evidence.cu:497-566): every job carries its own samples Y (a
fresh QAM block and noise through its cell's channel), so a job whose LLRs landed in another job's buffer would
fail; every served job's LLRs equal its cell's coefficients applied to its own samples alone, bit for bit, and a
sample of jobs (the largest cells and random others) passes the FP64 checks of (b). Shed jobs leave their
destination untouched."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
S = 11
EVM_BAR_DB = -60.0
COEF_BAR = 1e-4
EPS_REL = 1e-2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def t_c64(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)


def t_f32(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def fp64_checks(po, pe, cell, C_gpu, llr_gpu, Y, H, sigma2, prec):
    """(b): FP64 MMSE / product / demap of (H, sigma2, Y) against the GPU coefficients (complex64 numpy, None to
    skip that check) and LLRs (numpy). Returns a one-line summary."""
    q = cell["q"]
    Cr, nir, _, cond, bad = po.mmse_f64(H, sigma2, unbiased=True)
    assert bad == 0
    llr_r = po.llr_demap_f64(po.coef_apply_f64(Cr, Y), nir, q)
    eps = EPS_REL * np.median(np.abs(llr_r))
    flips = int((((llr_gpu < 0) != (llr_r < 0)) & (np.abs(llr_r) >= eps)).sum())
    msg = "A=%d L=%d q=%d nsc=%d max cond %.0f flips %d" % (cell["A"], cell["L"], q, cell["nsc"], cond.max(), flips)
    if C_gpu is not None:
        well = cond <= 1e4
        err = np.max(np.abs(C_gpu - Cr), axis=(1, 2)) / np.max(np.abs(Cr), axis=(1, 2))
        import torch
        dev = torch.device("cuda:0")
        yeq = pe.coef_apply(t_c64(C_gpu, dev), t_c64(Y, dev), prec=prec).cpu().numpy()
        _, db = po.evm(po.coef_apply_f64(C_gpu, Y), yeq)
        msg += " coef err %.2e apply EVM %.1f dB" % (err[well].max(), db)
        assert np.all(err[well] <= COEF_BAR), msg
        assert db <= EVM_BAR_DB, msg
    assert flips == 0, msg
    return msg


# ---------------------------------------------------------------------------------------------------------
# config 4: one slot of 24 cells
# ---------------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module", params=["f16", "mixed"])
def slot(request, po, pe, dev):
    """The 24-cell population with its inputs on the device and every cell's single-cell results. "mixed": every
    third cell f16_split, every third bf16_split, the others f16 (one apply launch per precision)."""
    import bench
    import torch
    mix = ("f16", "f16_split", "bf16_split")
    cells = []
    for i, sh in enumerate(bench.draw_shapes(24, 42)):
        inp = po.chain_inputs(sh["rng"], sh["nsc"], sh["A"], sh["L"], S, sh["q"], sh["snr"])
        prec = mix[i % 3] if request.param == "mixed" else "f16"
        c = dict(sh, prec=prec, H_np=inp["H"], Y_np=inp["Y"], s2_np=inp["sigma2"])
        c["H"], c["Y"], c["s2"] = t_c64(inp["H"], dev), t_c64(inp["Y"], dev), t_f32(inp["sigma2"], dev)
        C1, ni1, fl1 = pe.gram_mmse(c["H"], c["s2"], unbiased=True)
        c["C_alone"], c["ni_alone"] = C1, ni1
        c["llr_alone"] = pe.coef_apply_demap(C1, c["Y"], ni1, sh["q"], prec=prec)
        nsc, L, A, q = sh["nsc"], sh["L"], sh["A"], sh["q"]
        c["C"] = torch.zeros((nsc, L, A), dtype=torch.complex64, device=dev)
        c["ni"] = torch.zeros((nsc, L), dtype=torch.float32, device=dev)
        c["llr"] = torch.zeros((nsc, L, S, 2 * q), dtype=torch.float32, device=dev)
        cells.append(c)
    torch.cuda.synchronize()
    _say("config 4 population (%s): %d cells, %d subcarriers" % (request.param, len(cells),
                                                                   sum(c["nsc"] for c in cells)))
    return cells


def _descs(pe, cells):
    return [pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], c["prec"], c["H"], c["s2"]) for c in cells]


def _assert_alone(cells, what):
    import torch
    for i, c in enumerate(cells):
        assert torch.equal(c["llr"], c["llr_alone"]), "%s: cell %d LLRs differ from the cell alone" % (what, i)
        assert torch.equal(c["ni"], c["ni_alone"]), "%s: cell %d noise_inv differs from the cell alone" % (what, i)


@pytest.mark.parametrize("mode", ["eager", "graph", "records_eager", "records_graph"])
def test_config4_slot_equals_cells_alone(pe, dev, slot, mode):
    import torch
    _say("config 4 slot, mode %s" % mode)
    plan = pe.SlotPlan(_descs(pe, slot), solve_records=mode.startswith("records"))
    for rep in range(3):  # eager work between launches; every launch rewrites every output
        for c in slot:
            c["llr"].fill_(float(rep + 1))
            c["ni"].zero_()
        plan.graph_launch() if mode.endswith("graph") else plan.launch()
        torch.cuda.synchronize()
        _assert_alone(slot, "%s launch %d" % (mode, rep))
    if not mode.startswith("records"):  # the plan's coefficients are the cell-alone coefficients
        for i, c in enumerate(slot):
            assert torch.equal(c["C"], c["C_alone"]), "cell %d coefficients" % i
    plan.close()


def test_config4_slot_against_fp64_chain(po, pe, dev, slot):
    """(b) for every cell of the slot: the plan's coefficients, its apply precision and its LLRs against FP64."""
    import torch
    plan = pe.SlotPlan(_descs(pe, slot))
    plan.launch()
    torch.cuda.synchronize()
    for i, c in enumerate(slot):
        msg = fp64_checks(po, pe, c, c["C"].cpu().numpy(), c["llr"].cpu().numpy().astype(np.float64), c["Y_np"],
                          c["H_np"], c["s2_np"], c["prec"])
        _say("cell %2d %s: %s" % (i, c["prec"], msg))
    plan.close()


# ---------------------------------------------------------------------------------------------------------
# config 5: serving loop over 64 cells, each job with its own samples
# ---------------------------------------------------------------------------------------------------------
N_JOBS = 320


@pytest.fixture(scope="module")
def serving(po, pe, dev):
    """64 cells (H, sigma2 and the coefficients of the cell alone), N_JOBS jobs with their cells drawn uniformly
    and their own samples Y = H x + n (fresh QAM symbols and noise, drawn on the device), and each job's
    single-cell reference LLRs."""
    import bench
    import torch
    cells = []
    for sh in bench.draw_shapes(64, 42):
        H = po.rayleigh_h(sh["rng"], sh["nsc"], sh["A"], sh["L"])
        s2 = np.full(sh["nsc"], 10.0 ** (-sh["snr"] / 10.0), np.float32)
        c = dict(sh, H_np=H, s2_np=s2, H=t_c64(H, dev), s2=t_f32(s2, dev))
        c["C_alone"], c["ni_alone"], _ = pe.gram_mmse(c["H"], c["s2"], unbiased=True)
        cells.append(c)
    rng = np.random.default_rng(7)
    job_cells = rng.integers(0, len(cells), N_JOBS)
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    Ys, refs = [], []
    for j in range(N_JOBS):
        c = cells[job_cells[j]]
        bits = torch.randint(0, 2, (c["nsc"], c["L"], S, 2 * c["q"]), device=dev, generator=g, dtype=torch.uint8)
        Y = torch.bmm(c["H"], bench.qam_map_torch(bits, c["q"]).contiguous())
        Y = Y + torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=g) * float(c["s2_np"][0]) ** 0.5
        Y = Y.transpose(1, 2).contiguous()
        Ys.append(Y)
        refs.append(pe.coef_apply_demap(c["C_alone"], Y, c["ni_alone"], c["q"], prec="f16"))
    torch.cuda.synchronize()
    # the cells' own samples are never read (every job brings its own): a zero block each
    for c in cells:
        c["Y0"] = torch.zeros((c["nsc"], S, c["A"]), dtype=torch.complex64, device=dev)
    descs = [pe.job_desc(c["H"], c["s2"], c["Y0"], c["q"], "f16") for c in cells]
    _say("config 5 population: %d cells, %d jobs" % (len(cells), N_JOBS))
    return dict(cells=cells, job_cells=job_cells, Ys=Ys, refs=refs, descs=descs)


def _serve(pe, sv, rate, **kw):
    import torch
    outs = [torch.full((r.numel(),), -7.0, dtype=torch.float32, device=r.device) for r in sv["refs"]]
    rep, lat, cell = pe.serve_run(sv["descs"], N_JOBS, rate, deadline_s=0.5e-3, hbm_budget=2 << 30, max_batch=64,
                                  max_inflight=3, seed=5, job_llr=outs, job_cells=sv["job_cells"],
                                  job_H=[sv["cells"][c]["H"] for c in sv["job_cells"]],
                                  job_sigma2=[sv["cells"][c]["s2"] for c in sv["job_cells"]], job_Y=sv["Ys"], **kw)
    torch.cuda.synchronize()
    assert np.array_equal(cell, sv["job_cells"])
    return rep, lat, outs


@pytest.mark.parametrize("graph,shed", [(False, 0), (True, 0), (False, 2)])
def test_config5_serving_jobs_equal_cells_alone(pe, dev, serving, graph, shed):
    import torch
    _say("config 5 serving, graph=%s shed_late=%d" % (graph, shed))
    cap, _, _ = _serve(pe, serving, 0.0, use_graph=graph)  # every job at t = 0: capacity
    rate = 0.5 * cap["jobs_per_s"]
    rep, lat, outs = _serve(pe, serving, rate, use_graph=graph, shed_late=shed)
    _say("capacity %.0f jobs/s; at %.0f jobs/s: p50 %.3f ms p99 %.3f ms, misses %.2f%%, shed %d, batches %d" % (
        cap["jobs_per_s"], rate, rep["p50_s"] * 1e3, rep["p99_s"] * 1e3, 100 * rep["miss_rate"], rep["n_shed"],
        rep["n_batches"]))
    assert cap["n_done"] == N_JOBS
    assert rep["n_done"] + rep["n_shed"] == N_JOBS
    if shed == 0:
        assert rep["n_done"] == N_JOBS and np.all(lat > 0)
    for j in range(N_JOBS):
        if lat[j] > 0:
            assert torch.equal(outs[j], serving["refs"][j].reshape(-1)), "job %d (cell %d)" % (
                j, serving["job_cells"][j])
        else:  # shed: nothing written
            assert bool((outs[j] == -7.0).all()), "shed job %d was written" % j


def test_config5_sampled_jobs_against_fp64_chain(po, pe, dev, serving):
    """(b) for a sample of served jobs: the two largest cells' first jobs and four random others."""
    import torch
    rep, lat, outs = _serve(pe, serving, 0.0, use_graph=False)
    assert rep["n_done"] == N_JOBS
    cells, jc = serving["cells"], serving["job_cells"]
    size = [cells[c]["nsc"] * cells[c]["A"] * cells[c]["L"] for c in jc]
    order = list(np.argsort(size)[::-1])
    sample = [order[0]]
    for j in order:
        if jc[j] != jc[sample[0]]:
            sample.append(j)
            break
    sample += list(np.random.default_rng(3).choice(N_JOBS, 4, replace=False))
    for j in sample:
        c = cells[jc[j]]
        Y = serving["Ys"][j].cpu().numpy()
        llr = outs[j].reshape(serving["refs"][j].shape).cpu().numpy().astype(np.float64)
        msg = fp64_checks(po, pe, c, c["C_alone"].cpu().numpy(), llr, Y, c["H_np"], c["s2_np"], "f16")
        _say("job %d cell %d: %s" % (j, jc[j], msg))
