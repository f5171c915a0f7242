"""Serving loop: jobs from several cells arrive, are admitted by device-memory budget and run in overlapping
batches. Every job's LLRs must be identical, bit for bit, to its cell run alone through the single-cell entry
points, whatever batch and stream it landed in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(2, 16, 2, 2, "f16"), (4, 32, 4, 3, "f16"), (9, 64, 8, 4, "f16"), (3, 64, 4, 4, "f16_split"),
          (5, 16, 8, 4, "f16")]


def population(po, pe, dev, seed):
    import torch
    rng = np.random.default_rng(seed)
    cells, refs = [], []
    for (prb, A, L, q, prec) in SHAPES:
        nsc = 12 * prb
        inp = po.chain_inputs(rng, nsc, A, L, 11, q, float(rng.uniform(5, 25)))
        t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        H, s2, Y = t(inp["H"]), t(inp["sigma2"], np.float32), t(inp["Y"])
        C, ni, _ = pe.gram_mmse(H, s2, unbiased=True)
        refs.append(pe.coef_apply_demap(C, Y, ni, q, prec=prec).cpu().numpy())
        cells.append((pe.job_desc(H, s2, Y, q, prec), (H, s2, Y), refs[-1].shape))
    return cells, refs


@pytest.mark.parametrize("rate,max_batch,inflight,graph", [(0.0, 8, 3, True), (0.0, 3, 1, False),
                                                           (40000.0, 16, 2, True)])
def test_serve_results_equal_cells_alone(po, pe, dev, rate, max_batch, inflight, graph):
    import torch
    cells, refs = population(po, pe, dev, 21)
    n_jobs = 40
    size = max(int(np.prod(r.shape)) for r in refs)
    outs = [torch.full((size,), -7.0, dtype=torch.float32, device=dev) for _ in range(n_jobs)]
    rep, lat, cell = pe.serve_run([c[0] for c in cells], n_jobs, rate, hbm_budget=64 << 20, max_batch=max_batch,
                                  max_inflight=inflight, use_graph=graph, seed=3, job_llr=outs)
    print(rep)
    assert rep["n_done"] == n_jobs and rep["n_shed"] == 0
    assert np.all(lat > 0)
    assert rep["p50_s"] <= rep["p99_s"] <= rep["max_s"]
    assert rep["n_batches"] >= n_jobs / max_batch
    for j in range(n_jobs):
        r = refs[cell[j]]
        got = outs[j][: r.size].cpu().numpy().reshape(r.shape)
        assert np.array_equal(got, r), "job %d (cell %d)" % (j, cell[j])


def test_serve_budget_limits_inflight_jobs(po, pe, dev):
    """A budget that holds one job of the largest cell still serves every job (admission waits for memory)."""
    cells, refs = population(po, pe, dev, 22)
    largest = max(12 * p * L * A * 8 + 12 * p * L * 4 + 12 * p * L * 11 * 2 * q * 4 + 768
                  for (p, A, L, q, _) in SHAPES)
    rep, lat, cell = pe.serve_run([c[0] for c in cells], 30, 0.0, hbm_budget=largest, max_batch=8,
                                  max_inflight=2, seed=4)
    print(rep)
    assert rep["n_done"] == 30 and np.all(lat > 0)


def test_serve_deadline_aware_admission_under_overload(po, pe, dev):
    """Offered load far above capacity: deadline-aware admission (shed_late=2) sheds jobs that would finish
    late, so the served jobs' latency stays near the deadline, while serving every job (shed_late=0) lets the
    queue and the latency grow. Served jobs' LLRs stay bit-identical to the cells alone."""
    import torch
    cells, refs = population(po, pe, dev, 23)
    descs = [c[0] for c in cells]
    n_jobs = 1200  # at 3x capacity the serve-everything queue grows to several deadlines
    pe.serve_run(descs, 200, 0.0, hbm_budget=256 << 20, max_batch=16, max_inflight=2, use_graph=False, seed=4)  # warm-up
    cap, _, _ = pe.serve_run(descs, n_jobs, 0.0, hbm_budget=256 << 20, max_batch=16, max_inflight=2,
                             use_graph=False, seed=5)
    rate = 3.0 * cap["jobs_per_s"]
    deadline = 0.25e-3
    base, _, _ = pe.serve_run(descs, n_jobs, rate, deadline_s=deadline, hbm_budget=256 << 20, max_batch=16,
                              max_inflight=2, use_graph=False, seed=6)
    size = max(int(np.prod(r.shape)) for r in refs)
    outs = [torch.full((size,), -7.0, dtype=torch.float32, device=dev) for _ in range(n_jobs)]
    rep, lat, cell = pe.serve_run(descs, n_jobs, rate, deadline_s=deadline, hbm_budget=256 << 20, max_batch=16,
                                  max_inflight=2, use_graph=False, seed=6, shed_late=2, job_llr=outs)
    print("capacity", cap["jobs_per_s"], "serve-all", base, "deadline-aware", rep)
    assert rep["iso_job_s"] > 0 and rep["contention_index"] > 0
    assert rep["n_done"] + rep["n_shed"] == n_jobs and rep["n_shed"] > 0 and rep["n_done"] > 0
    # relative behaviour only (absolute latencies depend on host scheduling and are reported by bench.py): at 3x
    # capacity the serve-everything queue grows, so most of its jobs miss; admission serves with lower latency
    assert base["n_shed"] == 0 and base["miss_rate"] > 0.5
    assert rep["p99_s"] < base["p99_s"] and rep["p50_s"] < base["p50_s"]
    for j in np.nonzero(lat > 0)[0]:
        r = refs[cell[j]]
        got = outs[j][: r.size].cpu().numpy().reshape(r.shape)
        assert np.array_equal(got, r), "job %d (cell %d)" % (j, cell[j])


def test_serve_binary16_coefficient_cells(po, pe, dev):
    """Cells of precision f16_c16 in the serving loop (their jobs' coefficients take half the arena bytes):
    every job's LLRs equal pe_gram_mmse_c16 + pe_coef_apply_demap_c16 of its cell alone."""
    import torch
    rng = np.random.default_rng(31)
    descs, refs = [], []
    keep = []
    for (prb, A, L, q, _) in SHAPES:
        nsc = 12 * prb
        inp = po.chain_inputs(rng, nsc, A, L, 11, q, float(rng.uniform(5, 25)))
        t = lambda a, dt=np.complex64: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        H, s2, Y = t(inp["H"]), t(inp["sigma2"], np.float32), t(inp["Y"])
        C16, e, ni, _ = pe.gram_mmse_c16(H, s2, unbiased=True)
        refs.append(pe.coef_apply_demap_c16(C16, e, Y, ni, q).cpu().numpy())
        descs.append(pe.job_desc(H, s2, Y, q, "f16_c16"))
        keep.append((H, s2, Y))
    n_jobs = 40
    size = max(int(np.prod(r.shape)) for r in refs)
    outs = [torch.full((size,), -7.0, dtype=torch.float32, device=dev) for _ in range(n_jobs)]
    rep, lat, cell = pe.serve_run(descs, n_jobs, 30000.0, hbm_budget=64 << 20, max_batch=8, max_inflight=2,
                                  use_graph=False, seed=8, job_llr=outs)
    assert rep["n_done"] == n_jobs and np.all(lat > 0)
    for j in range(n_jobs):
        r = refs[cell[j]]
        assert np.array_equal(outs[j][: r.size].cpu().numpy().reshape(r.shape), r), "job %d" % j
