"""Benchmark of the PUSCH equalization hot path on one MI355X.

Workload (the chain at the headline shape): cells of 1584 subcarriers (132 PRB), 8 layers, 11 data symbols,
64 antennas, 256-QAM, Rayleigh CN(0, 1/64) channels at 20 dB SNR. Coefficients come from the library's
own MMSE stage (pe_gram_mmse); one step is one launch of the fused coefficient-apply + 256-QAM demapper
(pe_coef_apply_demap, binary16 MFMA) over a batch of cell-slots resident in HBM. Consecutive steps alternate
between two batch buffer sets of ~0.64 GB each, far beyond the 256 MB Infinity Cache, so every step streams
from HBM.

Printed: one JSON line (metric, value = subcarriers/s, roofline of the fused kernel, CPU baseline, EVM,
single-cell slot latency percentiles, plain-FP32 comparison). Synthetic data, generated on the GPU by torch
(plumbing); every measured operation runs in the HIP library.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "aerial-mimo-tc-optimization_amd")
sys.path.insert(0, PKG)

METRIC = "coef-apply+demap subcarriers/s & HBM GB/s at EVM<=-60dB; p99 slot latency"
HBM_PEAK_GBS = 8000.0
NSC_CELL, L, S, A, Q = 1584, 8, 11, 64, 4
SNR_DB = 20.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def qam_map_torch(bits, q):
    """TS 38.211 section 5.1 mapping of bits [..., 2q] (unit average power)."""
    import torch
    norm = (2.0 * (4.0 ** q - 1.0) / 3.0) ** 0.5
    axes = []
    for iq in range(2):
        c = bits[..., iq::2].to(torch.float32)
        t = 1.0 - 2.0 * c[..., q - 1]
        for i in range(q - 2, -1, -1):
            t = (1.0 - 2.0 * c[..., i]) * (2.0 ** (q - 1 - i) - t)
        axes.append(t / norm)
    return torch.complex(axes[0], axes[1])


def make_cells(n_cells, seed, dev):
    """H [n*1584, 64, 8] CN(0,1/64), x 256-QAM [.., 8, 11], Y = H x + n, sigma2 = 10^(-SNR/10)."""
    import torch
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    nsc = n_cells * NSC_CELL
    H = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A) ** 0.5
    bits = torch.randint(0, 2, (nsc, L, S, 2 * Q), device=dev, generator=g, dtype=torch.uint8)
    x = qam_map_torch(bits, Q)
    s2 = 10.0 ** (-SNR_DB / 10.0)
    Y = torch.bmm(H, x)  # [nsc, A, S]
    Y = Y + torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=g) * s2 ** 0.5
    Y = Y.transpose(1, 2).contiguous()  # [nsc, S, A]
    sigma2 = torch.full((nsc,), s2, dtype=torch.float32, device=dev)
    return H, Y, sigma2


def evm_db(ref, test):
    import torch
    d = (ref - test.to(ref.dtype)).abs().pow(2).sum()
    return float(10.0 * torch.log10(d / ref.abs().pow(2).sum()))


def torch_reference_chain(H, sigma2, Y):
    """FP64 MMSE + FP64 apply with torch on the GPU (for the EVM report only)."""
    import torch
    Hd = H.to(torch.complex128)
    Hh = Hd.conj().transpose(1, 2).contiguous()
    G = torch.bmm(Hh, Hd) + sigma2.to(torch.float64)[:, None, None] * torch.eye(L, dtype=torch.complex128,
                                                                                 device=H.device)
    W = torch.linalg.solve(G, Hh)
    Gi = torch.linalg.inv(G)
    mu = 1.0 - sigma2.to(torch.float64)[:, None] * torch.diagonal(Gi, dim1=1, dim2=2).real
    Cu = W / mu[:, :, None]
    return torch.bmm(Cu, Y.to(torch.complex128).transpose(1, 2).contiguous())


def time_launches(fn, bufs, steps, warmup):
    """Warmup, then `steps` launches alternating over `bufs`, bracketed by events on the current stream."""
    import torch
    for i in range(warmup):
        fn(*bufs[i % len(bufs)])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for i in range(steps):
        fn(*bufs[i % len(bufs)])
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / 1e3, wall


def mmse_stage_times(pe, cells, steps):
    """pe_gram_mmse and pe_gram_mmse_c16 over batches of (H, Y, sigma2), `steps` launches each alternating over
    the batches (beyond the Infinity Cache: every launch reads its H cold). Returns the two elapsed times in s."""
    bufs, bufs16 = [], []
    for H, _, s2 in cells:
        C, ni, fl = pe.gram_mmse(H, s2, unbiased=True)
        bufs.append((H, s2, C, ni, fl))
        bufs16.append((H, s2) + tuple(pe.gram_mmse_c16(H, s2, unbiased=True)))
    el, _ = time_launches(lambda H, s2, C, ni, fl: pe.gram_mmse(H, s2, unbiased=True, C=C, noise_inv=ni, flags=fl),
                          bufs, steps, 4)
    el16, _ = time_launches(lambda H, s2, *o: pe.gram_mmse_c16(H, s2, True, 0.0, *o), bufs16, steps, 4)
    return el, el16


def call_latency(fn, iters, spin):
    """Milliseconds from an event recorded before fn() to one recorded after it, one call at a time on an idle GPU
    (20 untimed calls first), sorted. spin: a spin kernel is queued before the start event, so fn's launches are
    already in the queue when it fires (device time); without it, the host's time to issue fn is included."""
    import torch
    out = []
    for i in range(iters + 20):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if spin:
            torch.cuda._sleep(200000)
        a.record()
        fn()
        b.record()
        b.synchronize()
        if i >= 20:
            out.append(a.elapsed_time(b))
    out.sort()
    return out


def percentile(v, p):
    """Nearest-rank percentile of the sorted list v (None if empty)."""
    return v[min(len(v) - 1, max(0, int(-(-p * len(v) // 100)) - 1))] if v else None


def cpu_model():
    """The host CPU as /proc/cpuinfo names it: model name, family / model / stepping, cache, cores per socket."""
    info = {}
    try:
        for ln in open("/proc/cpuinfo"):
            if not ln.strip():
                break  # first processor only
            k, _, v = ln.partition(":")
            info.setdefault(k.strip(), v.strip())
    except OSError:
        return None
    if "model name" not in info:
        return None
    extra = ["%s %s" % (k, info[k]) for k in ("vendor_id", "cpu family", "model", "stepping", "microcode",
                                               "cache size", "cpu cores", "siblings") if k in info]
    return "%s (%s)" % (info["model name"], ", ".join(extra))


def cpu_baseline(C, Y, ni, seconds):
    """The oracle's OpenMP restatement of the scalar coef-apply loop + FP32 demapper on one cell, repeated
    for about `seconds`, on the host cores (OMP_NUM_THREADS or all), then for a quarter of that on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pe_oracle as po
    Cn, Yn, nin = C.cpu().numpy(), Y.cpu().numpy(), ni.cpu().numpy()
    out = np.empty((NSC_CELL, L, S, 2 * Q), dtype=np.float32)
    bytes_sc = L * A * 8 + S * A * 8 + L * S * 2 * Q * 4 + L * 4

    def run(nthreads, budget):
        _, used = po.coef_apply_demap_f32_omp(Cn, Yn, nin, Q, nthreads, out)  # warm
        reps, t0 = 0, time.perf_counter()
        while True:
            po.coef_apply_demap_f32_omp(Cn, Yn, nin, Q, nthreads, out)
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget:
                break
        return reps * NSC_CELL / el, int(used), reps, el

    v, used, reps, el = run(0, seconds)
    v1, _, reps1, el1 = run(1, max(1.0, seconds / 4))
    return {"value": v, "unit": "subcarriers/s", "cores": used, "kind": "port",
            "sample": "%d repetitions of one 1584-subcarrier cell (8 layers, 11 symbols, 64 antennas, 256-QAM), "
                      "scalar antenna-serial FP32 chain + FP32 demapper, OpenMP over (subcarrier, layer), "
                      "%.1f s; then %d repetitions on 1 thread, %.1f s" % (reps, el, reps1, el1),
            "gbs": v * bytes_sc / 1e9, "single_thread_value": v1, "single_thread_gbs": v1 * bytes_sc / 1e9,
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "build": "oracle/Makefile: gcc -O3 -mfma -mavx2 -ffp-contract=off -fopenmp"}


def hbm_ceiling(dev):
    """Measured streaming ceilings of this GPU: the library's probe kernels (pe_hbm_probe: 16-byte nontemporal
    copy and read over 1 GiB, best grid of a small sweep) and, for comparison, torch's own copy_ and sum."""
    import torch
    import pusch_eq as pe
    n = 1 << 30
    a = torch.empty(n // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    sink = torch.empty(1 << 16, dtype=torch.int32, device=dev)

    def rate(fn, nbytes, reps=10):
        for _ in range(3):
            fn()
        el, _ = time_launches(fn, [()], reps, 0)
        return nbytes * reps / el / 1e9

    best = {"copy_gbs": 0.0, "read_gbs": 0.0}
    for grid in (1024, 2048, 4096, 8192, 16384):
        c = rate(lambda: pe.hbm_probe(a, b, grid=grid), 2 * n)
        r = rate(lambda: pe.hbm_probe(a, sink, read_only=True, grid=grid), n)
        if c > best["copy_gbs"]:
            best["copy_gbs"], best["copy_grid"] = c, grid
        if r > best["read_gbs"]:
            best["read_gbs"], best["read_grid"] = r, grid
    best["torch_copy_gbs"] = rate(lambda: b.copy_(a), 2 * n)
    best["torch_sum_gbs"] = rate(lambda: a.sum(), n)
    best["method"] = ("1 GiB buffers, 10 launches each after 3 warm-up; copy counts bytes read + written; "
                      "probe = pe_hbm_probe (16-byte nontemporal accesses, 256-lane workgroups, grid swept "
                      "1024..16384)")
    del a, b, sink
    torch.cuda.empty_cache()
    return best


def draw_shapes(n, seed):
    """Cell shapes of the multi-cell configurations: PRB in {24,52,106,273}, antennas in {16,32,64}, layers in
    {2,4,8}, QAM in {16,64,256} (nPamBits 2..4), SNR uniform in 5-25 dB, each drawn uniformly, in that order,
    from the cell's own generator seeded seed + 1000 c (the reference's per-cell convention, 42 + 1000 c,
    multicell_evidence.cu:110-132). The generator is returned with the shape, positioned after the draws."""
    import numpy as np
    out = []
    for ci in range(n):
        rng = np.random.default_rng(seed + 1000 * ci)
        prb = int(rng.choice([24, 52, 106, 273]))
        a = int(rng.choice([16, 32, 64]))
        l = int(rng.choice([2, 4, 8]))
        q = int(rng.choice([2, 3, 4]))
        snr = float(rng.uniform(5, 25))
        out.append(dict(prb=prb, nsc=12 * prb, A=a, L=l, q=q, snr=snr, seed=seed + 1000 * ci, rng=rng))
    return out


def draw_cells(n, seed, dev, prec="f16"):
    """Cells of the multi-cell configurations (shapes from draw_shapes), Rayleigh CN(0,1/A) channels, random QAM
    symbols of the cell's order through them plus AWGN at the cell's SNR, 11 data symbols; data drawn on the GPU
    by torch from the cell's seed."""
    import torch
    import pusch_eq as pe
    cells = []
    for sh in draw_shapes(n, seed):
        g = torch.Generator(device=dev)
        g.manual_seed(sh["seed"])
        a, l, q, snr, nsc = sh["A"], sh["L"], sh["q"], sh["snr"], sh["nsc"]
        H = torch.randn((nsc, a, l), dtype=torch.complex64, device=dev, generator=g) * (1.0 / a) ** 0.5
        bits = torch.randint(0, 2, (nsc, l, S, 2 * q), device=dev, generator=g, dtype=torch.uint8)
        s2v = 10.0 ** (-snr / 10.0)
        Y = torch.bmm(H, qam_map_torch(bits, q))
        Y = (Y + torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=g) * s2v ** 0.5)
        Y = Y.transpose(1, 2).contiguous()
        s2 = torch.full((nsc,), s2v, dtype=torch.float32, device=dev)
        C = torch.empty((nsc, l, a), dtype=torch.complex64, device=dev)
        ni = torch.empty((nsc, l), dtype=torch.float32, device=dev)
        llr = torch.empty((nsc, l, S, 2 * q), dtype=torch.float32, device=dev)
        cells.append(dict(H=H, s2=s2, Y=Y, C=C, ni=ni, llr=llr, q=q, nsc=nsc, A=a, L=l,
                          desc=pe.cell_desc(C, ni, Y, llr, q, prec, H, s2)))
    return cells


def codeword_bench(dev, C, Y, ni, steps=50):
    """Descrambling and the rate-recovery layout in the fused epilogue: the 32-cell batch as one allocation of
    50,688 subcarriers carrying one 8-layer codeword (G = 35.7 M bits), LLRs descrambled with its Gold sequence
    and written in codeword order, or de-interleaved per code block (C = ceil(G / 8448): LDPC base graph 1 blocks),
    against the same launch writing the per-symbol layout. LLR bytes are the same; the epilogue also reads one bit
    of the sequence per LLR."""
    import torch
    import pusch_eq as pe
    nsc = C.shape[0]
    G = nsc * S * L * 2 * Q
    c = pe.gold_sequence((20000 << 15) | 1, G)  # n_RNTI = 20000, n_ID = 1
    ncb = -(-G // 8448)
    outs = [torch.empty(G, dtype=torch.float32, device=dev)]
    llr = torch.empty((nsc, L, S, 2 * Q), dtype=torch.float32, device=dev)
    runs = {"per_symbol": lambda: pe.coef_apply_demap(C, Y, ni, Q, prec="f16", out=llr),
            "codeword_order": lambda: pe.coef_apply_demap_cw(C, Y, ni, Q, [0] * L, [c], [0], outs=outs),
            "deinterleaved": lambda: pe.coef_apply_demap_cw(C, Y, ni, Q, [0] * L, [c], [ncb], outs=outs)}
    base_bytes = nsc * (L * A * 8 + S * A * 8 + L * S * 2 * Q * 4 + L * 4)
    res = {}
    for name, fn in runs.items():
        el, _ = time_launches(lambda: fn(), [()], steps, 5)
        by = base_bytes + (0 if name == "per_symbol" else G // 8)
        res[name] = {"ms_per_launch": el / steps * 1e3, "bytes_per_launch": by, "hbm_gbs": by / (el / steps) / 1e9}
    gel, _ = time_launches(lambda: pe.gold_sequence((20000 << 15) | 1, G, out=c), [()], 20, 2)
    res["gold_sequence_ms"] = gel / 20 * 1e3
    res["workload"] = ("%d subcarriers x 8 layers x 11 symbols x 64 antennas, 256-QAM, f16 MFMA, FP32 LLRs; one "
                       "codeword of G = %d bits, %d code blocks when de-interleaved; the sequence (one bit per LLR, "
                       "%d bytes) is made once per UE and reused every slot" % (nsc, G, ncb, G // 8))
    return res


def slot_bytes(cells):
    """Algorithmic HBM bytes of one slot: MMSE (H, sigma2 in; C, noise_inv, flags out) + fused apply/demap
    (C, Y_rx, noise_inv in; LLR out)."""
    b = 0
    for c in cells:
        n, a, l, q = c["nsc"], c["A"], c["L"], c["q"]
        b += n * (a * l * 8 + 4 + l * a * 8 + l * 4 + 1)
        b += n * (l * a * 8 + S * a * 8 + l * 4 + l * S * 2 * q * 4)
    return b


def slot_bytes_records(cells):
    """The same slot with solve records (PE_SLOT_SOLVE_RECORDS; every cell of the draw qualifies): MMSE (H, sigma2
    in; record, noise_inv, flags out) + apply (H, record, flags, Y_rx, noise_inv in; LLR out)."""
    b = 0
    for c in cells:
        n, a, l, q = c["nsc"], c["A"], c["L"], c["q"]
        rec = l * (l + 2) * 4
        b += n * (a * l * 8 + 4 + rec + l * 4 + 1)
        b += n * (a * l * 8 + rec + 1 + S * a * 8 + l * 4 + l * S * 2 * q * 4)
    return b


def time_slots(plans, mode, steps):
    """Back-to-back slots alternating over the plans (seconds per slot), and the median latency of one slot on
    an idle GPU (ms; events around the call, so host submission is included)."""
    import torch
    for i in range(6):
        getattr(plans[i % 2], mode)()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(steps):
        getattr(plans[i % 2], mode)()
    e1.record()
    torch.cuda.synchronize()
    per = e0.elapsed_time(e1) / 1e3 / steps
    lat = []
    for i in range(40):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        getattr(plans[i % 2], mode)()
        b.record()
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat.sort()
    return per, lat[len(lat) // 2]


def multicell_bench(dev, steps):
    """24 uneven cells; one slot = the slot plan's kernel sequence (MMSE with its in-kernel FP64 fallback, one fused
    apply/demap launch) as one graph replay, and the same launched eagerly; two slot buffer sets alternate.
    Also the latency of one slot on an idle GPU (events around the call, so host submission is included)."""
    import torch
    with torch.cuda.stream(torch.cuda.Stream()):  # a receiver's own stream, not the legacy default stream
        return _multicell_bench(dev, steps)


def _multicell_bench(dev, steps):
    import torch
    import pusch_eq as pe
    sets = [draw_cells(24, 42 + k, dev) for k in range(2)]
    plans = [pe.SlotPlan([c["desc"] for c in cs]) for cs in sets]
    res = {}
    for mode in ("graph_launch", "launch"):
        res[mode], res[mode + "_latency"] = time_slots(plans, mode, steps)
    t = res["graph_launch"]
    # the two plans alternate: per-slot figures are their means
    nscs = [sum(c["nsc"] for c in cs) for cs in sets]
    bys = [slot_bytes(cs) for cs in sets]
    nsc, by = sum(nscs) / 2, sum(bys) / 2
    tiles, wg = plans[0].info()
    for p in plans:
        p.close()
    # the same plans with solve records: coefficients formed on chip by the apply
    plans_r = [pe.SlotPlan([c["desc"] for c in cs], solve_records=True) for cs in sets]
    rec = {}
    for mode in ("graph_launch", "launch"):
        rec[mode], rec[mode + "_latency"] = time_slots(plans_r, mode, steps)
    by_r = sum(slot_bytes_records(cs) for cs in sets) / 2
    for p in plans_r:
        p.close()
    solve_records = {"slot_ms": rec["graph_launch"] * 1e3, "slot_ms_eager": rec["launch"] * 1e3,
                     "subcarriers_per_s": nsc / rec["graph_launch"],
                     "subcarriers_per_s_eager": nsc / rec["launch"], "bytes_per_slot": by_r,
                     "hbm_gbs_eager": by_r / rec["launch"] / 1e9,
                     "idle_slot_latency_ms": {"graph": rec["graph_launch_latency"], "eager": rec["launch_latency"]},
                     "workload": "the same plans with PE_SLOT_SOLVE_RECORDS (stage 1 writes solve records, stage 2 "
                                 "forms the coefficients on chip)"}
    # the same cells with binary16 coefficients (PE_PREC_F16_C16): eager back-to-back slots
    for cs in sets:
        for c in cs:
            c["desc"] = pe.cell_desc(c["C"], c["ni"], c["Y"], c["llr"], c["q"], "f16_c16", c["H"], c["s2"])
    plans16 = [pe.SlotPlan([c["desc"] for c in cs]) for cs in sets]
    for i in range(6):
        plans16[i % 2].launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(steps):
        plans16[i % 2].launch()
    e1.record()
    torch.cuda.synchronize()
    t16 = e0.elapsed_time(e1) / 1e3 / steps
    for p in plans16:
        p.close()
    return {"binary16_coefficients": {"slot_ms_eager": t16 * 1e3, "subcarriers_per_s": nsc / t16,
                                      "workload": "the same plans with cells of precision f16_c16"},
            "solve_records": solve_records,
            "workload": "24 cells per slot, PRB {24,52,106,273}, antennas {16,32,64}, layers {2,4,8}, QAM "
                        "{16,64,256}, f16, cell c of a plan drawn from seed (42 or 43) + 1000 c; two plans "
                        "alternate; one slot = MMSE + fused apply/demap, back-to-back slots on a created stream",
            "subcarriers_per_plan": nscs, "bytes_per_plan": bys,
            "subcarriers_per_slot": nsc, "slot_ms": t * 1e3, "subcarriers_per_s": nsc / t,
            "hbm_gbs": by / t / 1e9, "hbm_frac": by / t / 1e9 / HBM_PEAK_GBS, "bytes_per_slot": by,
            "slot_ms_eager": res["launch"] * 1e3, "hbm_gbs_eager": by / res["launch"] / 1e9,
            "idle_slot_latency_ms": {"graph": res["graph_launch_latency"], "eager": res["launch_latency"]},
            "tiles": tiles, "max_workgroups": wg}


def srs_bench(dev, iters=500, warmup=50):
    """SRS filter multiply (pe_srs_filter) on the configurations of srs_wmma_bench.cu main (:347-353): 64
    antennas, 66 computation blocks; binary16 chain and MFMA. Per-launch times are medians of single launches
    bracketed by events (the reference's method); back-to-back launches give the device-side rate."""
    import torch
    import pusch_eq as pe
    g = torch.Generator(device=dev)
    g.manual_seed(42)
    nAnt, nBlocks = 64, 66

    def c16(*shape):
        v = ((torch.rand(shape + (2,), generator=g, device=dev) - 0.5) * 1.4).half()
        return v.view(torch.int32).reshape(shape)

    rows = []
    for K, P, FL in ((24, 4, 8), (24, 2, 8), (12, 4, 12), (12, 2, 12)):
        S, W, F = c16(nBlocks, nAnt, K), c16(K, K), c16(P, FL)
        outs = {}
        row = {"K": K, "ports": P, "focc_len": FL, "antennas": nAnt, "blocks": nBlocks}
        ops = nBlocks * nAnt * K * P * K * 8.0
        nbytes = 4 * (nBlocks * nAnt * K + nBlocks * nAnt * K * P + K * K + P * FL)
        for name, mf in (("f16_chain", False), ("mfma_f32acc", True)):
            out = torch.empty((nBlocks, nAnt, K, P), dtype=torch.int32, device=dev)
            for _ in range(warmup):
                pe.srs_filter(S, W, F, mfma=mf, out=out)
            torch.cuda.synchronize()
            ts = []
            for _ in range(iters):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                pe.srs_filter(S, W, F, mfma=mf, out=out)
                b.record()
                b.synchronize()
                ts.append(a.elapsed_time(b))
            ts.sort()
            el, _ = time_launches(lambda: pe.srs_filter(S, W, F, mfma=mf, out=out), [()], iters, 10)
            b2b_ms = el / iters * 1e3
            # device-side rate without host launch cost: 50 launches captured in one graph, replayed
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                for _ in range(50):
                    pe.srs_filter(S, W, F, mfma=mf, out=out)
            graph.replay()
            torch.cuda.synchronize()
            elg, _ = time_launches(graph.replay, [()], 20, 2)
            dev_ms = elg / (20 * 50) * 1e3
            row[name] = {"median_ms": ts[len(ts) // 2], "gflops_median": ops / (ts[len(ts) // 2] * 1e-3) / 1e9,
                         "back_to_back_ms": b2b_ms, "graph_ms_per_launch": dev_ms,
                         "gflops_graph": ops / (dev_ms * 1e-3) / 1e9, "gbs_graph": nbytes / (dev_ms * 1e-3) / 1e9}
            del graph
            outs[name] = out.view(torch.float16).float()
        a, b = outs["f16_chain"].flatten(), outs["mfma_f32acc"].flatten()
        row["cosine_chain_vs_mfma"] = float((a @ b) / (a.norm() * b.norm()))
        rows.append(row)
    return rows


def operators_bench(dev, n_cells, steps=50):
    """The standalone drop-ins of the reference's other kernels over the same batch size (n_cells x 1584
    subcarriers, 8 layers, 11 symbols, 64 antennas, 256-QAM): pe_gram (gram_scalar's M @ H, FP32, bit-exact
    order), pe_llr_demap_f16 (llr_v1_baseline's binary16 format, one noise value per symbol) and pe_llr_demap
    (FP32 on the equalizer layout). Mean of `steps` back-to-back launches rotating over buffer sets that together
    exceed the 256 MB Infinity Cache (2 sets for the Gram, 4 for the demappers), so the bytes come from HBM."""
    import torch
    import pusch_eq as pe
    nsc = n_cells * NSC_CELL
    g = torch.Generator(device=dev)
    g.manual_seed(77)
    out = {}

    def timed(fn, bufs):
        el, _ = time_launches(fn, bufs, steps, 5)
        return el / steps

    # Gram M @ H: M [nsc, L, A], H [nsc, A, L] -> G [nsc, L, L]
    gb = []
    for k in range(2):
        M = torch.randn((nsc, L, A), dtype=torch.complex64, device=dev, generator=g)
        Hh = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g)
        gb.append((M, Hh, torch.empty((nsc, L, L), dtype=torch.complex64, device=dev)))
    t = timed(lambda M, Hh, G: pe.gram(M, Hh, out=G), gb)
    by = nsc * (L * A * 8 + A * L * 8 + L * L * 8)
    out["gram_f32"] = {"ms_per_launch": t * 1e3, "subcarriers_per_s": nsc / t, "bytes_per_launch": by,
                       "hbm_gbs": by / t / 1e9}
    del gb
    # binary16 demapper, the reference's format: N symbols, noise per symbol
    N = nsc * L * S
    db = []
    for k in range(4):
        soft = (torch.randn((N, 2), device=dev, generator=g) * 0.7).to(torch.float16)
        nis = torch.rand((N,), device=dev, generator=g) * 10 + 1
        db.append((soft, nis, torch.empty((N, 2 * Q), dtype=torch.float16, device=dev)))
    t = timed(lambda soft, nis, o: pe.llr_demap_f16(soft, nis, Q, out=o), db)
    by = N * (4 + 4 + 2 * Q * 2)
    out["llr_demap_f16"] = {"ms_per_launch": t * 1e3, "symbols_per_s": N / t, "bytes_per_launch": by,
                            "hbm_gbs": by / t / 1e9, "symbols": N}
    del db
    # FP32 demapper on Y_eq [nsc, L, S], noise_inv [nsc, L]
    fb = []
    for k in range(4):
        Yeq = torch.randn((nsc, L, S), dtype=torch.complex64, device=dev, generator=g) * 0.7
        ni = torch.rand((nsc, L), device=dev, generator=g) * 10 + 1
        fb.append((Yeq, ni, torch.empty((nsc, L, S, 2 * Q), dtype=torch.float32, device=dev)))
    t = timed(lambda Yeq, ni, o: pe.llr_demap(Yeq, ni, Q, out=o), fb)
    by = nsc * (L * S * 8 + L * 4 + L * S * 2 * Q * 4)
    out["llr_demap_f32"] = {"ms_per_launch": t * 1e3, "symbols_per_s": N / t, "bytes_per_launch": by,
                            "hbm_gbs": by / t / 1e9}
    out["workload"] = ("%d x 1584 subcarriers, 8 layers, 11 symbols, 64 antennas, 256-QAM; random inputs; "
                       "buffer sets rotate beyond the Infinity Cache" % n_cells)
    return out


def fronthaul_bench(dev, n_cells, steps=50):
    """pe_bfp_decompress of the headline batch: n_cells x 132 PRB, 11 symbols, 64 antennas, 9-bit block
    floating point (28 bytes per PRB) -> complex64 Y_rx; random bytes (every byte pattern is a valid block)."""
    import torch
    import pusch_eq as pe
    nPrb = n_cells * NSC_CELL // 12
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    packed = torch.randint(0, 256, (S * A, nPrb * 28), dtype=torch.uint8, device=dev, generator=g)
    out = torch.empty((nPrb * 12, S, A), dtype=torch.complex64, device=dev)
    el, _ = time_launches(lambda: pe.bfp_decompress(packed, nPrb, S, A, 9, 2.0 ** -15, out=out), [()], steps, 5)
    per = el / steps
    nbytes = packed.numel() + out.numel() * 8
    del packed, out
    return {"workload": "%d PRB x 11 symbols x 64 antennas, 9-bit block floating point -> complex64 Y_rx" % nPrb,
            "ms_per_launch": per * 1e3, "subcarriers_per_s": nPrb * 12 / per, "hbm_gbs": nbytes / per / 1e9,
            "bytes_per_launch": nbytes}


FI_M, FI_K, FI_NC = 64, 96, 4           # the frequency-interpolation GEMM the reference names: [64 x 96] @ [96 x 4]
FI_BATCH = 14 * 64 * 64                  # 14 symbols x 64 antennas x 64 UE allocations


def interp_chain_bench(dev, steps=40, nblk=792):
    """DMRS estimates to LLRs for the 32-cell batch at 4 layers (50,688 subcarriers = 792 blocks of 64, 64
    antennas, 96 pilots per block, 11 data symbols, 256-QAM, bf16 interpolation, f16 apply), two buffer sets
    alternating: pe_interp_mmse_apply_demap (interpolation fused with the MMSE solve, H written once) against
    pe_freq_interp + the re-layout of H_f into the equalizer's H[sc][ant][layer] + pe_mmse_apply_demap, and each
    of those three alone. Y = H x + n with x uniform 256-QAM at 20 dB."""
    import torch
    import pusch_eq as pe
    A, K, S, q = 64, 96, 11, 4
    nsc = 64 * nblk
    g = torch.Generator(device=dev)
    g.manual_seed(4096)
    W = torch.randn((64, K), dtype=torch.complex64, device=dev, generator=g) / K ** 0.5
    s2v = 10.0 ** (-SNR_DB / 10.0)
    sets = []
    for k in range(2):
        X = torch.randn((nblk, A, K, 4), dtype=torch.complex64, device=dev, generator=g) / A ** 0.5
        s2 = torch.full((nsc,), s2v, dtype=torch.float32, device=dev)
        H = pe.freq_interp(W, X.reshape(nblk * A, K, 4), prec="bf16").reshape(nblk, A, 64, 4)
        H = H.permute(0, 2, 1, 3).reshape(nsc, A, 4).contiguous()
        bits = torch.randint(0, 2, (nsc, 4, S, 2 * q), device=dev, generator=g, dtype=torch.uint8)
        Y = torch.bmm(H, qam_map_torch(bits, q))
        Y = (Y + torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=g) * s2v ** 0.5)
        Y = Y.transpose(1, 2).contiguous()
        sets.append({"X": X, "s2": s2, "Y": Y, "H": torch.empty_like(H), "Hf": torch.empty((nblk * A, 64, 4),
                     dtype=torch.complex64, device=dev), "Hl": torch.empty_like(H),
                     "C": torch.empty((nsc, 4, A), dtype=torch.complex64, device=dev),
                     "ni": torch.empty((nsc, 4), dtype=torch.float32, device=dev),
                     "fl": torch.empty((nsc,), dtype=torch.uint8, device=dev),
                     "llr": torch.empty((nsc, 4, S, 2 * q), dtype=torch.float32, device=dev),
                     "llr2": torch.empty((nsc, 4, S, 2 * q), dtype=torch.float32, device=dev)})
        del H, bits

    def fused(b):
        pe.interp_mmse_apply_demap(W, b["X"], b["s2"], b["Y"], q, H=b["H"], C=b["C"], noise_inv=b["ni"],
                                   flags=b["fl"], out=b["llr"])

    def interp(b):
        pe.freq_interp(W, b["X"].reshape(nblk * A, K, 4), prec="bf16", out=b["Hf"])

    def relayout(b):
        b["Hl"].view(nblk, 64, A, 4).copy_(b["Hf"].view(nblk, A, 64, 4).permute(0, 2, 1, 3))

    def mmse(b):
        pe.mmse_apply_demap(b["Hl"], b["s2"], b["Y"], q, C=b["C"], noise_inv=b["ni"], flags=b["fl"], out=b["llr2"])

    def separate(b):
        interp(b)
        relayout(b)
        mmse(b)

    bufs = [(b,) for b in sets]
    t = {}
    for name, fn in (("fused", fused), ("separate", separate), ("interp", interp), ("relayout", relayout),
                     ("mmse_apply_demap", mmse)):
        el, _ = time_launches(fn, bufs, steps, 5)
        t[name] = el / steps * 1e6
    fused(sets[0])
    separate(sets[0])
    torch.cuda.synchronize()
    identical = bool(torch.equal(sets[0]["llr"], sets[0]["llr2"]) and torch.equal(sets[0]["H"], sets[0]["Hl"]))
    rec = 4 * (4 + 2) * 4  # solve record bytes at 4 layers
    b_x, b_h, b_y = nblk * A * K * 4 * 8, nsc * A * 4 * 8, nsc * S * A * 8
    b_llr, b_ni = nsc * 4 * S * 2 * q * 4, nsc * (4 * 4 + 1)
    bytes_fused = b_x + b_h + nsc * rec + b_ni + (b_h + nsc * rec + b_y + b_llr + b_ni)
    bytes_sep = b_x + b_h + 2 * b_h + (b_h + nsc * rec + b_ni) + (b_h + nsc * rec + b_y + b_llr + b_ni)
    return {"workload": "%d blocks of 64 subcarriers (%d), 64 antennas, 96 pilots, 4 layers, 11 symbols, 256-QAM at "
                        "20 dB; bf16 interpolation, f16 apply; two buffer sets alternating" % (nblk, nsc),
            "us_fused": t["fused"], "us_separate": t["separate"], "us_interp": t["interp"],
            "us_relayout": t["relayout"], "us_mmse_apply_demap": t["mmse_apply_demap"],
            "speedup_vs_separate": t["separate"] / t["fused"],
            "speedup_vs_separate_without_relayout": (t["interp"] + t["mmse_apply_demap"]) / t["fused"],
            "bytes_fused": bytes_fused, "bytes_separate": bytes_sep,
            "hbm_gbs_fused": bytes_fused / (t["fused"] * 1e-6) / 1e9,
            "subcarriers_per_s_fused": nsc / (t["fused"] * 1e-6),
            "llrs_and_h_identical": identical}


def freq_interp_bench(dev, steps=50, cpu_seconds=2.0, precs=("bf16", "bf16_split", "f32")):
    """pe_freq_interp over FI_BATCH items of [96 pilots x 4 ports] through one [64 x 96] filter, per precision:
    mean of `steps` launches rotating over 2 buffer sets (586 MB together, beyond the Infinity Cache). Bytes per
    item: X 3,072 + H_f 2,048 (W, 48 KB, once per launch). The oracle's FP32 chain (OpenMP over items) on a
    sample of items is the CPU figure."""
    import torch
    import pusch_eq as pe
    g = torch.Generator(device=dev)
    g.manual_seed(96)
    W = torch.randn((FI_M, FI_K), dtype=torch.complex64, device=dev, generator=g)
    sets = []
    for k in range(2):
        X = torch.randn((FI_BATCH, FI_K, FI_NC), dtype=torch.complex64, device=dev, generator=g)
        sets.append((X, torch.empty((FI_BATCH, FI_M, FI_NC), dtype=torch.complex64, device=dev)))
    item_bytes = FI_K * FI_NC * 8 + FI_M * FI_NC * 8
    nbytes = FI_BATCH * item_bytes + FI_M * FI_K * 8
    out = {"workload": "%d items (14 symbols x 64 antennas x 64 UE allocations) of [%d pilots x %d ports] through "
                       "one [%d x %d] filter, complex64 in and out" % (FI_BATCH, FI_K, FI_NC, FI_M, FI_K),
           "bytes_per_launch": nbytes, "flops_per_launch": 8 * FI_M * FI_K * FI_NC * FI_BATCH}
    for p in precs:
        el, _ = time_launches(lambda X, H: pe.freq_interp(W, X, prec=p, out=H), sets, steps, 5)
        per = el / steps
        out[p] = {"ms_per_launch": per * 1e3, "items_per_s": FI_BATCH / per, "hbm_gbs": nbytes / per / 1e9,
                  "tflops": out["flops_per_launch"] / per / 1e12}
    if cpu_seconds > 0:  # CPU baseline leg: the oracle, timed
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import numpy as np
        from pe_oracle import freq_interp_f32_omp
        Wh = W.cpu().numpy()
        Xh = sets[0][0][:4096].cpu().numpy()
        Hh = np.empty((Xh.shape[0], FI_M, FI_NC), dtype=np.complex64)
        freq_interp_f32_omp(Wh, Xh, Hh)  # warm-up
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            _, used = freq_interp_f32_omp(Wh, Xh, Hh)
            reps += 1
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"items_per_s": reps * Xh.shape[0] / el, "cores": used, "kind": "port",
                               "sample": "%d repetitions of %d items, oracle FP32 chain, OpenMP over items, %.1f s"
                                         % (reps, Xh.shape[0], el)}
    del sets
    return out


def serving_bench(dev, n_jobs):
    """64 cells; jobs (one slot of one cell: MMSE + fused apply/demap of 11 symbols) arrive as a Poisson
    process; capacity = throughput with every job queued at t = 0; then offered load 10-120% of capacity."""
    import torch
    import pusch_eq as pe
    cells = draw_cells(64, 42, dev)
    descs = [pe.job_desc(c["H"], c["s2"], c["Y"], c["q"], "f16") for c in cells]
    kw = dict(deadline_s=0.5e-3, hbm_budget=2 << 30, max_batch=64, max_inflight=3, shed_late=False,
              use_graph=False)
    pe.serve_run(descs, min(n_jobs, 2000), 0.0, seed=1, **kw)  # warm-up run (GPU clocks, first use)
    cap, _, _ = pe.serve_run(descs, n_jobs, 0.0, seed=1, **kw)
    capacity = cap["jobs_per_s"]
    descs16 = [pe.job_desc(c["H"], c["s2"], c["Y"], c["q"], "f16_c16") for c in cells]
    cap16, _, _ = pe.serve_run(descs16, n_jobs, 0.0, seed=1, **kw)
    sweep = []
    for f in (0.1, 0.3, 0.5, 0.7, 0.9, 1.0, 1.1, 1.2):
        rep, _, _ = pe.serve_run(descs, n_jobs, f * capacity, seed=2, **kw)
        sweep.append({"load": f, "offered_jobs_per_s": f * capacity, "jobs_per_s": rep["jobs_per_s"],
                      "p50_ms": rep["p50_s"] * 1e3, "p99_ms": rep["p99_s"] * 1e3, "max_ms": rep["max_s"] * 1e3,
                      "miss_rate": rep["miss_rate"], "late": rep["n_late"], "mean_batch_jobs": rep["mean_batch_jobs"]})
    # deadline-aware admission (shed jobs predicted to finish late, margin calibrated from observed lateness), same
    # arrival streams
    admission = []
    for f in (0.1, 0.3, 0.5, 0.7, 0.9, 1.0, 1.1, 1.2, 1.5):
        rep, _, _ = pe.serve_run(descs, n_jobs, f * capacity, seed=2, **dict(kw, shed_late=2))
        admission.append({"load": f, "jobs_per_s": rep["jobs_per_s"], "p50_ms": rep["p50_s"] * 1e3,
                          "p99_ms": rep["p99_s"] * 1e3, "max_ms": rep["max_s"] * 1e3,
                          "miss_rate": rep["miss_rate"], "late": rep["n_late"], "shed": rep["n_shed"],
                          "final_margin_ms": rep["model_margin_s"] * 1e3, "mean_batch_jobs": rep["mean_batch_jobs"]})
    # plain FP32 path, one job at a time (per-cell launches: MMSE then fused FP32 apply/demap)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for c in cells:
        pe.gram_mmse(c["H"], c["s2"], unbiased=True, C=c["C"], noise_inv=c["ni"])
        pe.coef_apply_demap(c["C"], c["Y"], c["ni"], c["q"], prec="f32", out=c["llr"])
    e1.record()
    torch.cuda.synchronize()
    f32_jobs_per_s = len(cells) / (e0.elapsed_time(e1) / 1e3)
    return {"workload": "64 cells drawn as the multi-cell batch, Poisson arrivals, 0.5 ms deadline, continuous "
                        "batching (<= 64 jobs per batch, 3 streams, 2 GiB budget, eager launches), f16",
            "jobs": n_jobs, "capacity_jobs_per_s": capacity,
            "capacity_subcarriers_per_s": cap["subcarriers_per_s"], "capacity_p99_ms": cap["p99_s"] * 1e3,
            "load_sweep": sweep, "deadline_aware_admission": admission,
            "iso_job_ms": cap["iso_job_s"] * 1e3, "contention_index_first5": cap["contention_index"],
            "capacity_jobs_per_s_binary16_coefficients": cap16["jobs_per_s"],
            "plain_f32_sequential_jobs_per_s": f32_jobs_per_s}


SWEEP_PRECS = ("f32", "bf16", "bf16_split", "f16", "f16_split")


def deployment_sweep(dev, seed=5, reps=10):
    """One cell per launch: antennas {8,16,32,64,128} x PRB {25,51,106,133,273} x layers {2,4,8}, 11 symbols;
    QAM order (QPSK..256-QAM) and SNR (0-30 dB) drawn per configuration; Rayleigh CN(0, 1/A) channels. Per
    configuration: apply EVM of each MFMA precision vs the FP64 product of the same coefficients, and the
    device time per launch of the fused apply + demap (median of `reps` single launches, each queued behind a
    short spin so that host launch overhead is excluded). The bf16 path at the -60 dB bar is plain bf16 where
    that meets it, else bf16_split; it is compared with the plain FP32 path. Also the chain from channel
    estimates at f16 (pe_gram_mmse + pe_coef_apply_demap against pe_mmse_apply_demap), timed the same way, and
    whether the two give identical LLRs.
    Roofline per configuration, f16 fused apply + demap: algorithmic bytes per launch (C and Y_rx in, LLRs and
    noise_inv), and GB/s and HBM fraction twice: warm (the single-launch times above: the cell's buffers, at most
    73 MB, stay in the 256 MB MALL between launches) and cold (each launch preceded by a 512 MB device write that
    evicts them, so the inputs come from HBM)."""
    import numpy as np
    import torch
    import pusch_eq as pe
    rng = np.random.default_rng(seed)
    rows = []
    flush = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
    for A_ in (8, 16, 32, 64, 128):
        for prb in (25, 51, 106, 133, 273):
            for L_ in (2, 4, 8):
                q = int(rng.integers(1, 5))
                snr = float(rng.uniform(0.0, 30.0))
                nsc = prb * 12
                g = torch.Generator(device=dev)
                g.manual_seed(int(rng.integers(1 << 30)))
                H = torch.randn((nsc, A_, L_), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A_) ** 0.5
                bits = torch.randint(0, 2, (nsc, L_, S, 2 * q), device=dev, generator=g, dtype=torch.uint8)
                x = qam_map_torch(bits, q)
                s2 = 10.0 ** (-snr / 10.0)
                Y = torch.bmm(H, x) + torch.randn((nsc, A_, S), dtype=torch.complex64, device=dev,
                                                  generator=g) * s2 ** 0.5
                Y = Y.transpose(1, 2).contiguous()
                sigma2 = torch.full((nsc,), s2, dtype=torch.float32, device=dev)
                C, ni, fl = pe.gram_mmse(H, sigma2, unbiased=True)
                yref = torch.bmm(C.to(torch.complex128), Y.to(torch.complex128).transpose(1, 2).contiguous())
                llr = torch.empty((nsc, L_, S, 2 * q), dtype=torch.float32, device=dev)
                row = {"A": A_, "prb": prb, "L": L_, "qam": 4 ** q, "snr_db": round(snr, 2), "nsc": nsc,
                       "fp64_mmse_scs": int((fl != 0).sum())}
                for prec in SWEEP_PRECS:
                    yeq = pe.coef_apply(C, Y, prec=prec)
                    row["evm_db_" + prec] = evm_db(yref, yeq)
                    ts = []
                    for i in range(reps + 2):
                        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        torch.cuda._sleep(100000)
                        a.record()
                        pe.coef_apply_demap(C, Y, ni, q, prec=prec, out=llr)
                        b.record()
                        b.synchronize()
                        if i >= 2:
                            ts.append(a.elapsed_time(b) * 1e3)
                    ts.sort()
                    row["us_" + prec] = ts[len(ts) // 2]
                # the chain from channel estimates, f16: two launches against the one pass (same device-time method)
                def dev_us(fn):
                    ts = []
                    for i in range(reps + 2):
                        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        torch.cuda._sleep(100000)
                        a.record()
                        fn()
                        b.record()
                        b.synchronize()
                        if i >= 2:
                            ts.append(a.elapsed_time(b) * 1e3)
                    ts.sort()
                    return ts[len(ts) // 2]

                C2, ni2, fl2 = torch.empty_like(C), torch.empty_like(ni), torch.empty_like(fl)
                llr1, llr2 = torch.empty_like(llr), torch.empty_like(llr)
                row["us_chain_two_launches"] = dev_us(lambda: (
                    pe.gram_mmse(H, sigma2, unbiased=True, C=C2, noise_inv=ni2, flags=fl2),
                    pe.coef_apply_demap(C2, Y, ni2, q, prec="f16", out=llr1)))
                row["us_chain_one_pass"] = dev_us(lambda: pe.mmse_apply_demap(H, sigma2, Y, q, prec="f16", C=C2,
                                                                              noise_inv=ni2, flags=fl2, out=llr2))
                row["one_pass_identical"] = bool(torch.equal(llr1, llr2))
                # roofline of the f16 fused apply + demap, warm (above) and cold (MALL evicted before each launch)
                ts = []
                for i in range(reps + 2):
                    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    flush.fill_(i & 255)
                    a.record()
                    pe.coef_apply_demap(C, Y, ni, q, prec="f16", out=llr)
                    b.record()
                    b.synchronize()
                    if i >= 2:
                        ts.append(a.elapsed_time(b) * 1e3)
                ts.sort()
                row["us_f16_cold"] = ts[len(ts) // 2]
                by = nsc * (L_ * A_ * 8 + S * A_ * 8 + L_ * S * 2 * q * 4 + L_ * 4)
                row["bytes_f16"] = by
                row["gbs_f16_warm"] = by / row["us_f16"] / 1e3
                row["gbs_f16_cold"] = by / row["us_f16_cold"] / 1e3
                row["hbm_frac_f16_warm"] = row["gbs_f16_warm"] / HBM_PEAK_GBS
                row["hbm_frac_f16_cold"] = row["gbs_f16_cold"] / HBM_PEAK_GBS
                row["speedup_one_pass"] = row["us_chain_two_launches"] / row["us_chain_one_pass"]
                bf = "bf16" if row["evm_db_bf16"] <= -60.0 else "bf16_split"
                row["bf16_at_bar"] = bf
                row["speedup_bf16_at_bar_vs_f32"] = row["us_f32"] / row["us_" + bf]
                row["speedup_f16_vs_f32"] = row["us_f32"] / row["us_f16"]
                rows.append(row)
    import statistics
    summ = {"configs": len(rows),
            "workload": "one cell per launch, antennas {8,16,32,64,128} x PRB {25,51,106,133,273} x layers "
                        "{2,4,8}, 11 symbols, QAM and SNR (0-30 dB) drawn per config, fused apply + demap",
            "evm_bar_db": -60.0}
    for prec in SWEEP_PRECS[1:]:
        evs = [r["evm_db_" + prec] for r in rows]
        summ["worst_evm_db_" + prec] = max(evs)
        summ["configs_at_bar_" + prec] = sum(e <= -60.0 for e in evs)
    summ["bf16_at_bar_speedup_vs_f32_median"] = statistics.median(r["speedup_bf16_at_bar_vs_f32"] for r in rows)
    summ["bf16_at_bar_speedup_vs_f32_min"] = min(r["speedup_bf16_at_bar_vs_f32"] for r in rows)
    summ["f16_speedup_vs_f32_median"] = statistics.median(r["speedup_f16_vs_f32"] for r in rows)
    sp = [r["speedup_one_pass"] for r in rows]
    summ["one_pass_chain_speedup_median"] = statistics.median(sp)
    summ["one_pass_chain_speedup_min_max"] = [min(sp), max(sp)]
    summ["one_pass_chain_speedup_median_64_antennas"] = statistics.median(
        r["speedup_one_pass"] for r in rows if r["A"] == 64)
    summ["one_pass_llrs_identical_all"] = all(r["one_pass_identical"] for r in rows)
    for kind in ("cold", "warm"):
        fr = sorted(rows, key=lambda r: r["hbm_frac_f16_" + kind])
        summ["hbm_frac_median_f16_" + kind] = statistics.median(r["hbm_frac_f16_" + kind] for r in rows)
        w = fr[0]
        summ["hbm_frac_worst_f16_" + kind] = {"frac": w["hbm_frac_f16_" + kind], "A": w["A"], "prb": w["prb"],
                                             "L": w["L"], "qam": w["qam"], "us": w["us_f16_" + kind]
                                             if kind == "cold" else w["us_f16"], "bytes": w["bytes_f16"]}
        b = fr[-1]
        summ["hbm_frac_best_f16_" + kind] = {"frac": b["hbm_frac_f16_" + kind], "A": b["A"], "prb": b["prb"],
                                            "L": b["L"], "qam": b["qam"]}
    summ["roofline_note"] = ("one cell per launch: cold = MALL evicted before each launch (inputs from HBM), warm = "
                             "the cell's buffers resident in the 256 MB MALL from the previous launch; "
                             "small cells are launch-latency-bound either way")
    big = [r for r in rows if r["nsc"] == 3276]
    summ["largest_cells_bf16_at_bar_speedup_vs_f32_median"] = statistics.median(
        r["speedup_bf16_at_bar_vs_f32"] for r in big)
    return summ, rows


def load_traffic(workload_key):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(workload_key)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=32, help="cell-slots per launch")
    ap.add_argument("--prec", default="f16")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-iters", type=int, default=300)
    ap.add_argument("--no-multicell", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=10000, help="jobs per serving-loop run (0 = skip)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the deployment sweep")
    ap.add_argument("--no-srs", action="store_true", help="skip the SRS filter section")
    ap.add_argument("--srs-only", action="store_true", help="only the SRS filter section, printed as JSON")
    ap.add_argument("--mmse-only", action="store_true",
                    help="only the MMSE stage over the 32-cell batch (10 launches each output format), for profiling")
    ap.add_argument("--chain-only", action="store_true",
                    help="only the chain from channel estimates over the 32-cell batch (10 two-launch chains, then "
                         "10 pe_mmse_apply_demap calls), for profiling")
    ap.add_argument("--fronthaul-only", action="store_true",
                    help="only the fronthaul decode of the 32-cell batch, for profiling")
    ap.add_argument("--freq-interp-only", action="store_true",
                    help="only the DMRS frequency interpolation (10 launches per precision), for profiling")
    ap.add_argument("--sweep-out", default=None, help="write the deployment sweep table (JSON) here")
    ap.add_argument("--sweep-only", action="store_true", help="only the deployment sweep, summary printed as JSON")
    ap.add_argument("--multicell-only", action="store_true",
                    help="only the 24-cell slot section (graph and eager slots), printed as JSON (for profiling)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed headline launches (fused apply + demap over the 32-cell batches), for "
                         "profiling")
    ap.add_argument("--serving-only", action="store_true", help="only the serving-loop section, printed as JSON")
    ap.add_argument("--codeword-only", action="store_true",
                    help="only the descrambling / codeword-layout launches over the first 32-cell batch, for profiling")
    ap.add_argument("--interp-chain-only", action="store_true",
                    help="only the DMRS-estimates-to-LLRs chain (fused interpolation + MMSE against separate calls)")
    ap.add_argument("--ceiling-only", action="store_true",
                    help="only the measured streaming ceilings (pe_hbm_probe), printed as JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 plain 256 MiB device copies (scale reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (cells shard across GPUs with "
              "no collective; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": a.numel(), "copies": 5}))
        return
    import pusch_eq as pe
    dev = torch.device("cuda:0")
    pe.init(0)
    if args.ceiling_only:
        print(json.dumps(hbm_ceiling(dev)), flush=True)
        return
    if args.interp_chain_only:
        print(json.dumps(interp_chain_bench(dev)), flush=True)
        print(json.dumps(interp_chain_bench(dev, nblk=768)), flush=True)
        return
    if args.serving_only:
        print(json.dumps(serving_bench(dev, args.serve_jobs)), flush=True)
        return
    if args.multicell_only:
        print(json.dumps(multicell_bench(dev, 20)), flush=True)
        return
    if args.srs_only:
        print(json.dumps(srs_bench(dev)), flush=True)
        return
    if args.sweep_only:
        sweep, rows = deployment_sweep(dev)
        if args.sweep_out:
            with open(args.sweep_out, "w") as f:
                json.dump({"summary": sweep, "rows": rows}, f, indent=1)
        print(json.dumps(sweep), flush=True)
        return
    if args.fronthaul_only:
        print(json.dumps(fronthaul_bench(dev, args.cells, 10)), flush=True)
        return
    if args.freq_interp_only:
        print(json.dumps(freq_interp_bench(dev, 10, 0.0)), flush=True)
        return
    if args.chain_only:  # two buffer sets alternate, as in the full run: every launch reads its inputs cold
        nsc = args.cells * NSC_CELL
        cbufs = []
        for k in range(2):
            H, Y, s2 = make_cells(args.cells, 1234 + k, dev)
            cbufs.append((H, s2, torch.empty((nsc, L, A), dtype=torch.complex64, device=dev),
                          torch.empty((nsc, L), dtype=torch.float32, device=dev),
                          torch.empty((nsc,), dtype=torch.uint8, device=dev), Y,
                          torch.empty((nsc, L, S, 2 * Q), dtype=torch.float32, device=dev)))

        def two(H, s2, C, ni, fl, Y, llr):
            pe.gram_mmse(H, s2, unbiased=True, C=C, noise_inv=ni, flags=fl)
            pe.coef_apply_demap(C, Y, ni, Q, prec=args.prec, out=llr)

        def one(H, s2, C, ni, fl, Y, llr):
            pe.mmse_apply_demap(H, s2, Y, Q, prec=args.prec, C=C, noise_inv=ni, flags=fl, out=llr)

        el2, _ = time_launches(two, cbufs, 40, 5)
        el1, _ = time_launches(one, cbufs, 40, 5)
        print(json.dumps({"subcarriers": nsc, "two_launches_ms": el2 / 40 * 1e3, "one_pass_ms": el1 / 40 * 1e3}),
              flush=True)
        return
    if args.mmse_only:
        el_m, el_m16 = mmse_stage_times(pe, [make_cells(args.cells, 1234 + k, dev) for k in range(2)], 40)
        print(json.dumps({"subcarriers": args.cells * NSC_CELL, "c64_ms": el_m / 40 * 1e3,
                          "c16_ms": el_m16 / 40 * 1e3}), flush=True)
        return
    n = args.cells
    nsc = n * NSC_CELL
    log("generating %d cells x 2 buffer sets" % n)
    sets = []
    cells = [make_cells(n, 1234 + k, dev) for k in range(2)]
    # the MMSE stage over the same batches, complex64 and binary16 coefficient output, the two sets alternating
    el_m, el_m16 = mmse_stage_times(pe, cells, 20)
    mmse_ms = {"c64": el_m / 20 * 1e3, "c16": el_m16 / 20 * 1e3}
    for k in range(2):
        H, Y, s2 = cells[k]
        cells[k] = None
        C, ni, fl = pe.gram_mmse(H, s2, unbiased=True)
        H0 = H[:NSC_CELL].clone() if k == 0 else None  # channels are not an input of the timed launch: keep one cell
        del H
        llr = torch.empty((nsc, L, S, 2 * Q), dtype=torch.float32, device=dev)
        sets.append({"H": H0, "Y": Y, "s2": s2, "C": C, "ni": ni, "fl": fl, "llr": llr})
    del cells
    torch.cuda.synchronize()
    n_fp64 = int(sum(int((s["fl"] != 0).sum()) for s in sets))

    # EVM at the benchmark shape (first cell of set 0): apply precision alone, and the whole chain
    s0 = sets[0]
    c1 = slice(0, NSC_CELL)
    yeq = pe.coef_apply(s0["C"][c1].contiguous(), s0["Y"][c1].contiguous(), prec=args.prec)
    yref_apply = torch.bmm(s0["C"][c1].to(torch.complex128),
                           s0["Y"][c1].to(torch.complex128).transpose(1, 2).contiguous())
    yref_chain = torch_reference_chain(s0["H"][c1], s0["s2"][c1], s0["Y"][c1])
    evm_apply = evm_db(yref_apply, yeq)
    evm_chain = evm_db(yref_chain, yeq)
    d = (yeq.to(torch.complex128) - yref_apply)
    acc = {"max_abs_err_vs_fp64": float(d.abs().max()),
           "cosine_vs_fp64": float((torch.vdot(yref_apply.flatten(), yeq.to(torch.complex128).flatten())).real
                                   / (yref_apply.abs().pow(2).sum() * yeq.to(torch.complex128).abs().pow(2).sum())
                                   .sqrt())}
    log("EVM (%s apply vs FP64) %.1f dB, chain vs FP64 %.1f dB" % (args.prec, evm_apply, evm_chain))

    def fused(C, Y, ni, llr, prec=args.prec):
        pe.coef_apply_demap(C, Y, ni, Q, prec=prec, out=llr)

    bufs = [(s["C"], s["Y"], s["ni"], s["llr"]) for s in sets]
    log("timing fused %s: %d warmup + %d steps of %d subcarriers" % (args.prec, args.warmup, args.steps, nsc))
    el, wall = time_launches(fused, bufs, args.steps, args.warmup)
    per_launch = el / args.steps
    bytes_launch = nsc * (L * A * 8 + S * A * 8 + L * S * 2 * Q * 4 + L * 4)
    flops_launch = 8.0 * nsc * L * S * A
    achieved_gbs = bytes_launch / per_launch / 1e9
    value = nsc / per_launch
    if args.headline_only:
        print(json.dumps({"subcarriers": nsc, "ms_per_launch": per_launch * 1e3, "hbm_gbs": achieved_gbs,
                          "bytes_per_launch": bytes_launch, "steps": args.steps, "warmup": args.warmup}), flush=True)
        return
    if args.codeword_only:
        print(json.dumps(codeword_bench(dev, sets[0]["C"], sets[0]["Y"], sets[0]["ni"], 10)), flush=True)
        return
    log("descrambling + codeword layout in the fused epilogue")
    codeword = codeword_bench(dev, sets[0]["C"], sets[0]["Y"], sets[0]["ni"])
    log("per-symbol %.1f us, codeword order %.1f us, de-interleaved %.1f us" % tuple(
        codeword[k]["ms_per_launch"] * 1e3 for k in ("per_symbol", "codeword_order", "deinterleaved")))

    # per-launch distribution (events between back-to-back launches, same alternation), cold = from HBM
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record()
    for i in range(args.steps):
        fused(*bufs[i % 2])
        evs[i + 1].record()
    torch.cuda.synchronize()
    per = sorted(evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps))
    launch_ms = {"p50": per[len(per) // 2], "p99": per[min(len(per) - 1, int(-(-99 * len(per) // 100)) - 1)],
                 "launches": len(per)}
    # warm: eight cells (159 MB) fit in the 256 MB Infinity Cache; the same launch repeated over them
    c8 = slice(0, 8 * NSC_CELL)
    wbuf = [(sets[0]["C"][c8], sets[0]["Y"][c8], sets[0]["ni"][c8], sets[0]["llr"][c8])]
    elw, _ = time_launches(fused, wbuf, 200, 20)
    warm = {"subcarriers_per_s": 8 * NSC_CELL / (elw / 200), "gbs": 8 * bytes_launch / n / (elw / 200) / 1e9,
            "ms_per_launch": elw / 200 * 1e3,
            "workload": "8 cells (159 MB) relaunched over the same buffers, resident in the Infinity Cache"}
    ceiling = hbm_ceiling(dev)

    # plain FP32 path on the same batches
    el32, _ = time_launches(lambda C, Y, ni, llr: fused(C, Y, ni, llr, "f32"), bufs, max(20, args.steps // 4),
                            5)
    per32 = el32 / max(20, args.steps // 4)
    # compensated variant
    elsp, _ = time_launches(lambda C, Y, ni, llr: fused(C, Y, ni, llr, "f16_split"), bufs,
                            max(20, args.steps // 4), 5)
    persp = elsp / max(20, args.steps // 4)
    # binary16 LLR wire format (the decoder input) on the same batches
    llr16 = [torch.empty((nsc, L, S, 2 * Q), dtype=torch.float16, device=dev) for _ in sets]
    bufs16 = [(s["C"], s["Y"], s["ni"], o) for s, o in zip(sets, llr16)]
    el16, _ = time_launches(lambda C, Y, ni, o: pe.coef_apply_demap_f16(C, Y, ni, Q, prec=args.prec, out=o),
                            bufs16, max(20, args.steps // 4), 5)
    per16 = el16 / max(20, args.steps // 4)
    bytes16 = nsc * (L * A * 8 + S * A * 8 + L * S * 2 * Q * 2 + L * 4)
    # coefficients kept in HBM as binary16 pairs + exponent (pe_coef_to_c16 outside the timed region), FP32 and
    # binary16 LLRs
    c16s = [pe.coef_to_c16(s["C"]) for s in sets]
    bufs_c16 = [(c16, ce, s["Y"], s["ni"], s["llr"]) for (c16, ce), s in zip(c16s, sets)]
    elc, _ = time_launches(lambda c16, ce, Y, ni, o: pe.coef_apply_demap_c16(c16, ce, Y, ni, Q, out=o), bufs_c16,
                           max(20, args.steps // 4), 5)
    perc = elc / max(20, args.steps // 4)
    bytes_c16 = nsc * (L * A * 4 + 4 + S * A * 8 + L * S * 2 * Q * 4 + L * 4)
    llr16 = [torch.empty((nsc, L, S, 2 * Q), dtype=torch.float16, device=dev) for _ in sets]
    bufs_c16w = [(c16, ce, s["Y"], s["ni"], o) for (c16, ce), s, o in zip(c16s, sets, llr16)]
    elcw, _ = time_launches(lambda c16, ce, Y, ni, o: pe.coef_apply_demap_c16(c16, ce, Y, ni, Q, llr_f16=True,
                                                                                out=o),
                            bufs_c16w, max(20, args.steps // 4), 5)
    percw = elcw / max(20, args.steps // 4)
    bytes_c16w = nsc * (L * A * 4 + 4 + S * A * 8 + L * S * 2 * Q * 2 + L * 4)
    del llr16, bufs16, bufs_c16, bufs_c16w, c16s

    # the whole chain from channel estimates: two launches with C in HBM, and the one-pass entry point
    # (pe_mmse_apply_demap: the MMSE writes a compact solve record, the apply forms the coefficients on chip)
    def chain2(Hf, s2, C, ni, fl, Y, llr):
        pe.gram_mmse(Hf, s2, unbiased=True, C=C, noise_inv=ni, flags=fl)
        pe.coef_apply_demap(C, Y, ni, Q, prec=args.prec, out=llr)

    def chain1(Hf, s2, C, ni, fl, Y, llr):
        pe.mmse_apply_demap(Hf, s2, Y, Q, prec=args.prec, C=C, noise_inv=ni, flags=fl, out=llr)

    hfull = [make_cells(n, 1234 + k, dev)[0] for k in range(2)]  # the channels of the two sets again
    cbufs = [(h, s["s2"], s["C"], s["ni"], s["fl"], s["Y"], s["llr"]) for h, s in zip(hfull, sets)]
    nch = max(20, args.steps // 4)
    el_c2, _ = time_launches(chain2, cbufs, nch, 5)
    el_c1, _ = time_launches(chain1, cbufs, nch, 5)
    ref_llr = sets[0]["llr"].clone()
    chain2(*cbufs[0])
    ref_llr.copy_(sets[0]["llr"])
    chain1(*cbufs[0])
    one_pass_identical = bool(torch.equal(ref_llr, sets[0]["llr"]))
    del ref_llr
    chain2(*cbufs[0])  # set 0's coefficients again (the one-pass call keeps its solve records in C)
    rec = L * (L + 2) * 4
    bytes_one_pass = nsc * (2 * A * L * 8 + 2 * rec + S * A * 8 + 2 * L * 4 + L * S * 2 * Q * 4 + 4 + 2)
    bytes_two = nsc * (2 * A * L * 8 + A * L * 8 + S * A * 8 + 2 * L * 4 + L * S * 2 * Q * 4 + 4 + 1)
    one_pass = {"ms_per_chain_two_launches": el_c2 / nch * 1e3, "ms_per_chain_one_pass": el_c1 / nch * 1e3,
                "subcarriers_per_s_one_pass": nsc / (el_c1 / nch),
                "speedup": el_c2 / el_c1, "llrs_identical": one_pass_identical,
                "bytes_per_chain_one_pass": bytes_one_pass, "bytes_per_chain_two_launches": bytes_two,
                "hbm_gbs_one_pass": bytes_one_pass / (el_c1 / nch) / 1e9,
                "workload": "H -> LLR for the whole batch: pe_gram_mmse + pe_coef_apply_demap against "
                            "pe_mmse_apply_demap (%s, FP32 LLRs)" % args.prec}
    log("chain from H: two launches %.1f us, one pass %.1f us (identical LLRs: %s)" % (
        el_c2 / nch * 1e6, el_c1 / nch * 1e6, one_pass_identical))
    del cbufs, hfull

    # single-cell slot latency: MMSE + fused apply/demap for one cell, one stream; as the caller sees it (from an
    # event recorded before the Python call, so the binding's argument checks and the launches are included) and
    # on the device alone (the call queued behind a spin kernel, so its launches wait in the queue when the start
    # event fires)
    cell = {k: (v[c1].contiguous() if torch.is_tensor(v) and v.shape[0] == nsc else v) for k, v in s0.items()}

    def two_calls():
        pe.gram_mmse(cell["H"], cell["s2"], unbiased=True, C=cell["C"], noise_inv=cell["ni"], flags=cell["fl"])
        fused(cell["C"], cell["Y"], cell["ni"], cell["llr"])

    def one_call():
        pe.mmse_apply_demap(cell["H"], cell["s2"], cell["Y"], Q, prec=args.prec, C=cell["C"], noise_inv=cell["ni"],
                            flags=cell["fl"], out=cell["llr"])

    lat = call_latency(two_calls, args.latency_iters, False)
    lat1 = call_latency(one_call, args.latency_iters, False)
    lat_dev = call_latency(two_calls, args.latency_iters, True)
    lat1_dev = call_latency(one_call, args.latency_iters, True)
    pe.gram_mmse(cell["H"], cell["s2"], unbiased=True, C=cell["C"], noise_inv=cell["ni"], flags=cell["fl"])
    pct = lambda p: percentile(lat, p)
    pct1 = lambda p: percentile(lat1, p)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%.0f s)" % args.cpu_seconds)
        cpu = cpu_baseline(cell["C"], cell["Y"], cell["ni"], args.cpu_seconds)

    multicell = None if args.no_multicell else multicell_bench(dev, 50)
    srs = None if args.no_srs else srs_bench(dev)
    fronthaul = fronthaul_bench(dev, n)
    operators = operators_bench(dev, n)
    log("standalone operators: gram %.1f GB/s, demap f16 %.1f GB/s, demap f32 %.1f GB/s" % (
        operators["gram_f32"]["hbm_gbs"], operators["llr_demap_f16"]["hbm_gbs"], operators["llr_demap_f32"]["hbm_gbs"]))
    log("fronthaul decompress %.1f GB/s" % fronthaul["hbm_gbs"])
    freq_interp = freq_interp_bench(dev)
    interp_chain = interp_chain_bench(dev)
    interp_chain["blocks_768"] = interp_chain_bench(dev, nblk=768)  # 3 blocks per compute unit: no partial round
    log("DMRS estimates to LLRs: fused %.1f us, separate calls %.1f us (interpolation %.1f, re-layout %.1f, "
        "MMSE + apply %.1f; identical: %s)" % (interp_chain["us_fused"], interp_chain["us_separate"],
                                                interp_chain["us_interp"], interp_chain["us_relayout"],
                                                interp_chain["us_mmse_apply_demap"],
                                                interp_chain["llrs_and_h_identical"]))
    log("frequency interpolation: bf16 %.1f GB/s, bf16 split %.1f GB/s, f32 %.1f GB/s (%.1f TFLOP/s)" % (
        freq_interp["bf16"]["hbm_gbs"], freq_interp["bf16_split"]["hbm_gbs"], freq_interp["f32"]["hbm_gbs"],
        freq_interp["f32"]["tflops"]))
    serving = serving_bench(dev, args.serve_jobs) if args.serve_jobs > 0 else None
    sweep = None
    if not args.no_sweep:
        sweep, rows = deployment_sweep(dev)
        log("deployment sweep: %d configs, bf16 at -60 dB vs f32 median speedup %.2fx; f16 HBM fraction cold "
            "median %.3f worst %.3f (A=%d PRB=%d L=%d)" % (
                sweep["configs"], sweep["bf16_at_bar_speedup_vs_f32_median"], sweep["hbm_frac_median_f16_cold"],
                sweep["hbm_frac_worst_f16_cold"]["frac"], sweep["hbm_frac_worst_f16_cold"]["A"],
                sweep["hbm_frac_worst_f16_cold"]["prb"], sweep["hbm_frac_worst_f16_cold"]["L"]))
        if args.sweep_out:
            with open(args.sweep_out, "w") as f:
                json.dump({"summary": sweep, "rows": rows}, f, indent=1)
    if multicell:
        log("multicell slot %.3f ms, %.1f GB/s" % (multicell["slot_ms"], multicell["hbm_gbs"]))
    if serving:
        log("serving capacity %.0f jobs/s" % serving["capacity_jobs_per_s"])

    workload = "fused coef-apply + 256-QAM max-log demap, %d cell-slots x 1584 SC x 8 layers x 11 sym x 64 ant " \
               "per launch (complex64 C from the library MMSE stage, complex64 Y_rx, FP32 LLR out)" % n
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "subcarriers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_launch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16" if args.prec.startswith("f16") else args.prec,
        "data": "synthetic (Rayleigh CN(0,1/64) channels, random 256-QAM, AWGN at 20 dB; generated on GPU)",
        "config": {"workload": workload, "cells_per_launch": n, "subcarriers_per_launch": nsc, "layers": L,
                   "symbols": S, "antennas": A, "qam": 256, "prec": args.prec, "accumulate": "f32",
                   "buffer_sets": 2, "bytes_per_set": bytes_launch},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": load_traffic("fused_%s" % args.prec)},
        "cpu_baseline": cpu,
        "hbm_gbs": achieved_gbs,
        "cold_hbm": {"note": "value and roofline are cold: two 637 MB buffer sets alternate, beyond the Infinity "
                             "Cache", "launch_ms": launch_ms},
        "warm_infinity_cache": warm,
        "measured_ceiling": ceiling,
        "gflops": flops_launch / per_launch / 1e9,
        "evm_db": {"apply_vs_fp64": evm_apply, "chain_vs_fp64": evm_chain},
        "accuracy": acc,
        "fp64_fallback_subcarriers": n_fp64,
        "plain_f32": {"value": nsc / per32, "ms_per_step": per32 * 1e3,
                      "hbm_gbs": bytes_launch / per32 / 1e9},
        "f16_split": {"value": nsc / persp, "ms_per_step": persp * 1e3, "hbm_gbs": bytes_launch / persp / 1e9},
        "llr_codeword": codeword,
        "llr_f16_wire": {"value": nsc / per16, "ms_per_step": per16 * 1e3, "hbm_gbs": bytes16 / per16 / 1e9,
                         "bytes_per_launch": bytes16,
                         "workload": "same launch, LLRs written as binary16 (pe_coef_apply_demap_f16)"},
        "mmse_stage": {"ms_per_launch": mmse_ms, "subcarriers": nsc,
                       "hbm_gbs_c64": nsc * (A * L * 8 + 4 + L * A * 8 + L * 4 + 1) / (mmse_ms["c64"] * 1e-3) / 1e9,
                       "hbm_gbs_c16": nsc * (A * L * 8 + 4 + L * A * 4 + 4 + L * 4 + 1) / (mmse_ms["c16"] * 1e-3)
                       / 1e9,
                       "workload": "pe_gram_mmse / pe_gram_mmse_c16 over the batch (H read, C written), two "
                                   "batches alternating (H read cold)"},
        "c16_coefficients": {"value": nsc / perc, "ms_per_step": perc * 1e3, "hbm_gbs": bytes_c16 / perc / 1e9,
                             "bytes_per_launch": bytes_c16,
                             "workload": "same batches, coefficients held as binary16 pairs + exponent "
                                         "(pe_coef_apply_demap_c16), FP32 LLRs"},
        "c16_coefficients_llr_f16": {"value": nsc / percw, "ms_per_step": percw * 1e3,
                                     "hbm_gbs": bytes_c16w / percw / 1e9, "bytes_per_launch": bytes_c16w,
                                     "workload": "binary16 coefficients and binary16 LLRs"},
        "slot_latency_ms": {"p50": pct(50), "p99": pct(99), "max": lat[-1] if lat else None, "iters": len(lat),
                            "workload": "one cell: pe_gram_mmse + fused apply/demap, events on one stream around "
                                        "the Python calls (argument checks and launches included)",
                            "one_pass": {"p50": pct1(50), "p99": pct1(99), "max": lat1[-1] if lat1 else None,
                                         "workload": "one cell: pe_mmse_apply_demap"},
                            "device": {"p50": percentile(lat_dev, 50), "p99": percentile(lat_dev, 99),
                                       "one_pass_p50": percentile(lat1_dev, 50),
                                       "one_pass_p99": percentile(lat1_dev, 99),
                                       "workload": "the same calls queued behind a spin kernel: device time from "
                                                   "the first launch's start to the last one's end"}},
        "chain_from_channel_estimates": one_pass,
        "multicell_slot": multicell,
        "serving": serving,
        "srs_filter": srs,
        "fronthaul_bfp_decompress": fronthaul,
        "interp_mmse_chain": interp_chain,
        "freq_interp": freq_interp,
        "standalone_operators": operators,
        "deployment_sweep": sweep,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
