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


def cpu_baseline(C, Y, ni, seconds):
    """The oracle's OpenMP restatement of the scalar coef-apply loop + FP32 demapper on one cell, repeated
    for about `seconds`, on the host cores (OMP_NUM_THREADS or all)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pe_oracle as po
    Cn, Yn, nin = C.cpu().numpy(), Y.cpu().numpy(), ni.cpu().numpy()
    out = np.empty((NSC_CELL, L, S, 2 * Q), dtype=np.float32)
    _, used = po.coef_apply_demap_f32_omp(Cn, Yn, nin, Q, 0, out)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        po.coef_apply_demap_f32_omp(Cn, Yn, nin, Q, 0, out)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": reps * NSC_CELL / el, "unit": "subcarriers/s", "cores": int(used), "kind": "port",
            "sample": "%d repetitions of one 1584-subcarrier cell (8 layers, 11 symbols, 64 antennas, 256-QAM), "
                      "scalar antenna-serial FP32 chain + FP32 demapper, OpenMP over (subcarrier, layer), "
                      "%.1f s" % (reps, el)}


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
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=32, help="cell-slots per launch")
    ap.add_argument("--prec", default="f16")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-iters", type=int, default=300)
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
    n = args.cells
    nsc = n * NSC_CELL
    log("generating %d cells x 2 buffer sets" % n)
    sets = []
    for k in range(2):
        H, Y, s2 = make_cells(n, 1234 + k, dev)
        C, ni, fl = pe.gram_mmse(H, s2, unbiased=True)
        llr = torch.empty((nsc, L, S, 2 * Q), dtype=torch.float32, device=dev)
        sets.append({"H": H, "Y": Y, "s2": s2, "C": C, "ni": ni, "fl": fl, "llr": llr})
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

    # plain FP32 path on the same batches
    el32, _ = time_launches(lambda C, Y, ni, llr: fused(C, Y, ni, llr, "f32"), bufs, max(20, args.steps // 4),
                            5)
    per32 = el32 / max(20, args.steps // 4)
    # compensated variant
    elsp, _ = time_launches(lambda C, Y, ni, llr: fused(C, Y, ni, llr, "f16_split"), bufs,
                            max(20, args.steps // 4), 5)
    persp = elsp / max(20, args.steps // 4)

    # single-cell slot latency: MMSE + fused apply/demap for one cell, one stream
    cell = {k: (v[c1].contiguous() if torch.is_tensor(v) and v.shape[0] == nsc else v) for k, v in s0.items()}
    lat = []
    for i in range(args.latency_iters + 20):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        pe.gram_mmse(cell["H"], cell["s2"], unbiased=True, C=cell["C"], noise_inv=cell["ni"], flags=cell["fl"])
        fused(cell["C"], cell["Y"], cell["ni"], cell["llr"])
        b.record()
        b.synchronize()
        if i >= 20:
            lat.append(a.elapsed_time(b))
    lat.sort()
    pct = lambda p: lat[min(len(lat) - 1, max(0, int(-(-p * len(lat) // 100)) - 1))] if lat else None

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%.0f s)" % args.cpu_seconds)
        cpu = cpu_baseline(cell["C"], cell["Y"], cell["ni"], args.cpu_seconds)

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
        "gflops": flops_launch / per_launch / 1e9,
        "evm_db": {"apply_vs_fp64": evm_apply, "chain_vs_fp64": evm_chain},
        "fp64_fallback_subcarriers": n_fp64,
        "plain_f32": {"value": nsc / per32, "ms_per_step": per32 * 1e3,
                      "hbm_gbs": bytes_launch / per32 / 1e9},
        "f16_split": {"value": nsc / persp, "ms_per_step": persp * 1e3, "hbm_gbs": bytes_launch / persp / 1e9},
        "slot_latency_ms": {"p50": pct(50), "p99": pct(99), "max": lat[-1] if lat else None, "iters": len(lat),
                            "workload": "one cell: pe_gram_mmse + fused apply/demap, events on one stream"},
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
