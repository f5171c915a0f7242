"""Autotuner for the tile table of libpusch_eq.so (run on the target GPU).

For every shape class (layers L, symbols S, antennas A, precision) it times the candidate launch shapes of
the fused coefficient-apply + demap kernel and of the FP32 kernel, and the lanes-per-subcarrier choice of the
MMSE kernel at a single-cell size (latency) and a batch size (throughput); the fastest candidate of each
class becomes a line of tune_table.txt beside the library, which the library reads at first use.

    python autotune.py [--out tune_table.txt] [--quick]

Candidates (see pe_tune_load in include/pusch_eq.h):
    MFMA apply: waves per workgroup {1, 2, 4} x subcarriers per wave {1, 2}
    FP32 apply: antennas per LDS chunk {4, 8, 16}
    MMSE:       lanes per subcarrier {4, 8, 16} (one per shape: it fixes the summation order) x waves per
                workgroup {1, 2, 4, 8} x {one Gram round per wave, two, one with the antenna rows kept in
                registers} (tile = rounds x waves x 64 / lanes <= 128 subcarriers) per batch bucket (single
                cell, large batch)
    solve:      tile of pe_mmse_apply_demap's solve kernel, (subcarriers, waves) in {(8,1) .. (128,8)}, per
                (L, A) that takes solve records (L >= 4, 32 <= A <= 64) and batch bucket (one cell, 32 cells),
                timed on the whole one-pass chain
    slot:       MMSE (tile, waves, rows kept) in {(4,1,0) .. (32,8,1)} (24-cell plans)
Timing: HIP events around 20 launches after 5 warm-up launches, alternating two buffer sets (for the apply
their sum exceeds the 256 MB Infinity Cache, so launches stream from HBM as in production); single-cell MMSE
launches are timed one at a time (median of 15).
"""
import argparse
import itertools
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PRECS = {"f32": 0, "f16": 1, "bf16": 2, "f16_split": 3, "bf16_split": 4}


def timed(fn, sets, iters=20, warm=5):
    import torch
    for i in range(warm):
        fn(*sets[i % 2])
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for i in range(iters):
        fn(*sets[i % 2])
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / iters * 1e-3


def timed_single(fn, sets, iters=15, warm=3):
    """Median device time of one launch on an otherwise idle GPU. A short spin kernel is queued first so that
    the launch is already in the queue when the start event executes: host launch overhead is excluded."""
    import torch
    for i in range(warm):
        fn(*sets[i % 2])
    torch.cuda.synchronize()
    ts = []
    for i in range(iters):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(200000)
        a.record()
        fn(*sets[i % 2])
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    ts.sort()
    return ts[len(ts) // 2]


def with_table(pe, lines):
    fd, path = tempfile.mkstemp(suffix=".txt")
    with os.fdopen(fd, "w") as f:
        f.write("\n".join(lines) + "\n")
    if pe.lib().pe_tune_load(path.encode()) < 0:
        raise RuntimeError("cannot load candidate table")
    os.unlink(path)


def tune_apply(pe, dev, L, S, A, prec, target_bytes):
    import torch
    per_sc = 8 * (L * A + S * A) + 4 * L * S * 8 + 4 * L
    nsc = max(1584, int(target_bytes // per_sc) // 1584 * 1584)
    sets = []
    for k in range(2):
        g = torch.Generator(device=dev).manual_seed(k)
        C = torch.randn((nsc, L, A), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A) ** 0.5
        Y = torch.randn((nsc, S, A), dtype=torch.complex64, device=dev, generator=g)
        ni = torch.full((nsc, L), 40.0, device=dev)
        llr = torch.empty((nsc, L, S, 8), device=dev)
        sets.append((C, Y, ni, llr))
    if prec == "f32":
        cands = [(kc, 1, 1) for kc in (4, 8, 16)]
    else:
        cands = [(8, w, s) for w, s in itertools.product((1, 2, 4), (1, 2))]
    best = None
    res = []
    for kc, w, s in cands:
        with_table(pe, ["coef %d %d %d %d %d %d %d" % (L, S, A, PRECS[prec], kc, w, s)])
        t = timed(lambda C, Y, ni, llr: pe.coef_apply_demap(C, Y, ni, 4, prec=prec, out=llr), sets)
        res.append({"kc": kc, "wpb": w, "spw": s, "us": t * 1e6, "gbs": nsc * per_sc / t / 1e9})
        if best is None or t < best[0]:
            best = (t, kc, w, s)
    del sets
    torch.cuda.empty_cache()
    return best, res, nsc


def mmse_sets(dev, L, A, nsc):
    import torch
    sets = []
    for k in range(2):
        g = torch.Generator(device=dev).manual_seed(10 + k)
        H = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A) ** 0.5
        s2 = torch.full((nsc,), 0.03, device=dev)
        C = torch.empty((nsc, L, A), dtype=torch.complex64, device=dev)
        ni = torch.empty((nsc, L), device=dev)
        fl = torch.empty((nsc,), dtype=torch.uint8, device=dev)
        sets.append((H, s2, C, ni, fl))
    return sets


def tune_mmse(pe, dev, L, A, base):
    """Lanes per subcarrier P (fixes the summation order, so one choice per shape) and subcarriers per tile T
    per batch bucket. P is the one with the smallest sum over buckets of time / best time of the bucket."""
    import torch
    buckets = ((1584, 4096), (50688, -1))
    res = {}
    for P in (4, 8, 16):
        for nsc, bucket in buckets:
            sets = mmse_sets(dev, L, A, nsc)
            clock = timed_single if nsc <= 4096 else timed
            for W in (1, 2, 4, 8):
                for R, K in ((1, 0), (2, 0), (1, 1)):  # Gram rounds per wave, rows kept in registers
                    T = R * W * 64 // P
                    if T > 128 or (K and A > 4 * P):
                        continue
                    with_table(pe, ["mmse %d %d -1 %d %d %d %d" % (L, A, P, T, W, K)] + base)
                    t = clock(lambda H, s2, C, ni, fl: pe.gram_mmse(H, s2, True, 0.0, C, ni, fl), sets)
                    res[(P, nsc, T, W, K)] = t
            del sets
            torch.cuda.empty_cache()
    best_b = {nsc: min(t for k, t in res.items() if k[1] == nsc) for nsc, _ in buckets}
    score = {}
    for P in (4, 8, 16):
        score[P] = sum(min(t for k, t in res.items() if k[0] == P and k[1] == nsc) / best_b[nsc]
                       for nsc, _ in buckets)
    P = min(score, key=score.get)
    lines, chosen = [], []
    for nsc, bucket in buckets:
        (T, W, K), t = min((((k[2], k[3], k[4]), t) for k, t in res.items() if k[0] == P and k[1] == nsc),
                           key=lambda x: x[1])
        lines.append("mmse %d %d %d %d %d %d %d" % (L, A, bucket, P, T, W, K))
        chosen.append((nsc, T, W, K, t))
    report = [{"P": k[0], "nsc": k[1], "T": k[2], "W": k[3], "keep": k[4], "us": t * 1e6}
              for k, t in sorted(res.items())]
    return lines, P, chosen, report


def slot_cells(dev, n, seed):
    """Cells of the multi-cell batch: PRB {24,52,106,273}, antennas {16,32,64}, layers {2,4,8}, QAM 16-256."""
    import numpy as np
    import torch
    import pusch_eq as pe
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=dev).manual_seed(seed)
    cells, keep = [], []
    for _ in range(n):
        prb, a, l, q = (int(rng.choice(v)) for v in ([24, 52, 106, 273], [16, 32, 64], [2, 4, 8], [2, 3, 4]))
        nsc = 12 * prb
        H = torch.randn((nsc, a, l), dtype=torch.complex64, device=dev, generator=g) * (1.0 / a) ** 0.5
        s2 = torch.full((nsc,), 10 ** (-float(rng.uniform(5, 25)) / 10), device=dev)
        Y = torch.randn((nsc, 11, a), dtype=torch.complex64, device=dev, generator=g)
        C = torch.empty((nsc, l, a), dtype=torch.complex64, device=dev)
        ni = torch.empty((nsc, l), device=dev)
        llr = torch.empty((nsc, l, 11, 2 * q), device=dev)
        keep.append((H, s2, Y, C, ni, llr))
        cells.append(pe.cell_desc(C, ni, Y, llr, q, "f16", H, s2))
    return cells, keep


def tune_solve(pe, dev, L, A, base):
    """Subcarriers and waves per workgroup of the one-pass chain's solve kernel, timed on the whole chain
    (pe_mmse_apply_demap, 11 symbols, QPSK, two input sets alternating) per batch bucket: one cell of 1584
    subcarriers (single launches, latency) and 32 cells (throughput). Returns the table lines (one per bucket)."""
    import torch
    S = 11
    lines, report = [], []
    for nsc, bucket in ((1584, 4096), (32 * 1584, -1)):
        sets = []
        for k in range(2):
            g = torch.Generator(device=dev).manual_seed(20 + k)
            H = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A) ** 0.5
            s2 = torch.full((nsc,), 0.03, device=dev)
            Y = torch.randn((nsc, S, A), dtype=torch.complex64, device=dev, generator=g)
            sets.append((H, s2, Y))
        clock = timed_single if nsc <= 4096 else timed
        res = []
        for T, W in ((8, 1), (8, 2), (16, 2), (16, 4), (32, 4), (32, 8), (64, 4), (64, 8), (128, 4), (128, 8)):
            with_table(pe, ["solve %d %d %d %d %d" % (L, A, bucket, T, W)] + base)
            t = clock(lambda H, s2, Y: pe.mmse_apply_demap(H, s2, Y, 1, prec="f16"), sets)
            res.append({"nsc": nsc, "T": T, "W": W, "us": t * 1e6})
        del sets
        torch.cuda.empty_cache()
        best = min(res, key=lambda r: r["us"])
        lines.append("solve %d %d %d %d %d" % (L, A, bucket, best["T"], best["W"]))
        report.extend(res)
    return lines, report


def tune_slot(pe, dev, base):
    """Subcarriers and waves per MMSE tile of the multi-cell executor, on two alternating 24-cell slot plans
    replayed as graphs."""
    import torch
    side = torch.cuda.Stream()  # a created stream, as a receiver runs its slots
    torch.cuda.set_stream(side)
    sets = [slot_cells(dev, 24, 77 + k) for k in range(2)]
    res = []
    for mt, mw, mk in ((4, 1, 0), (8, 1, 0), (8, 2, 0), (16, 1, 0), (16, 2, 0), (16, 4, 0), (32, 2, 0), (32, 4, 0),
                       (32, 8, 0), (64, 4, 0), (64, 8, 0), (128, 8, 0), (8, 2, 1), (16, 4, 1), (32, 8, 1)):
        with_table(pe, ["slot %d %d %d" % (mt, mw, mk)] + base)
        plans = [pe.SlotPlan(c) for c, _ in sets]
        t = timed(lambda pl: pl.graph_launch(), [(plans[0],), (plans[1],)], iters=30, warm=6)
        for pl in plans:
            pl.close()
        res.append({"mmse_tile": mt, "mmse_waves": mw, "mmse_keep": mk, "us": t * 1e6})
    torch.cuda.synchronize()
    torch.cuda.set_stream(torch.cuda.default_stream())
    best = min(res, key=lambda r: r["us"])
    return "slot %d %d %d" % (best["mmse_tile"], best["mmse_waves"], best["mmse_keep"]), best, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune_table.txt"))
    ap.add_argument("--report", default=None, help="write the per-candidate timings as JSON")
    ap.add_argument("--quick", action="store_true", help="A in {16, 64}, L = 8, f16 only")
    ap.add_argument("--precs", default=None, help="comma-separated precisions to tune (default: all)")
    ap.add_argument("--keep-coef", action="store_true",
                    help="keep the coef lines of the existing table (--out) and tune only MMSE and slot tiles")
    args = ap.parse_args()
    import torch
    import pusch_eq as pe
    dev = torch.device("cuda:0")
    pe.init(0)
    As = (16, 64) if args.quick else (8, 16, 32, 64, 128)
    Ls = (8,) if args.quick else (2, 4, 8)
    precs = ("f16",) if args.quick else ("f16", "f16_split", "bf16", "bf16_split", "f32")
    if args.precs:
        precs = tuple(args.precs.split(","))
    S = 11
    lines = ["# tile table written by autotune.py on %s" % torch.cuda.get_device_name(0),
             "# coef <L> <S> <A> <prec> <f32_kc> <mfma_wpb> <mfma_spw>",
             "# mmse <L> <A> <nSC_max> <lanes per subcarrier> <subcarriers per tile>",
             "# solve <L> <A> <nSC_max> <subcarriers per tile> <waves>",
             "# slot <mmse subcarriers per tile> <mmse waves> <keep rows>"]
    report = {"apply": [], "mmse": [], "solve": [], "slot": []}
    kept = []
    if args.keep_coef and os.path.exists(args.out):
        kept = [ln.strip() for ln in open(args.out) if ln.startswith("coef ")]
        lines.extend(kept)
    for A, L, prec in ([] if kept else itertools.product(As, Ls, precs)):
        (t, kc, w, s), res, nsc = tune_apply(pe, dev, L, S, A, prec, 600e6)
        lines.append("coef %d %d %d %d %d %d %d" % (L, S, A, PRECS[prec], kc, w, s))
        report["apply"].append({"L": L, "S": S, "A": A, "prec": prec, "nsc": nsc, "candidates": res})
        print("apply L=%d A=%3d %-10s best kc=%d wpb=%d spw=%d %.1f us" % (L, A, prec, kc, w, s, t * 1e6), flush=True)
    if not kept:
        lines.append("coef -1 -1 -1 -1 8 1 1")
    for A, L in itertools.product(As, Ls):
        mlines, P, chosen, rep = tune_mmse(pe, dev, L, A, lines)
        lines.extend(mlines)
        report["mmse"].append({"L": L, "A": A, "P": P, "candidates": rep})
        print("mmse  L=%d A=%3d lanes=%2d  %s" % (L, A, P, "  ".join(
            "nsc %d: T=%d W=%d keep=%d %.1f us" % (n, T, W, K, t * 1e6) for n, T, W, K, t in chosen)), flush=True)
    for A, L in itertools.product(As, Ls):
        if L < 4 or not 32 <= A <= 64:  # shapes without solve records run the coefficient kernels
            continue
        sl, rep = tune_solve(pe, dev, L, A, lines)
        lines.extend(sl)
        report["solve"].append({"L": L, "A": A, "candidates": rep})
        print("solve L=%d A=%3d  %s" % (L, A, "  ".join(sl)), flush=True)
    sline, best, rep = tune_slot(pe, dev, lines)
    lines.append(sline)
    report["slot"] = rep
    print("slot  %s  %.1f us per 24-cell slot" % (sline, best["us"]), flush=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    pe.lib().pe_tune_load(args.out.encode())
    if args.report:
        json.dump(report, open(args.report, "w"), indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
