"""Autotuner for the tile table of libpusch_eq.so (run on the target GPU).

For every shape class (layers L, symbols S, antennas A, precision) it times the candidate launch shapes of
the fused coefficient-apply + demap kernel and of the FP32 kernel, and the lanes-per-subcarrier choice of the
MMSE kernel at a single-cell size (latency) and a batch size (throughput); the fastest candidate of each
class becomes a line of tune_table.txt beside the library, which the library reads at first use.

    python autotune.py [--out tune_table.txt] [--quick]

Candidates (see pe_tune_load in include/pusch_eq.h):
    MFMA apply: waves per workgroup {1, 2, 4} x subcarriers per wave {1, 2}
    FP32 apply: antennas per LDS chunk {8, 16}
    MMSE:       lanes per subcarrier {8, 16}
Timing: HIP events around 20 launches after 5 warm-up launches, alternating two buffer sets whose sum exceeds
the 256 MB Infinity Cache, so launches stream from HBM as in production.
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
        cands = [(kc, 1, 1) for kc in (8, 16)]
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


def tune_mmse(pe, dev, L, A, nsc):
    import torch
    sets = []
    for k in range(2):
        g = torch.Generator(device=dev).manual_seed(10 + k)
        H = torch.randn((nsc, A, L), dtype=torch.complex64, device=dev, generator=g) * (1.0 / A) ** 0.5
        s2 = torch.full((nsc,), 0.01, device=dev)
        C = torch.empty((nsc, L, A), dtype=torch.complex64, device=dev)
        ni = torch.empty((nsc, L), device=dev)
        fl = torch.empty((nsc,), dtype=torch.uint8, device=dev)
        sets.append((H, s2, C, ni, fl))
    best, res = None, []
    clock = timed_single if nsc <= 4096 else timed
    for lanes in (8, 16):
        with_table(pe, ["mmse %d %d -1 %d" % (L, A, lanes)])
        t = clock(lambda H, s2, C, ni, fl: pe.gram_mmse(H, s2, True, 0.0, C, ni, fl), sets)
        res.append({"lanes": lanes, "us": t * 1e6})
        if best is None or t < best[0]:
            best = (t, lanes)
    del sets
    torch.cuda.empty_cache()
    return best, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune_table.txt"))
    ap.add_argument("--report", default=None, help="write the per-candidate timings as JSON")
    ap.add_argument("--quick", action="store_true", help="A in {16, 64}, L = 8, f16 only")
    args = ap.parse_args()
    import torch
    import pusch_eq as pe
    dev = torch.device("cuda:0")
    pe.init(0)
    As = (16, 64) if args.quick else (8, 16, 32, 64, 128)
    Ls = (8,) if args.quick else (2, 4, 8)
    precs = ("f16",) if args.quick else ("f16", "f16_split", "bf16", "bf16_split", "f32")
    S = 11
    lines = ["# tile table written by autotune.py on %s" % torch.cuda.get_device_name(0),
             "# coef <L> <S> <A> <prec> <f32_kc> <mfma_wpb> <mfma_spw>; mmse <L> <A> <nSC_max> <lanes>"]
    report = {"apply": [], "mmse": []}
    for A, L, prec in itertools.product(As, Ls, precs):
        (t, kc, w, s), res, nsc = tune_apply(pe, dev, L, S, A, prec, 600e6)
        lines.append("coef %d %d %d %d %d %d %d" % (L, S, A, PRECS[prec], kc, w, s))
        report["apply"].append({"L": L, "S": S, "A": A, "prec": prec, "nsc": nsc, "candidates": res})
        print("apply L=%d A=%3d %-10s best kc=%d wpb=%d spw=%d %.1f us" % (L, A, prec, kc, w, s, t * 1e6), flush=True)
    for A, L in itertools.product(As, Ls):
        for nsc, bucket in ((1584, 4096), (50688, -1)):
            (t, lanes), res = tune_mmse(pe, dev, L, A, nsc)
            lines.append("mmse %d %d %d %d" % (L, A, bucket, lanes))
            report["mmse"].append({"L": L, "A": A, "nsc": nsc, "candidates": res})
            print("mmse  L=%d A=%3d nsc=%5d best lanes=%d %.1f us" % (L, A, nsc, lanes, t * 1e6), flush=True)
    lines.append("coef -1 -1 -1 -1 8 1 1")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    pe.lib().pe_tune_load(args.out.encode())
    if args.report:
        json.dump(report, open(args.report, "w"), indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
