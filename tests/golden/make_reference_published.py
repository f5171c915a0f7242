"""Writes reference_published.json: the outputs the reference itself recorded for this path (A100 runs of its own
programs), read from its result file and README. Needs the reference checkout; the JSON it writes is committed
and is what the tests read (the reference is not present where the GPU tests run).

    python tests/golden/make_reference_published.py /path/to/aerial-mimo-tc-optimization
"""
import json
import os
import re
import sys


def main(ref):
    rigor_path = os.path.join(ref, "results", "coef_rigor_results.txt")
    readme_path = os.path.join(ref, "README.md")
    rigor = open(rigor_path).read().splitlines()
    readme = open(readme_path).read().splitlines()
    out = {}
    for i, ln in enumerate(rigor, 1):
        m = re.search(r"EVM\(V2 vs V1\) = ([0-9.]+)%\s+\((-?[0-9.]+) dB\)\s+max_rel_err = ([0-9.]+)%", ln)
        if m:
            out["coef_rigor_evm_pct"] = float(m.group(1))
            out["coef_rigor_evm_db"] = float(m.group(2))
            out["coef_rigor_max_rel_err_pct"] = float(m.group(3))
            out["coef_rigor_source"] = "results/coef_rigor_results.txt:%d" % i
            out["coef_rigor_inputs"] = "srand(7), ((float)rand()/RAND_MAX - 0.5f) * 1.4f, C then Y (coef_rigor.cu:101-104)"
            out["coef_rigor_shape"] = [1584, 8, 11, 64]
    for i, ln in enumerate(readme, 1):
        m = re.search(r"Cosine similarity\s*:\s*([0-9.]+)", ln)
        if m:
            out["readme_cosine"] = float(m.group(1))
            out["readme_cosine_source"] = "README.md:%d" % i
        m = re.search(r"Max abs error\s*:\s*([0-9.]+)", ln)
        if m:
            out["readme_max_abs"] = float(m.group(1))
            out["readme_max_abs_source"] = "README.md:%d" % i
    out["readme_inputs"] = "srand(42), (float)rand()/RAND_MAX - 0.5f, C then Y (aerial_coef_wmma.cu:383-387)"
    need = ("coef_rigor_evm_db", "readme_cosine", "readme_max_abs")
    missing = [k for k in need if k not in out]
    if missing:
        raise SystemExit("not found in the reference files: %s" % missing)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_published.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
