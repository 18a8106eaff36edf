#!/usr/bin/env python3
"""Extract golden vectors from the reference repository's own result files.

Run once in a container where the reference is mounted (default
/root/reference); it writes small JSON fixtures next to this script.  The
fixtures are data (inputs and recorded outputs) only.

Sources (paths relative to the reference root):
  distance_distribution/src/solutions.h:18-114       golden vectors (AN, Ext. Hamming)
  distance_distribution/results/As/*.csv             exhaustive AN runs
  distance_distribution/results/runtimes/*/*.csv     exhaustive + grid runs
  distance_distribution/results/converge/geom/*.csv  deterministic grid1d runs
  distance_distribution/results/GridMs/*grid1d*.csv  deterministic grid1d runs
  distance_distribution/results/superAs/*.txt        super-A sweep summaries
  distance_distribution/legacy/print_codeword_matrix.out   k=5 full histograms
  results/runtimes/4/ancoding_gpu_A61_n16_h6.csv, results/converge/geom/
    ancoding_grid1d_gpu_m1001_A61_n16_h6.csv,
  results/runtimes/4/ancoding_cpu_A61_n{8,16,24}_h6.csv
                                                     CSV text (row and summary lines)
  distance_distribution/legacy/hamming_count_undetectable_errors.cpp
                                                     (40,33) and (52,45) weight tables

A run file is kept only when its integer column is complete (the histogram
sum matches the expected total); truncated runs are listed as excluded.
"""
import glob
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
DD = os.path.join(REF, "distance_distribution")
OUT = os.path.dirname(os.path.abspath(__file__))


def parse_solutions():
    txt = open(os.path.join(DD, "src", "solutions.h")).read()
    out = {}
    for m in re.finditer(r"const\s+uint128_t\s+(\w+)\s*\[\]\s*=\s*\{([^}]*)\}", txt):
        vals = [int(x.strip().rstrip("ull").rstrip("u")) for x in m.group(2).split(",") if x.strip()]
        out[m.group(1)] = vals
    return out


def parse_run(path):
    """Histogram rows 'bin, count, prob, base[, err, err]' up to the first blank line.
    Returns (counts, exact) where exact is False if a count was printed as a float."""
    counts, exact = [], True
    stats = {}
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip():
        parts = [p.strip() for p in lines[i].split(",")]
        b = int(parts[0])
        assert b == len(counts), (path, b)
        c = parts[1]
        if re.fullmatch(r"\d+", c):
            counts.append(int(c))
        else:
            counts.append(None)
            exact = False
        i += 1
    for line in lines[i:]:
        m = re.match(r'\s*\d+,"([^"]+)",([^,]+),', line)
        if m:
            stats[m.group(1)] = float(m.group(2))
    return counts, exact, stats


def name_params(path):
    b = os.path.basename(path)
    d = {}
    for key, pat in (("A", r"_A(\d+)"), ("n", r"_n(\d+)"), ("h", r"_h(\d+)"), ("m", r"_m(\d+)")):
        m = re.search(pat, b)
        if m:
            d[key] = int(m.group(1))
    return d


def exhaustive_runs():
    files = sorted(glob.glob(os.path.join(DD, "results", "As", "ancoding_gpu_*.csv")))
    files += sorted(glob.glob(os.path.join(DD, "results", "runtimes", "*", "ancoding_*_A61_n*_h6.csv")))
    keep, excluded = [], []
    for p in files:
        if "grid" in os.path.basename(p):
            continue
        prm = name_params(p)
        counts, exact, stats = parse_run(p)
        n = prm["n"]
        ok = exact and None not in counts and sum(counts) == 4 ** n
        rel = os.path.relpath(p, REF)
        if ok:
            keep.append(dict(src=rel, A=prm["A"], n=n, h=prm["h"], counts=counts,
                             kernel_s=stats.get("Kernel Runtime"), total_s=stats.get("Total Runtime")))
        else:
            excluded.append(rel)
    return keep, excluded


def grid_runs():
    pats = ["results/converge/geom/ancoding_grid1d_*.csv", "results/GridMs/ancoding_grid1d_*.csv",
            "results/runtimes/*/ancoding_grid1d_*.csv", "results/As/ancoding_grid1d_*.csv"]
    files = []
    for pat in pats:
        files += sorted(glob.glob(os.path.join(DD, pat)))
    runs = []
    for p in files:
        prm = name_params(p)
        counts, exact, stats = parse_run(p)
        if not exact or None in counts:
            continue
        runs.append(dict(src=os.path.relpath(p, REF), A=prm["A"], n=prm["n"], h=prm["h"], m=prm["m"],
                         counts=counts, kernel_s=stats.get("Kernel Runtime"),
                         four_gpu="4gpu" in os.path.basename(p)))
    return runs


def supera():
    out = []
    for p in sorted(glob.glob(os.path.join(DD, "results", "superAs", "superA_n*_h*.txt"))):
        rows, summary = [], None
        for line in open(p):
            m = re.match(r"(\d+): c\[(\d+)\] = (\d+)", line.strip())
            if m:
                rows.append([int(m.group(1)), int(m.group(2)), int(m.group(3))])
            elif line.startswith("n,"):
                summary = line.strip()
        out.append(dict(src=os.path.relpath(p, REF), rows=rows, summary=summary))
    return out


def codeword_matrices():
    p = os.path.join(DD, "legacy", "print_codeword_matrix.out")
    txt = open(p).read()
    res = []
    for block in re.finditer(r"Settings: A=(\d+)\. n=(\d+)\. k=\|A\|=(\d+)(.*?)total edges: (\d+)", txt, re.S):
        A, nn, ka = int(block.group(1)), int(block.group(2)), int(block.group(3))
        hist = {}
        for m in re.finditer(r"^\s*(\d+): (\d+)$", block.group(4), re.M):
            hist[int(m.group(1))] = int(m.group(2))
        h = [hist.get(i, 0) for i in range(max(hist) + 1)]
        # bin 0 holds the 2^k diagonal entries of the ordered-pair matrix, which
        # fixes the data width (the printed "k=|A|" label is not the data width)
        k = h[0].bit_length() - 1
        res.append(dict(src=os.path.relpath(p, REF), A=A, codeword_bits=nn, k=k,
                        hist=h, total=int(block.group(5))))
    return res


def legacy_hamming_weights():
    """The weight distributions of the non-extended (39,33) and (51,45) Hamming
    codes that the legacy table program seeds its (40,33) and (52,45) runs with
    (`current.counts[i] = ...` and the commented `current2.counts[i] = ...`)."""
    p = os.path.join(DD, "legacy", "hamming_count_undetectable_errors.cpp")
    lines = open(p).read().split("\n")
    res = []
    for var, k in (("current", 33), ("current2", 45)):
        got, first, last = {}, None, None
        for no, ln in enumerate(lines, 1):
            m = re.match(r"\s*(?://\s*)?" + var + r"\.counts\[(\d+)\] = (\d+)(?:ull)?;", ln)
            if m:
                got[int(m.group(1))] = int(m.group(2))
                first = first or no
                last = no
        counts = [got[i] for i in range(max(got) + 1)]
        res.append(dict(src=f"{os.path.relpath(p, REF)}:{first}-{last}", k=k, hamming_bits=7, counts=counts))
    return res


def recorded_csv_text():
    """Row lines and summary lines of two recorded result files whose every
    column is deterministic (counts, probabilities, binomial bases, and for the
    grid run the error columns), for line-by-line comparison of the CSV
    layout; the runtime lines are left out."""
    files = {
        "exhaustive_A61_n16": "results/runtimes/4/ancoding_gpu_A61_n16_h6.csv",
        "grid1d_m1001_A61_n16": "results/converge/geom/ancoding_grid1d_gpu_m1001_A61_n16_h6.csv",
    }
    # the reference CPU path's own files (run_ancoding_cpu: ancoding_cpu prefix, one timing row)
    for n in (8, 16, 24):
        files[f"cpu_A61_n{n}"] = f"results/runtimes/4/ancoding_cpu_A61_n{n}_h6.csv"
    for A in (1, 3, 17, 63, 99, 127):
        for n in (8, 12, 16):
            for f in glob.glob(os.path.join(DD, "results", "As", f"ancoding_gpu_A{A}_n{n}_h*.csv")):
                files[f"As_A{A}_n{n}"] = os.path.relpath(f, DD)
    out = {}
    for name, rel in files.items():
        lines = open(os.path.join(DD, rel)).read().split("\n")
        rows = []
        for ln in lines:
            if not ln.strip():
                break
            rows.append(ln)
        summary = [ln for ln in lines if ln.startswith('"Total"') or ln.startswith('"MaxAbsErr"')
                   or ln.startswith('"MaxRelErr"')]
        out[name] = dict(src="distance_distribution/" + rel, rows=rows, summary=summary)
    # k=40 grid: counts above 2^64 are printed with 12 digits; its error
    # columns hold uninitialised memory, so only the first four columns count
    rel = "results/runtimes/4/ancoding_grid1d_gpu_m101_A61_n40_h6.csv"
    lines = open(os.path.join(DD, rel)).read().split("\n")
    rows = []
    for ln in lines:
        if not ln.strip():
            break
        rows.append(",".join(ln.split(",")[:4]))
    out["grid1d_m101_A61_n40"] = dict(src="distance_distribution/" + rel, rows=rows, cols=4,
                                      summary=[ln for ln in lines if ln.startswith('"Total"')])
    return out


def main():
    sol = parse_solutions()
    runs, excluded = exhaustive_runs()
    json.dump(dict(source="distance_distribution/src/solutions.h", vectors={k: [str(x) for x in v] for k, v in sol.items()}),
              open(os.path.join(OUT, "solutions.json"), "w"), indent=0)
    json.dump(dict(runs=runs, excluded_truncated=excluded), open(os.path.join(OUT, "an_exhaustive_runs.json"), "w"))
    json.dump(dict(runs=grid_runs()), open(os.path.join(OUT, "an_grid1d_runs.json"), "w"))
    json.dump(dict(files=supera()), open(os.path.join(OUT, "supera_runs.json"), "w"))
    json.dump(dict(matrices=codeword_matrices()), open(os.path.join(OUT, "codeword_matrices.json"), "w"), indent=1)
    json.dump(recorded_csv_text(), open(os.path.join(OUT, "recorded_csv_text.json"), "w"), indent=1)
    json.dump(dict(tables=legacy_hamming_weights()), open(os.path.join(OUT, "legacy_hamming_weights.json"), "w"),
              indent=1)
    print(f"solutions: {len(sol)} vectors; exhaustive runs kept {len(runs)}, excluded {len(excluded)}")


if __name__ == "__main__":
    main()
