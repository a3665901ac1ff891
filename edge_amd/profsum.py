"""Summaries of rocprofv3 CSV output for the kernels of this library.

    python -m edge_amd.profsum stats <run_kernel_stats.csv> [--top K]
    python -m edge_amd.profsum counters <run_counter_collection.csv> [--match REGEX]

`stats` prints the per-kernel duration table of a `--kernel-trace --stats`
run.  `counters` prints, per kernel name, the number of dispatches, the mean
dispatch duration (from the counter run's own timestamps) and the mean of
every counter per dispatch, and derived figures where their inputs are
present:
  mfma_busy   = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x 4 SIMDs x CUs)
  mfma_tflops = (SQ_INSTS_VALU_MFMA_MOPS_F32 + _F64) x 512 / duration
  (MOPS count matrix operations in units of 512 FLOPs)
  gui_ghz     = GRBM_GUI_ACTIVE / duration (the counter's clock as seen per
                dispatch; a multiple of the engine clock means the counter is
                summed over dies)
"""
import argparse
import collections
import csv
import re
import sys

N_CU = 256


def short_name(name):
    """The kernel's template name without namespaces, return type and
    arguments: `void edge::(anonymous namespace)::k_local<5, true>(...)` ->
    `k_local<5, true>`."""
    n = name[5:] if name.startswith("void ") else name
    n = n.replace("(anonymous namespace)::", "").replace("edge::", "")
    depth = 0
    for i, ch in enumerate(n):  # cut at the argument list: the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def read_stats(path):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]), "avg_ms": float(r["AverageNs"]) * 1e-6,
                         "min_ms": float(r["MinNs"]) * 1e-6, "max_ms": float(r["MaxNs"]) * 1e-6,
                         "pct": float(r["Percentage"])})
    rows.sort(key=lambda r: -r["pct"])
    return rows


def read_counters(path, match=None):
    """{kernel: {"dispatches": n, "ms": mean duration, counter: mean value}}"""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(dict)
    pat = re.compile(match) if match else None
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = short_name(r["Kernel_Name"])
            if pat and not pat.search(name):
                continue
            did = int(r["Dispatch_Id"])
            per[(name, did)][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[name][did] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
    out = {}
    for name, ds in disp.items():
        keys = sorted({k for (n, d) in per if n == name for k in per[(n, d)]})
        agg = {"dispatches": len(ds), "ms": sum(ds.values()) / len(ds)}
        for k in keys:
            agg[k] = sum(per[(name, d)].get(k, 0.0) for d in ds) / len(ds)
        gui = agg.get("GRBM_GUI_ACTIVE")
        if gui and "SQ_VALU_MFMA_BUSY_CYCLES" in agg:
            agg["mfma_busy"] = agg["SQ_VALU_MFMA_BUSY_CYCLES"] / (gui * 4 * N_CU)
        mops = agg.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0) + agg.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0)
        if mops and agg["ms"] > 0:
            agg["mfma_tflops"] = mops * 512 / (agg["ms"] * 1e-3) / 1e12
        if gui and agg["ms"] > 0:
            agg["gui_ghz"] = gui / (agg["ms"] * 1e-3) / 1e9
        out[name] = agg
    return out


def _fmt(v):
    if isinstance(v, int):
        return f"{v:,}"
    if abs(v) >= 1e4:
        return f"{v:,.0f}"
    return f"{v:.4g}"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=12)
    b = sub.add_parser("counters")
    b.add_argument("csv")
    b.add_argument("--match", default=None)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print("| Kernel | Calls | Average ms | Min ms | Max ms | % of GPU time |")
        print("|---|---|---|---|---|---|")
        for r in read_stats(args.csv)[:args.top]:
            print(f"| `{r['name']}` | {r['calls']} | {r['avg_ms']:.3f} | {r['min_ms']:.3f} | {r['max_ms']:.3f} | "
                  f"{r['pct']:.1f} |")
        return 0
    res = read_counters(args.csv, args.match)
    cols = sorted({k for v in res.values() for k in v} - {"dispatches", "ms"})
    print("| Kernel | dispatches | ms | " + " | ".join(cols) + " |")
    print("|---|---|---|" + "---|" * len(cols))
    for name, v in sorted(res.items(), key=lambda kv: -kv[1]["ms"]):
        print(f"| `{name}` | {v['dispatches']} | {v['ms']:.3f} | " +
              " | ".join(_fmt(v[c]) if c in v else "" for c in cols) + " |")
    return 0


if __name__ == "__main__":
    sys.exit(main())
