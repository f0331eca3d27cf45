"""Summaries of rocprofv3 CSV output for the profiles kept with the benchmark.

    python -m bbwadg_amd.profsum stats <run_kernel_stats.csv>
    python -m bbwadg_amd.profsum trace <run_kernel_trace.csv> [kernel substring ...]
    python -m bbwadg_amd.profsum pmc <run_counter_collection.csv> <kernel substring>

``stats`` prints a markdown table of the per-kernel launch count and durations; ``trace``
prints the same per (kernel, grid size), so launches over meshes of different sizes are kept
apart; ``pmc`` prints the mean per dispatch of
every counter collected for kernels whose name contains the substring (FETCH_SIZE and
WRITE_SIZE are in KiB).
"""
import csv
import sys
from collections import defaultdict


def stats(path):
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"].split("(")[0]
        out.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.1f} |")
    return "\n".join(out)


def trace(path, kernels=()):
    rows = list(csv.DictReader(open(path)))
    groups = defaultdict(list)
    for r in rows:
        name = r["Kernel_Name"].split("(")[0]
        if kernels and not any(k in name for k in kernels):
            continue
        key = (name, int(r["Grid_Size_X"]) // int(r["Workgroup_Size_X"]))
        groups[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = ["| kernel | workgroups | calls | avg us | min us | max us |", "|---|---|---|---|---|---|"]
    for key in sorted(groups, key=lambda k: -sum(groups[k])):
        d = groups[key]
        out.append(f"| `{key[0]}` | {key[1]} | {len(d)} | {sum(d) / len(d):.1f} | {min(d):.1f} | "
                   f"{max(d):.1f} |")
    return "\n".join(out)


def pmc(path, kernel):
    rows = list(csv.DictReader(open(path)))
    agg = defaultdict(list)
    for r in rows:
        if kernel in r["Kernel_Name"]:
            agg[r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}, {k: len(v) for k, v in agg.items()}


def main():
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2]))
    elif sys.argv[1] == "trace":
        print(trace(sys.argv[2], sys.argv[3:]))
    else:
        means, counts = pmc(sys.argv[2], sys.argv[3])
        for k in sorted(means):
            print(f"{k}: mean {means[k]:.6g} over {counts[k]} dispatches")


if __name__ == "__main__":
    main()
