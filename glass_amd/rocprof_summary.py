"""Per-kernel summaries of rocprofv3 CSV output (the tables kept with the
measurements).

    python -m glass_amd.rocprof_summary counters run_counter_collection.csv
    python -m glass_amd.rocprof_summary stats run_kernel_stats.csv
    python -m glass_amd.rocprof_summary trace run_kernel_trace.csv

counters: per kernel (template arguments kept, grid size in blocks shown so
batched launches stand apart from per-call ones) the number of dispatches,
the mean dispatch duration and the mean of every counter per dispatch
(summed over the counter's instances).  stats: the kernel-trace summary
(calls, mean and total duration) in a short table.  trace: the same per
kernel and grid size from the per-dispatch kernel trace.
"""
import collections
import csv
import re
import sys


def _short(name):
    """Kernel name without return type, namespaces and argument list,
    template arguments kept: 'eval_kernel<512,256,false,1>'."""
    n = re.sub(r"^void\s+", "", name.replace("(anonymous namespace)::", ""))
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            n = n[:i]
            break
    n = re.sub(r"^(glass::)?(wdm::)?", "", n)
    return n.replace(" ", "")[:80]


def counters(path):
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    dur = collections.defaultdict(float)
    names = set()
    for r in csv.DictReader(open(path)):
        wg = max(int(r.get("Workgroup_Size", 1) or 1), 1)
        key = (_short(r["Kernel_Name"]), int(r["Grid_Size"]) // wg)
        agg[key][r["Counter_Name"]] += float(r["Counter_Value"])
        names.add(r["Counter_Name"])
        if r["Dispatch_Id"] not in disp[key]:
            disp[key].add(r["Dispatch_Id"])
            dur[key] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    cols = sorted(names)
    print("kernel | blocks | dispatches | mean_ms | " + " | ".join(cols))
    for key in sorted(agg, key=lambda k: -dur[k]):
        n = len(disp[key])
        vals = " | ".join(f"{agg[key][c] / n:.4e}" for c in cols)
        print(f"{key[0]} | {key[1]} | {n} | {dur[key] / n / 1e6:.4f} | {vals}")


def _blocks(r):
    """Workgroups of a dispatch from a kernel-trace or counter row."""
    if "Grid_Size" in r:
        g, w = int(r["Grid_Size"]), int(r.get("Workgroup_Size", 1) or 1)
    else:
        g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        w = int(r["Workgroup_Size_X"]) * int(r["Workgroup_Size_Y"]) * int(r["Workgroup_Size_Z"])
    return g // max(w, 1)


def trace(path):
    """Per (kernel, grid size in blocks) dispatch count and mean duration
    from run_kernel_trace.csv, so launches of one kernel at different sizes
    stand apart."""
    n = collections.Counter()
    dur = collections.defaultdict(float)
    for r in csv.DictReader(open(path)):
        key = (_short(r["Kernel_Name"]), _blocks(r))
        n[key] += 1
        dur[key] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    print("kernel | blocks | dispatches | mean_us | total_ms")
    for key in sorted(n, key=lambda k: -dur[k]):
        print(f"{key[0]} | {key[1]} | {n[key]} | {dur[key] / n[key] / 1e3:.2f} | {dur[key] / 1e6:.3f}")


def stats(path):
    print("kernel | calls | mean_us | total_ms | percent")
    for r in csv.DictReader(open(path)):
        print(f"{_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f}")


if __name__ == "__main__":
    {"counters": counters, "stats": stats, "trace": trace}[sys.argv[1]](sys.argv[2])
