"""Summaries of `rocprofv3 --kernel-trace --marker-trace --output-format csv` traces.

    python -m pbsam_amd.rocprof <run_kernel_trace.csv> [run_marker_api_trace.csv [bench.json]]

Prints a Markdown table of per-kernel dispatch statistics over the whole trace and, when
the marker trace holds bench.py's "timed_region" range, over the dispatches that ran
inside it; with bench.py's JSON line it also lists bench.py's own HIP-event averages of
the instrumented kernels beside rocprofv3's.
"""
import csv
import json
import re
import sys
from collections import defaultdict

INSTRUMENTED = {"xform": ("k_xform",), "p2l": ("k_p2l", "k_p2lgemv"), "solve": ("k_solve", "k_solve3"),
                "surface": ("k_surface",), "xgrad": ("k_xgrad",), "surfgrad": ("k_surfgrad", "k_surfgemv")}


def short(name):
    """'void pbsam::k_xform<128>(pbsam::Tabs, ...)' -> 'k_xform<128>'"""
    name = name.split("(")[0]
    name = re.sub(r"^void ", "", name)
    return name.replace("pbsam::", "")


def load_kernels(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((short(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    return rows


def find_range(path, label):
    with open(path) as f:
        for r in csv.DictReader(f):
            if any(label in str(v) for v in r.values()):
                return int(r["Start_Timestamp"]), int(r["End_Timestamp"])
    return None


def table(rows, title):
    agg = defaultdict(list)
    for name, s, e in rows:
        agg[name].append(e - s)
    print(title)
    print()
    print("| kernel | calls | total ms | avg us | min us | max us |")
    print("|---|---|---|---|---|---|")
    for name, d in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print("| %s | %d | %.3f | %.1f | %.1f | %.1f |" % (name, len(d), sum(d) / 1e6, sum(d) / len(d) / 1e3,
                                                          min(d) / 1e3, max(d) / 1e3))
    print()
    return agg


def main(argv):
    rows = load_kernels(argv[1])
    table(rows, "All dispatches of the run")
    if len(argv) < 3:
        return
    rng = find_range(argv[2], "timed_region")
    if rng is None:
        print("no timed_region range in the marker trace")
        return
    inside = [r for r in rows if r[1] >= rng[0] and r[2] <= rng[1]]
    agg = table(inside, "Dispatches inside the timed region (%.3f ms of wall time)" % ((rng[1] - rng[0]) / 1e6))
    if len(argv) < 4:
        return
    with open(argv[3]) as f:
        line = [json.loads(x) for x in f if x.startswith("{")][-1]
    print("Instrumented kernels: rocprofv3 against bench.py's HIP events (%d steps)" % line["steps"])
    print()
    print("| kernel | rocprofv3 dispatches | rocprofv3 avg us | bench.py launches | bench.py avg us |")
    print("|---|---|---|---|---|")
    for key, k in line["kernels"].items():
        if not k["launches"]:
            continue
        d = [x for name, v in agg.items() if name.split("<")[0] in INSTRUMENTED[key] for x in v]
        ev = k["ms_per_step"] * line["steps"] / k["launches"] * 1e3
        print("| %s | %d | %.1f | %d | %.1f |" % ("/".join(INSTRUMENTED[key]), len(d), sum(d) / max(len(d), 1) / 1e3,
                                                 k["launches"], ev))


if __name__ == "__main__":
    main(sys.argv)
