"""Summaries of `rocprofv3 --kernel-trace --marker-trace --output-format csv` traces.

    python -m pbsam_amd.rocprof <run_kernel_trace.csv> [run_marker_api_trace.csv [bench.json]]

Prints a Markdown table of per-kernel dispatch statistics over the whole trace and, when
the marker trace holds bench.py's "timed_region" (and "batch_region", "mixed_region", "bd_region",
"large_region", "tight_region", "exact_region", "large_exact_region", "kernel_timing_region") range, over the dispatches that ran inside it; with bench.py's JSON line it also
lists bench.py's own HIP-event averages of the instrumented kernels (timed in the instrumented repeat of
the timed evaluations) beside rocprofv3's over the same dispatches.

    python -m pbsam_amd.rocprof pmc <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv> <copy_bytes>
        [--after KERNEL]

Per-kernel HBM traffic from `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` counter
collections (KiB per dispatch), each scaled by what the same counter reads for one
device-to-device copy of copy_bytes (bench.py --copy-calibration); with --after, only the
dispatches from the first dispatch of KERNEL on (k_rotcoeff: bench.py's evaluation, past
the molecule types' set-up).
"""
import csv
import json
import re
import sys
from collections import defaultdict

INSTRUMENTED = {"xform": ("k_xform", "k_xapply"), "p2l": ("k_p2l", "k_p2lgemv", "k_p2lgemm"), "solve": ("k_solve", "k_solve3"),
                "surface": ("k_surface",), "xgrad": ("k_xgrad",), "surfgrad": ("k_surfgemm", "k_surfgemm_k4")}


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
    for label, what in (("batch_region", "bench.py's batched leg"), ("mixed_region", "bench.py's mixed-order batch leg"),
                        ("bd_region", "bench.py's Brownian-dynamics leg"),
                        ("large_region", "bench.py's large-system force evaluations"),
                        ("tight_region", "bench.py's tight-tolerance mutual solve"),
                        ("exact_region", "bench.py's exact evaluations (GMRES solve + adjoint)"),
                        ("large_exact_region", "bench.py's large-system exact evaluation (GMRES solve + "
                                               "adjoint of all 48 coordinates)")):
        brng = find_range(argv[2], label)
        if brng is not None:
            inb = [r for r in rows if r[1] >= brng[0] and r[2] <= brng[1]]
            table(inb, "Dispatches inside %s (%.3f ms of wall time)" % (what, (brng[1] - brng[0]) / 1e6))
    rng = find_range(argv[2], "timed_region")
    if rng is None:
        print("no timed_region range in the marker trace")
        return
    inside = [r for r in rows if r[1] >= rng[0] and r[2] <= rng[1]]
    agg = table(inside, "Dispatches inside the timed region (%.3f ms of wall time)" % ((rng[1] - rng[0]) / 1e6))
    # bench.py's HIP events time a repeat of the timed evaluations ("kernel_timing_region")
    erng = find_range(argv[2], "kernel_timing_region")
    if erng is not None:
        ins = [r for r in rows if r[1] >= erng[0] and r[2] <= erng[1]]
        agg = table(ins, "Dispatches inside the instrumented repeat (%.3f ms of wall time)" %
                    ((erng[1] - erng[0]) / 1e6))
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


def load_counters(path, counter, after=None):
    """{kernel: [value per dispatch]} of one counter (KiB) from a counter-collection CSV;
    with `after`, only dispatches from the first dispatch of that kernel on"""
    per = defaultdict(float)
    names = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            d = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
            per[d] += float(r["Counter_Value"])
            names[d] = short(r["Kernel_Name"])
    first = min((d for d, n in names.items() if n == after), default=None) if after else None
    out = defaultdict(list)
    for d, v in per.items():
        if first is None or d >= first:
            out[names[d]].append(v)
    return out


def pmc(argv):
    after = None
    if "--after" in argv:
        i = argv.index("--after")
        after = argv[i + 1]
        argv = argv[:i] + argv[i + 2:]
    fetch, write = load_counters(argv[0], "FETCH_SIZE", after), load_counters(argv[1], "WRITE_SIZE", after)
    cf, cw = load_counters(argv[2], "FETCH_SIZE"), load_counters(argv[3], "WRITE_SIZE")
    nbytes = float(argv[4])
    copy = [k for k in cf if "copy" in k.lower()]
    if not copy:
        print("no copy kernel in the calibration run")
        return
    kf, kw = max(cf[copy[0]]), max(cw[copy[0]])
    sf, sw = nbytes / (kf * 1024.0), nbytes / (kw * 1024.0)
    print("Calibration: %s of %d bytes reads FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB; scale %.4f (fetch), %.4f "
          "(write)" % (copy[0], nbytes, kf, kw, sf, sw))
    print()
    print("| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | HBM bytes/dispatch (scaled) |")
    print("|---|---|---|---|---|")
    rows = []
    for k in fetch:
        f_ = sum(fetch[k]) / len(fetch[k])
        w_ = sum(write.get(k, [0.0])) / max(len(write.get(k, [])), 1)
        rows.append((f_ * 1024 * sf + w_ * 1024 * sw, k, len(fetch[k]), f_, w_))
    for b, k, n, f_, w_ in sorted(rows, reverse=True):
        print("| %s | %d | %.1f | %.1f | %.0f |" % (k, n, f_, w_, b))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "pmc":
        pmc(sys.argv[2:])
    else:
        main(sys.argv)
