"""Summarise rocprofv3 CSV output of bench.py into profiles/ text.

    python -m openpfc_amd.profile_summary --trace DIR [--fetch DIR --write DIR]
        --cmd "<command that produced them>" --out profiles/NAME.md
        [--traffic-json profiles/NAME.json]

--trace   a `rocprofv3 --kernel-trace --stats -f csv` output directory
--fetch   a `rocprofv3 --pmc FETCH_SIZE -f csv` output directory
--write   a `rocprofv3 --pmc WRITE_SIZE -f csv` output directory

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same
counter reads for the plain device-to-device copy of `bench.py
--calibrate-copy`, which reads and writes 4 GiB per dispatch and runs last:
the calibration value is the median over the trailing __amd_rocclr_copyBuffer
dispatches with the grid size of the last one (other copies, such as the
autotuner's backup of the field, come earlier and differ in size).  The scaled bytes per dispatch of every kernel go to the
markdown table and, for the FP64 kernels, with --traffic-json to a JSON file
({name: [GB read, GB written]}) bench.py uses for roofline.traffic.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
from collections import defaultdict

COPY_BYTES = 4 << 30  # bench.py --calibrate-copy: 4 GiB read + 4 GiB written


def _csv(d: str, suffix: str):
    f = glob.glob(os.path.join(d, f"*{suffix}"))
    if not f:
        raise FileNotFoundError(f"no *{suffix} in {d}")
    with open(f[0]) as fh:
        return list(csv.DictReader(fh))


def short_name(name: str) -> str:
    """pfc_amd kernels -> 'k_tz<double, r16x16x4/E16, G8, split>' style names."""
    m = re.match(r"void pfc_amd::(k_\w+)<(.*)>\(", name)
    if not m:
        plain = name.replace("(anonymous namespace)::", "")
        return plain.split("(")[0][:80]
    args = m.group(2)
    t = "double" if args.startswith("double") else ("float" if args.startswith("float") else "")
    # one or two (paired) line plans: "N512/E8,r8x8x8[+N1024/E16,r16x16x4]"; a
    # plan with double butterflies (its last template argument true) is
    # "N1024/E8/xp,r8x8x4x4"
    plan_re = r"FftPlan<(\d+), (\d+), pfc_amd::RadixList<([\d, ]+)>(?:, (true|false))? ?>"
    plans = [f"N{n}/E{e}{'/xp' if xp == 'true' else ''},r" + r.replace(", ", "x")
             for n, e, r, xp in re.findall(plan_re, args)]
    rest = re.sub(r".*RadixList<[\d, ]+>(?:, (?:true|false))? ?>,?\s*", "", args)
    desc = []
    if plans:
        desc.append("+".join(plans))
    if rest:
        desc.append(rest.replace(" ", ""))
    return f"{m.group(1)}<{t}{',' if desc else ''}{','.join(desc)}>"


def kernel_stats(trace_dir: str):
    rows = _csv(trace_dir, "kernel_stats.csv")
    out = []
    for r in rows:
        out.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "avg_ms": float(r["AverageNs"]) / 1e6,
                    "min_ms": float(r["MinNs"]) / 1e6, "max_ms": float(r["MaxNs"]) / 1e6,
                    "pct": float(r["Percentage"])})
    return out


def counter_means(pmc_dir: str):
    rows = _csv(pmc_dir, "counter_collection.csv")
    per = defaultdict(list)
    copy = []
    for r in rows:
        v = float(r["Counter_Value"])
        if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer"):
            copy.append((int(r["Dispatch_Id"]), r["Grid_Size"], v))
        per[short_name(r["Kernel_Name"])].append(v)
    cal = None
    if copy:
        copy.sort()
        grid = copy[-1][1]
        tail = []
        for _, g, v in reversed(copy):
            if g != grid:
                break
            tail.append(v)
        tail.sort()
        cal = tail[len(tail) // 2]
    return {k: sum(v) / len(v) for k, v in per.items()}, cal


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cmd", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic-json")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--min-pct", type=float, default=0.5,
                    help="list kernels with at least this share of GPU time")
    a = ap.parse_args(argv)

    stats = kernel_stats(a.trace)
    lines = [f"# {a.title}", "", "Command (run from the repository root on one MI355X):", "",
             "```", a.cmd, "```", "",
             "| kernel | calls | avg ms | min ms | max ms | % time |",
             "|---|---:|---:|---:|---:|---:|"]
    shown = set()
    for s in sorted(stats, key=lambda s: -s["pct"]):
        if s["pct"] < a.min_pct:
            continue
        shown.add(s["name"])
        lines.append(f"| `{s['name']}` | {s['calls']} | {s['avg_ms']:.4f} | {s['min_ms']:.4f} "
                     f"| {s['max_ms']:.4f} | {s['pct']:.2f} |")
    traffic = {}
    if a.fetch and a.write:
        fm, fcopy = counter_means(a.fetch)
        wm, wcopy = counter_means(a.write)
        if not fcopy or not wcopy:
            raise RuntimeError("no copy-kernel dispatch in the PMC runs (use --calibrate-copy)")
        sf = COPY_BYTES / (fcopy * 1024.0)
        sw = COPY_BYTES / (wcopy * 1024.0)
        lines += ["", "## HBM traffic per dispatch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)", "",
                  f"Calibration: the 4 GiB device copy reads {fcopy:.0f} KiB (FETCH_SIZE) and "
                  f"writes {wcopy:.0f} KiB (WRITE_SIZE), so FETCH_SIZE is scaled by {sf:.4f} "
                  f"and WRITE_SIZE by {sw:.4f}.", "",
                  "| kernel | read GB | write GB | total GB |", "|---|---:|---:|---:|"]
        for k in sorted(set(fm) & set(wm)):
            if not k.startswith("k_"):
                continue
            rb, wb = fm[k] * 1024 * sf, wm[k] * 1024 * sw
            traffic[k] = {"read_bytes": rb, "write_bytes": wb, "traffic_bytes": rb + wb}
            if k in shown:
                lines.append(f"| `{k}` | {rb / 1e9:.3f} | {wb / 1e9:.3f} | "
                             f"{(rb + wb) / 1e9:.3f} |")
        if a.traffic_json:
            # FP64 step kernels only (the table bench.py reads), GB per dispatch
            table = {k: [round(v["read_bytes"] / 1e9, 4), round(v["write_bytes"] / 1e9, 4)]
                     for k, v in sorted(traffic.items()) if "<double" in k}
            with open(a.traffic_json, "w") as fh:
                json.dump({"command": a.cmd, "scale": {"fetch": sf, "write": sw},
                           "units": "GB read, GB written per dispatch",
                           "kernels": table}, fh, indent=0)
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
