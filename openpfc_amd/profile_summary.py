"""Summarise rocprofv3 CSV output of bench.py into profiles/ text.

    python -m openpfc_amd.profile_summary --trace DIR [--fetch DIR --write DIR]
        --cmd "<command that produced them>" --out profiles/NAME.md
        [--traffic-json profiles/NAME.json]

--trace   a `rocprofv3 --kernel-trace --stats -f csv` output directory
--fetch   a `rocprofv3 --pmc FETCH_SIZE -f csv` output directory
--write   a `rocprofv3 --pmc WRITE_SIZE -f csv` output directory
--gap-window MS [--window-steps K]   the idle share of the kernel activity
          after the last idle gap longer than MS in the trace (the graph steps
          of `bench.py --ensemble --ens-trace-steps K`)

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


def busy_window(trace_dir: str, gap_ms: float):
    """Kernel activity after the last idle gap longer than gap_ms in a
    `rocprofv3 --kernel-trace` run: the window's span, the time covered by
    at least one kernel (the union of the dispatch intervals; kernels on
    parallel branches overlap) and the launches in it."""
    rows = _csv(trace_dir, "kernel_trace.csv")
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows)
    start = 0
    reach = iv[0][1]
    for i in range(1, len(iv)):
        if iv[i][0] - reach > gap_ms * 1e6:
            start = i
        reach = max(reach, iv[i][1])
    win = iv[start:]
    covered, cur_s, cur_e = 0, win[0][0], win[0][1]
    for b, e in win[1:]:
        if b > cur_e:
            covered += cur_e - cur_s
            cur_s, cur_e = b, e
        else:
            cur_e = max(cur_e, e)
    covered += cur_e - cur_s
    span = max(e for _, e in win) - win[0][0]
    return {"kernels": len(win), "span_ms": span / 1e6, "busy_ms": covered / 1e6,
            "idle_frac": 1.0 - covered / span}


def kernel_vgprs(lib: str):
    """VGPRs per kernel (short names) from the gfx950 code objects of a
    library built here: the objects are unbundled with llvm-objdump
    --offloading into a scratch directory and their notes read with
    llvm-readelf."""
    import shutil
    import subprocess
    import tempfile
    llvm = "/opt/rocm/llvm/bin"
    out = {}
    with tempfile.TemporaryDirectory() as d:
        copy = os.path.join(d, "lib.so")
        shutil.copy(lib, copy)
        subprocess.run([f"{llvm}/llvm-objdump", "--offloading", copy], check=True,
                       capture_output=True)
        for obj in glob.glob(os.path.join(d, "lib.so.*gfx*")):
            notes = subprocess.run([f"{llvm}/llvm-readelf", "--notes", obj], check=True,
                                   capture_output=True, text=True).stdout
            for blk in notes.split(".agpr_count:")[1:]:
                nm = re.search(r"\.name:\s+(\S+)", blk)
                vg = re.search(r"\.vgpr_count:\s+(\d+)", blk)
                if nm and vg:
                    out[nm.group(1)] = int(vg.group(1))
    # short names of the demangled symbols (one c++filt run for all)
    names = sorted(out)
    dem = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True,
                         text=True, check=True).stdout.splitlines()
    short = {}
    for n, d in zip(names, dem):
        d = d.strip()
        short[short_name(d if d.startswith("void ") else "void " + d)] = out[n]
    return short


def sq_table(dir_a: str, dir_b: str, lib: str | None):
    """Per-kernel SQ counter table from two `rocprofv3 --pmc` runs of the same
    command (their counters joined per kernel; means over dispatches)."""
    vg = kernel_vgprs(lib) if lib else {}
    per = defaultdict(lambda: defaultdict(list))
    meta = {}
    for d in (dir_a, dir_b):
        for r in _csv(d, "counter_collection.csv"):
            k = short_name(r["Kernel_Name"])
            if not k.startswith("k_t"):
                continue
            per[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            mangled = re.sub(r"\(.*", "", r["Kernel_Name"])
            meta[k] = (int(r["Workgroup_Size"]), int(r["LDS_Block_Size"]), mangled)
    rows = []
    for k, c in per.items():
        x = {n: sum(v) / len(v) for n, v in c.items()}
        need = ("SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
                "SQ_ACTIVE_INST_LDS")
        if any(n not in x for n in need):
            continue
        wg, lds, _ = meta[k]
        rows.append((k, wg, lds, vg.get(k), x))
    return rows


def waves_per_simd(vgpr, lds, wg):
    """Waves a SIMD holds for this kernel: 512 VGPRs per lane in granules of
    8, 160 KiB of LDS per CU, at most 8 waves, whole blocks."""
    if not vgpr:
        return None
    by_vgpr = min(8, 512 // (8 * ((vgpr + 7) // 8)))
    wpb = max(1, wg // 64)
    blocks = (4 * by_vgpr) // wpb
    if lds:
        blocks = min(blocks, (160 * 1024) // lds)
    return blocks * wpb / 4


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cmd", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic-json")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--min-pct", type=float, default=0.5,
                    help="list kernels with at least this share of GPU time")
    ap.add_argument("--gap-window", type=float, metavar="MS",
                    help="also report the idle share of the kernel activity after the last "
                         "idle gap longer than MS (bench.py --ensemble --ens-trace-steps)")
    ap.add_argument("--window-steps", type=int, default=0,
                    help="steps in that window (for the per-step figures)")
    ap.add_argument("--sq", nargs=2, metavar="DIR",
                    help="two rocprofv3 --pmc output directories with the SQ counters "
                         "(SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY "
                         "SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_ACTIVE_INST_VMEM; "
                         "SQ_INSTS_VALU SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS "
                         "SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INST_LEVEL_VMEM SQ_LEVEL_WAVES)")
    ap.add_argument("--lib", help="the built library, for the kernels' VGPR counts")
    a = ap.parse_args(argv)
    if a.sq:
        return sq_main(a)

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
    if a.gap_window:
        w = busy_window(a.trace, a.gap_window)
        lines += ["", "## Kernel activity of the traced steps", "",
                  f"After the last idle gap longer than {a.gap_window:g} ms: {w['kernels']} "
                  f"kernel launches over {w['span_ms']:.3f} ms, of which {w['busy_ms']:.3f} ms "
                  f"have at least one kernel running; idle share {100 * w['idle_frac']:.2f} %."]
        if a.window_steps:
            lines.append(f"Per step ({a.window_steps} steps): {w['kernels'] / a.window_steps:.0f} "
                         f"launches, {w['span_ms'] / a.window_steps:.3f} ms span, "
                         f"{(w['span_ms'] - w['busy_ms']) / a.window_steps:.4f} ms idle.")
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


def sq_main(a):
    rows = sq_table(a.sq[0], a.sq[1], a.lib)
    lines = [f"# {a.title}", "", "Commands (run from the repository root on one MI355X):", "",
             "```", a.cmd, "```", "",
             "Per kernel, means over its dispatches.  Percentages are of SQ_WAVE_CYCLES "
             "(wave-cycles) for SQ_ACTIVE_INST_VALU, SQ_ACTIVE_INST_LDS, SQ_WAIT_INST_ANY, "
             "SQ_WAIT_ANY and SQ_WAIT_INST_LDS.  conflict/LDS: SQ_LDS_BANK_CONFLICT over "
             "SQ_ACTIVE_INST_LDS (cycles lost to bank conflicts per LDS-active cycle).  "
             "VGPRs from the built code object; waves/SIMD from VGPRs, LDS and the block "
             "size (512 VGPRs per lane in granules of 8, 160 KiB LDS per CU).", "",
             "| kernel | block | LDS B | VGPRs | waves/SIMD | VALU/wave | LDS inst/wave | "
             "VALU % | LDS % | WAIT_INST_ANY % | WAIT_ANY % | WAIT_INST_LDS % | conflict/LDS |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for k, wg, lds, vgpr, x in sorted(rows, key=lambda r: r[0]):
        wc = x["SQ_WAVE_CYCLES"]
        w = x["SQ_WAVES"]
        occ = waves_per_simd(vgpr, lds, wg)

        def pct(n):
            return f"{100 * x[n] / wc:.1f}" if n in x and wc else "-"
        lines.append(
            f"| `{k}` | {wg} | {lds} | {vgpr if vgpr else '-'} | "
            f"{occ if occ is not None else '-'} | "
            f"{x.get('SQ_INSTS_VALU', 0) / w:.0f} | {x['SQ_INSTS_LDS'] / w:.0f} | "
            f"{pct('SQ_ACTIVE_INST_VALU')} | {pct('SQ_ACTIVE_INST_LDS')} | "
            f"{pct('SQ_WAIT_INST_ANY')} | {pct('SQ_WAIT_ANY')} | {pct('SQ_WAIT_INST_LDS')} | "
            f"{x['SQ_LDS_BANK_CONFLICT'] / max(1.0, x['SQ_ACTIVE_INST_LDS']):.2f} |")
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
