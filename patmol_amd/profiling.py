"""Summaries of rocprofv3 CSV output for the J/K kernels.

    python -m patmol_amd.profiling stats DIR [--title T] [--command C]
        per-kernel duration table from DIR/*_kernel_stats.csv and the
        register/scratch columns of DIR/*_kernel_trace.csv (markdown)
    python -m patmol_amd.profiling traffic FETCH_DIR WRITE_DIR CFETCH_DIR CWRITE_DIR
                                   --copy-bytes B --json OUT.json [--command C]
        per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of every kernel, scaled by
        the factor that makes a plain device copy of B bytes read B and write
        B bytes (markdown table on stdout, per-kernel bytes in OUT.json)
    python -m patmol_amd.profiling counters DIR [DIR ...] --json OUT.json [--command C]
        per-dispatch means of every counter of --pmc runs (one or more
        directories, e.g. one per pass), per kernel (markdown, OUT.json)
    python -m patmol_amd.profiling valu --fp64 DIR [DIR ...] --fp32 DIR [DIR ...]
                                   --bench BENCH.json [--command C]
        per J/K class: VALU instructions and executed FP64 / FP32 FLOPs of
        the FP64 and the FP32 kernel (SQ_INSTS_VALU, SQ_INSTS_VALU_FLOPS_*,
        64 lanes per wave instruction), their durations in the counter runs,
        executed TFLOP/s, and the algorithmic FLOPs of the class from a
        bench.py JSON line (headline per_class_fp64)

Only reads the CSV files rocprofv3 writes with --output-format csv.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys


def _one(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, f"*{suffix}")))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits[0]


def _short(name, n=72):
    return name if len(name) <= n else name[: n - 3] + "..."


def stats(args):
    rows = list(csv.DictReader(open(_one(args.dir, "kernel_stats.csv"))))
    meta = {}
    for r in csv.DictReader(open(_one(args.dir, "kernel_trace.csv"))):
        meta.setdefault(r["Kernel_Name"], (r["VGPR_Count"], r["Accum_VGPR_Count"], r["Scratch_Size"],
                                           int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))))
    out = []
    if args.title:
        out.append(f"# {args.title}\n")
    if args.command:
        out.append(f"Command: `{args.command}`\n")
    out.append("| kernel | calls | total ms | avg us | % | vgpr | agpr | scratch B/lane | workgroups |")
    out.append("|---|---|---|---|---|---|---|---|---|")
    for r in rows:
        v = meta.get(r["Name"], ("?", "?", "?", "?"))
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} | {v[0]} | {v[1]} | {v[2]} | {v[3]} |")
    print("\n".join(out))


def _per_dispatch(d):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(_one(d, "counter_collection.csv"))):
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def traffic(args):
    cf, cw = _per_dispatch(args.cfetch), _per_dispatch(args.cwrite)
    kib = args.copy_bytes / 1024.0

    def copy_reading(agg):
        vals = [v for k, lst in agg.items() if "copyBuffer" in k or "copy" in k.lower() for v in lst
                if v > 0.25 * kib]
        if not vals:
            raise SystemExit("no copy dispatch found in the calibration run")
        return sum(vals) / len(vals)

    f_fac = kib / copy_reading(cf)
    w_fac = kib / copy_reading(cw)
    fe, wr = _per_dispatch(args.fetch), _per_dispatch(args.write)
    res = {"copy_bytes": args.copy_bytes, "fetch_factor": f_fac, "write_factor": w_fac, "kernels": {}}
    out = []
    if args.command:
        out.append(f"Command: `{args.command}`\n")
    out.append(f"Copy calibration: a device copy of {args.copy_bytes} B reads FETCH_SIZE x {f_fac:.3f} "
               f"and WRITE_SIZE x {w_fac:.3f} of the true bytes; the factors are applied below.\n")
    out.append("| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | total MB/dispatch |")
    out.append("|---|---|---|---|---|")
    names = sorted(set(fe) | set(wr), key=lambda k: -(sum(fe.get(k, [0])) / max(1, len(fe.get(k, [1])))
                                                      + sum(wr.get(k, [0])) / max(1, len(wr.get(k, [1])))))
    for k in names:
        f = sum(fe.get(k, [0.0])) / max(1, len(fe.get(k, [1]))) * 1024.0 * f_fac
        w = sum(wr.get(k, [0.0])) / max(1, len(wr.get(k, [1]))) * 1024.0 * w_fac
        res["kernels"][k] = {"fetch_bytes": f, "write_bytes": w, "bytes": f + w}
        out.append(f"| `{_short(k)}` | {len(fe.get(k, []))} | {f / 1e6:.1f} | {w / 1e6:.1f} | {(f + w) / 1e6:.1f} |")
    print("\n".join(out))
    with open(args.json, "w") as fh:
        json.dump(res, fh, indent=1)


def counters(args):
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    names = []
    for d in args.dirs:
        for r in csv.DictReader(open(_one(d, "counter_collection.csv"))):
            k, c = r["Kernel_Name"], r["Counter_Name"]
            if c not in names:
                names.append(c)
            vals[k][c] += float(r["Counter_Value"])
            disp[k][c].add(r.get("Dispatch_Id", r.get("Correlation_Id", "")))
    res = {}
    out = []
    if args.command:
        out.append(f"Command: `{args.command}`\n")
    out.append("| kernel | " + " | ".join(names) + " |")
    out.append("|---|" + "---|" * len(names))
    order = sorted(vals, key=lambda k: -max(vals[k].values()))
    for k in order:
        row = {c: vals[k][c] / max(1, len(disp[k][c])) for c in names if c in vals[k]}
        res[k] = row
        out.append(f"| `{_short(k)}` | " + " | ".join(f"{row[c]:.4g}" if c in row else "-" for c in names) + " |")
    print("\n".join(out))
    with open(args.json, "w") as fh:
        json.dump(res, fh, indent=1)


_L = "spd"


def _class_of(kernel):
    """'(ab|cd)' and precision of a J/K kernel symbol, or None."""
    import re

    m = re.search(r"jk_kernel3?<(\d), (\d), (\d), (\d), (double|float)>", kernel)
    if not m:
        return None
    la, lb, lc, ld = (int(m.group(i)) for i in range(1, 5))
    return f"({_L[la]}{_L[lb]}|{_L[lc]}{_L[ld]})", m.group(5)


def _collect(dirs):
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(lambda: collections.defaultdict(set))
    dur = collections.defaultdict(list)
    for d in dirs:
        for r in csv.DictReader(open(_one(d, "counter_collection.csv"))):
            vals[r["Kernel_Name"]][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[r["Kernel_Name"]][r["Counter_Name"]].add(r.get("Dispatch_Id", ""))
        for r in csv.DictReader(open(_one(d, "kernel_trace.csv"))):
            dur[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    out = {}
    for k in vals:
        row = {c: vals[k][c] / max(1, len(disp[k][c])) for c in vals[k]}
        row["seconds"] = sum(dur[k]) / max(1, len(dur[k]))
        out[k] = row
    return out


def valu(args):
    f64, f32 = _collect(args.fp64), _collect(args.fp32)
    bench = json.load(open(args.bench))
    per = bench["headline"]["per_class_fp64"]
    rows = {}
    for data, prec in ((f64, "double"), (f32, "float")):
        for k, row in data.items():
            c = _class_of(k)
            if c and c[1] == prec:
                rows.setdefault(c[0], {})[prec] = row
    out = []
    if args.command:
        out.append(f"Command: `{args.command}`\n")
    out.append("Executed FLOPs = 64 x SQ_INSTS_VALU_FLOPS_FP64 (FP32) (2 per FMA, 1 per add / multiply / "
               "transcendental, per wave instruction); durations from the counter runs' kernel trace; "
               "algorithmic FLOPs of the class from the bench line.\n")
    out.append("| class | alg. GFLOP | FP64 kernel: VALU instr (M) | exec. FP64 GFLOP | ms | exec. TFLOP/s | exec./alg. "
               "| FP32 kernel: VALU instr (M) | exec. FP32 GFLOP | exec. FP64 GFLOP | ms | FP64/FP32 time |")
    out.append("|---|---|---|---|---|---|---|---|---|---|---|---|")
    res = {}
    t64 = t32 = 0.0
    for cls in sorted(rows, key=lambda c: -rows[c].get("double", {}).get("seconds", 0)):
        d, f = rows[cls].get("double"), rows[cls].get("float")
        alg = per.get(cls, {}).get("tflops", 0) * per.get(cls, {}).get("ms", 0) * 1e9 / 1e9   # GFLOP
        line = f"| {cls} | {alg:.1f} |"
        r = {"alg_gflop": alg}
        if d:
            ex = 64 * d.get("SQ_INSTS_VALU_FLOPS_FP64", 0) / 1e9
            t64 += d["seconds"]
            r.update(fp64_valu=d.get("SQ_INSTS_VALU", 0), fp64_exec_gflop=ex, fp64_s=d["seconds"])
            line += (f" {d.get('SQ_INSTS_VALU', 0) / 1e6:.0f} | {ex:.1f} | {d['seconds'] * 1e3:.3f} | "
                     f"{ex / max(d['seconds'], 1e-12) / 1e3:.2f} | {ex / max(alg, 1e-12):.2f} |")
        else:
            line += " - | - | - | - | - |"
        if f:
            ex32 = 64 * f.get("SQ_INSTS_VALU_FLOPS_FP32", 0) / 1e9
            ex64 = 64 * f.get("SQ_INSTS_VALU_FLOPS_FP64", 0) / 1e9
            t32 += f["seconds"]
            r.update(fp32_valu=f.get("SQ_INSTS_VALU", 0), fp32_exec_gflop=ex32, fp32_s=f["seconds"])
            ratio = d["seconds"] / f["seconds"] if d and f["seconds"] > 0 else 0
            line += (f" {f.get('SQ_INSTS_VALU', 0) / 1e6:.0f} | {ex32:.1f} | {ex64:.1f} | {f['seconds'] * 1e3:.3f} |"
                     f" {ratio:.2f} |")
        else:
            line += " - | - | - | - | - |"
        res[cls] = r
        out.append(line)
    out.append(f"\nSum of the class kernels: FP64 {t64 * 1e3:.2f} ms, FP32 (where a FP32 kernel ran) "
               f"{t32 * 1e3:.2f} ms.")
    print("\n".join(out))
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(res, fh, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m patmol_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    for x in ("fetch", "write", "cfetch", "cwrite"):
        t.add_argument(x)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--json", required=True)
    t.add_argument("--command", default="")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--json", required=True)
    c.add_argument("--command", default="")
    v = sub.add_parser("valu")
    v.add_argument("--fp64", nargs="+", required=True)
    v.add_argument("--fp32", nargs="+", required=True)
    v.add_argument("--bench", required=True)
    v.add_argument("--json", default="")
    v.add_argument("--command", default="")
    args = ap.parse_args(argv)
    {"stats": stats, "traffic": traffic, "counters": counters, "valu": valu}[args.cmd](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
