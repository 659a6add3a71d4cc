"""Summaries of rocprofv3 CSV output for the J/K kernels.

    python -m patmol_amd.profiling stats DIR [--title T] [--command C]
        per-kernel duration table from DIR/*_kernel_stats.csv and the
        register/scratch columns of DIR/*_kernel_trace.csv (markdown)
    python -m patmol_amd.profiling traffic FETCH_DIR WRITE_DIR CFETCH_DIR CWRITE_DIR
                                   --copy-bytes B --json OUT.json [--command C]
        per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of every kernel, scaled by
        the factor that makes a plain device copy of B bytes read B and write
        B bytes (markdown table on stdout, per-kernel bytes in OUT.json)

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
    args = ap.parse_args(argv)
    stats(args) if args.cmd == "stats" else traffic(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
