"""Per-class launch-shape autotuner for the J/K kernels (jk_kernel3).

Every class (la lb | lc ld) of the LDS-staged J/K kernel compiles in one of
several launch shapes (the shape code of jk3_code in csrc/eri_kernel3.h:
Rys digestion order, density-read fence, waves per SIMD, bra-component
split over the waves, two workgroups per item, Rys coefficients in LDS or
global memory).  Which shape is fastest depends on the class's register
footprint, LDS use and item count, separately for the FP64 and the FP32
instantiation.  The tool:

  build    compiles one library per variant (hipcc cross-compiles, no GPU
           needed): variant v gives every class its v-th candidate shape,
           so a handful of builds covers every class's candidates;
  measure  (GPU) times every class in every variant on a workload, classes
           serialised on one stream, all items FP64 and then all items FP32,
           and checks each variant's G against variant 0's;
  select   picks per class and precision the fastest candidate, marks the
           classes whose FP32 kernel is faster than their FP64 kernel (only
           those run items in FP32 in mixed precision), and writes
           csrc/jk3_tuned.inc; rebuild the library afterwards.

The candidates of a class are its current shape and the shapes one change
away from it (one coordinate-descent step per build/measure/select pass).

Usage:
  python -m patmol_amd.autotune build [--out DIR] [--jobs N]
  python -m patmol_amd.autotune measure [--dir DIR] [--system water16] [--out FILE]
  python -m patmol_amd.autotune select [--timings FILE] [--table PATH]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
TABLE = os.path.join(CSRC, "jk3_tuned.inc")

# the classes with a jk_kernel3 instantiation (the (dd|dd) class runs the
# ket-pair-per-lane kernel, whose shape is fixed; it has an FP32 policy only)
CLASSES = ["0000", "1000", "1010", "1100", "1110", "1111", "2000", "2010", "2011", "2020", "2100", "2110", "2111",
           "2120", "2121", "2200", "2210", "2211", "2220", "2221", "0010", "0020", "1020"]
DDDD = "2222"
LNAME = "spd"

RYS_GLOBAL, RYS_LDS, KSPLIT, SPLIT_B, SPLIT_A, FENCE, ROOTWISE = 256, 128, 64, 32, 16, 8, 1


def cls_index(c):
    la, lb, lc, ld = (int(x) for x in c)
    return ((la * 3 + lb) * 3 + lc) * 3 + ld


def cls_name(c):
    la, lb, lc, ld = (int(x) for x in c)
    return f"({LNAME[la]}{LNAME[lb]}|{LNAME[lc]}{LNAME[ld]})"


def waves(code):
    return (code >> 1) & 3


def with_waves(code, w):
    return (code & ~6) | (w << 1)


def read_table(path=TABLE):
    """{(class index, fp32): code} and {class index: fp32 pays} of a table."""
    txt = open(path).read()
    shapes, pays = {}, {}
    for m in re.finditer(r"case (\d+): return (-?\d+);\s*//\s*shape", txt):
        k, v = int(m.group(1)), int(m.group(2))
        shapes[(k // 2, k % 2 == 1)] = v
    for m in re.finditer(r"case (\d+): return (true|false);\s*//\s*fp32", txt):
        pays[int(m.group(1))] = m.group(2) == "true"
    return shapes, pays


def effective(code, c, fp32):
    """The code with the bits that do not change the compiled kernel
    cleared (the Rys placement bits for FP32, the fence without rootwise)."""
    if not code & ROOTWISE:
        code &= ~FENCE
    if fp32:
        code &= ~(RYS_GLOBAL | RYS_LDS)
    la, lb, lc, ld = (int(x) for x in c)
    nr = (la + lb + lc + ld) // 2 + 1
    if nr <= 3:
        code &= ~RYS_LDS
    else:
        code &= ~RYS_GLOBAL
    return code


def candidates(c, base, fp32):
    """base first, then the shapes one change away from it."""
    la, lb = int(c[0]), int(c[1])
    out = [base]

    def add(code):
        e = effective(code, c, fp32)
        if e not in [effective(x, c, fp32) for x in out]:
            out.append(code)

    add(with_waves(base, 1 if waves(base) == 2 else 2))
    add(base ^ ROOTWISE)
    if base & ROOTWISE:
        add(base ^ FENCE)
    add(base ^ KSPLIT)
    if la == 2:
        add(base ^ SPLIT_A)
    if lb == 2:
        add(base ^ SPLIT_B)
    if not fp32:
        nr = (sum(int(x) for x in c)) // 2 + 1
        add(base ^ (RYS_GLOBAL if nr <= 3 else RYS_LDS))
    return out


def write_table(path, shapes, pays, notes):
    """Write jk3_tuned.inc: shapes {(class index, fp32): code}, pays {class
    index: bool}, notes {(class index, fp32) or class index: comment}."""
    lines = ["// jk3_tuned.inc -- written by python -m patmol_amd.autotune select; do not edit by hand.",
             "// Launch shape per J/K class and precision (the code of jk3_code in",
             "// eri_kernel3.h) and the classes whose FP32 kernel is faster than their",
             "// FP64 kernel (mixed precision runs items in FP32 only there).",
             "#pragma once", "", "namespace pm {", "",
             "constexpr int jk3_tuned_shape(int cls, bool fp32)", "{", "    switch (cls * 2 + (fp32 ? 1 : 0)) {"]
    for (k, f) in sorted(shapes):
        note = notes.get((k, f), "")
        lines.append(f"    case {2 * k + int(f)}: return {shapes[(k, f)]};  // shape {note}".rstrip())
    lines += ["    default: return -1;", "    }", "}", "",
              "constexpr bool jk3_fp32_pays(int cls)", "{", "    switch (cls) {"]
    for k in sorted(pays):
        note = notes.get(k, "")
        lines.append(f"    case {k}: return {'true' if pays[k] else 'false'};  // fp32 {note}".rstrip())
    lines += ["    default: return true;", "    }", "}", "", "}  // namespace pm", ""]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def variant_tables(shapes):
    """[{(class index, fp32): code}] per variant, from the current table."""
    per = {}
    for c in CLASSES:
        k = cls_index(c)
        for fp32 in (False, True):
            base = shapes.get((k, fp32), 3)
            per[(k, fp32)] = candidates(c, base, fp32)
    nv = max(len(v) for v in per.values())
    return [{key: (cl[v] if v < len(cl) else cl[0]) for key, cl in per.items()} for v in range(nv)]


def cmd_build(args):
    shapes, pays = read_table()
    variants = variant_tables(shapes)
    os.makedirs(args.out, exist_ok=True)
    print(f"{len(variants)} variants into {args.out}", flush=True)
    for v, tab in enumerate(variants):
        d = os.path.abspath(os.path.join(args.out, f"v{v}"))
        os.makedirs(d, exist_ok=True)
        full = dict(shapes)
        full.update(tab)
        for attempt in range(3):
            write_table(os.path.join(d, "jk3_tuned.inc"), full, pays, {})
            t0 = time.time()
            r = subprocess.run(["make", "-k", f"-j{args.jobs}", "-C", CSRC, f"OUT={d}", f"OBJ={d}/_build",
                               f"TUNE_INC=-I{d}"], capture_output=True, text=True)
            if r.returncode == 0:
                break
            # a candidate the compiler rejects (e.g. LDS over the limit) falls
            # back to the class's current shape in this variant
            failed = [c for c in CLASSES if not os.path.exists(os.path.join(d, "_build", f"cls_{c}.o"))]
            if not failed:
                sys.stderr.write(r.stdout[-4000:] + r.stderr[-4000:])
                raise SystemExit(f"variant {v}: build failed")
            for c in failed:
                k = cls_index(c)
                for f in (False, True):
                    full[(k, f)] = shapes.get((k, f), 3)
            print(f"  variant {v}: {', '.join(cls_name(c) for c in failed)} back to the current shape", flush=True)
        used = {f"{k}:{int(f)}": full[(k, f)] for (k, f) in tab}
        with open(os.path.join(d, "variant.json"), "w") as fh:
            json.dump(used, fh)
        print(f"  variant {v} built in {time.time() - t0:.0f} s", flush=True)


def _measure_one(args):
    """In a process whose libraries are the variant's: per-class ms (min
    over repeats) with all items FP64, then all items FP32, and G."""
    import numpy as np

    import patmol_amd as pm
    from patmol_amd import systems

    s = getattr(systems, args.system)()
    pm.load_system(s)
    # every class on its Rys kernel: the shapes tuned here are jk_kernel3's
    pm.set_options(tau=1e-12, precision=0, strict=False, timing=True, md_engine=0)
    S, _, H = pm.one_electron_ints()
    n, nocc = s.norb, s.nelec // 2
    il = np.tril_indices(n)

    def full(p):
        F = np.zeros((n, n))
        F[il] = p
        F[(il[1], il[0])] = p
        return F

    sv, U = np.linalg.eigh(full(S))
    X = U / np.sqrt(sv)
    _, V = np.linalg.eigh(X.T @ full(H) @ X)
    C = (X @ V)[:, :nocc]
    P = (C @ C.T)[il].copy()
    st = pm.EriStore()
    out = {}
    G64 = None
    pm.lib()[1].patmol_hip_set_fp32_policy(1)
    for key, prec, thr in (("fp64", 0, 1e-6), ("fp32", 1, 1e300)):
        pm.set_options(precision=prec, mixed_threshold=thr, timing=True)
        best = None
        for rep in range(args.repeats + 1):
            G = st.rhf_gmatrix(P)
            _, cls = st.last_timing()
            cls = np.asarray(cls, dtype=float)
            if rep > 0:
                best = cls if best is None else np.minimum(best, cls)
        out[key] = best.tolist()
        if prec == 0:
            G64 = G
    pm.lib()[1].patmol_hip_set_fp32_policy(0)
    st.close()
    np.save(args.g_out, G64)
    print(json.dumps(out))


def cmd_measure(args):
    import numpy as np

    dirs = sorted((d for d in os.listdir(args.dir) if re.fullmatch(r"v\d+", d)), key=lambda d: int(d[1:]))
    res = {}
    ref = None
    for d in dirs:
        path = os.path.abspath(os.path.join(args.dir, d))
        env = dict(os.environ, PATMOL_AMD_LIB_DIR=path)
        gpath = os.path.join(path, "G.npy")
        print(f"measuring {d} on {args.system} ...", flush=True)
        r = subprocess.run([sys.executable, "-m", "patmol_amd.autotune", "_measure_one", "--system", args.system,
                            "--repeats", str(args.repeats), "--g-out", gpath], env=env, capture_output=True,
                           text=True, timeout=args.timeout)
        if r.returncode != 0:
            sys.stderr.write(r.stdout[-2000:] + r.stderr[-2000:])
            raise SystemExit(f"{d}: measurement failed ({r.returncode})")
        t = json.loads(r.stdout.strip().splitlines()[-1])
        G = np.load(gpath)
        if ref is None:
            ref = G
        dev = float(np.abs(G - ref).max())
        t["max_abs_dG_vs_v0"] = dev
        t["ok"] = dev < 1e-10
        with open(os.path.join(path, "variant.json")) as fh:
            t["codes"] = json.load(fh)
        res[d] = t
        print(f"  {d}: FP64 {sum(t['fp64']):.2f} ms, FP32 {sum(t['fp32']):.2f} ms (classes serialised), "
              f"max|dG| vs v0 {dev:.1e}", flush=True)
    res["_system"] = args.system
    with open(args.out, "w") as fh:
        json.dump(res, fh)
    print(f"timings -> {args.out}")


def cmd_select(args):
    with open(args.timings) as fh:
        res = json.load(fh)
    system = res.pop("_system", "?")
    shapes, pays = read_table(args.table)
    notes = {}
    ok = {d: t for d, t in res.items() if t.get("ok")}
    pc = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    pcx = {p: i for i, p in enumerate(pc)}
    for c in CLASSES + [DDDD]:
        k = cls_index(c)
        la, lb, lc, ld = (int(x) for x in c)
        slot = pcx[(la, lb)] * 6 + pcx[(lc, ld)]   # the 36-entry timing index (bra pc * 6 + ket pc)
        best = {}
        for f, key in ((False, "fp64"), (True, "fp32")):
            cands = [(t[key][slot], t["codes"].get(f"{k}:{int(f)}", shapes.get((k, f), 3)), d) for d, t in ok.items()
                     if t[key][slot] > 0]
            if not cands:
                continue
            ms, code, d = min(cands)
            best[f] = ms
            if c != DDDD:
                shapes[(k, f)] = code
                notes[(k, f)] = f"{cls_name(c)} {'FP32' if f else 'FP64'} {ms:.3f} ms on {system}"
        if False in best and True in best:
            pays[k] = best[True] < 0.97 * best[False]
            notes[k] = f"{cls_name(c)} FP32 {best[True]:.3f} / FP64 {best[False]:.3f} ms"
    write_table(args.table, shapes, pays, notes)
    print(f"wrote {args.table}")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--out", default=os.path.join(os.getcwd(), "_tune"))
    b.add_argument("--jobs", type=int, default=8)
    m = sub.add_parser("measure")
    m.add_argument("--dir", default=os.path.join(os.getcwd(), "_tune"))
    m.add_argument("--system", default="water16")
    m.add_argument("--repeats", type=int, default=3)
    m.add_argument("--timeout", type=int, default=300)
    m.add_argument("--out", default=os.path.join(os.getcwd(), "_tune", "timings.json"))
    s = sub.add_parser("select")
    s.add_argument("--timings", default=os.path.join(os.getcwd(), "_tune", "timings.json"))
    s.add_argument("--table", default=TABLE)
    o = sub.add_parser("_measure_one")
    o.add_argument("--system", default="water16")
    o.add_argument("--repeats", type=int, default=3)
    o.add_argument("--g-out", required=True)
    args = ap.parse_args()
    {"build": cmd_build, "measure": cmd_measure, "select": cmd_select, "_measure_one": _measure_one}[args.cmd](args)


if __name__ == "__main__":
    main()
