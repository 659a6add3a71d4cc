#!/usr/bin/env python3
"""ERI + Fock (G) build per SCF iteration on one MI355X.

Workload: (H2O)8 on a 2x2x2 lattice (2.9 A spacing, waters rigidly rotated
by seeded uniform rotations, seed 1001), even-tempered s/p/d Cartesian basis:
232 shells, 448 functions, 80 electrons.  One step is one direct G build
G = 2J[P] - K[P] (every surviving shell quartet's integrals recomputed and
contracted), with the density P and G already resident in HBM, through the
library's device entry point.  The step is timed in FP64 and in mixed
precision (work items whose Schwarz bound times max|P| is below 1e-6 run in
FP32, accumulated in FP64); |dE| of the mixed build against the FP64 build
is reported.  value = surviving shell quartets per second of the mixed
build (of the FP64 build if the mixed one misses |dE| <= 1e-6, runs no FP32
work item or is not faster).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
       python bench.py --calibrate-copy BYTES   (plain device copy of BYTES,
       three launches, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ERI+Fock s/SCF-iter, mixed vs FP64 on 1 GPU at |dE|<=1e-6 Eh; quartets/s"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector = matrix)
TAU = 1e-12
MIXED_THR = 1e-6
CPU_SAMPLE_MOD = 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def core_guess_density(S, H, n, nocc):
    """Core-Hamiltonian guess with canonical orthogonalisation (input prep)."""
    il = np.tril_indices(n)

    def full(p):
        F = np.zeros((n, n))
        F[il] = p
        F[(il[1], il[0])] = p
        return F

    s, U = np.linalg.eigh(full(S))
    X = U / np.sqrt(s)
    e, V = np.linalg.eigh(X.T @ full(H) @ X)
    C = X @ V
    P = C[:, :nocc] @ C[:, :nocc].T
    return P[il].copy()


def elec_energy(P, H, G, n):
    """2 * calc_energy (scfaux.c:99-118): sum_full P (2H + G) / ... as the reference."""
    diag = np.array([i * (i + 1) // 2 + i for i in range(n)])
    w = np.ones_like(P)
    w[diag] = 0.5
    s1 = float(np.sum(w * P * H))
    s2 = float(np.sum(w * P * G))
    return 2.0 * (s1 + s1 + s2)


def time_builds(torch, store, dP, dG, steps, warmup, stream):
    for _ in range(warmup):
        dG.zero_()
        store.rhf_gmatrix_device(dP.data_ptr(), dG.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        dG.zero_()
        store.rhf_gmatrix_device(dP.data_ptr(), dG.data_ptr(), stream.cuda_stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_water8_fp64.json")


def dominant_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), if it lists it."""
    try:
        with open(TRAFFIC_FILE) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None
    for name, v in data.get("kernels", {}).items():
        if kernel and name.startswith(kernel):
            return v
    return None


def cpu_baseline(system, P, kept_total):
    """The oracle's direct G build (the reference's Rys algorithm, canonical
    loop, in-memory scatter) on host cores, over the same Schwarz-screened
    quartet set as the device build, timed over one full build
    (CPU_SAMPLE_MOD = 1, about 10 s on 16 host threads; a larger value times
    the class-representative subset of one canonical tuple in CPU_SAMPLE_MOD)."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    b = oracle.Basis(system)
    q = b.pair_schwarz(nthreads=threads)
    t0 = time.time()
    _, done = b.direct_gmatrix_sampled(P, q, TAU, CPU_SAMPLE_MOD, nthreads=threads)
    dt = time.time() - t0
    rate = done / dt
    which = ("all of them: one full G build" if CPU_SAMPLE_MOD == 1 else
             f"every canonical tuple with (I+3J+7K+13L) % {CPU_SAMPLE_MOD} == 0 of the same G build")
    return {"value": rate, "unit": "quartets/s", "cores": threads, "kind": "port",
            "sample": f"{done} Schwarz-kept shell quartets (tau {TAU:g}; {which}) in {dt:.1f} s",
            "s_per_build_est": kept_total / rate}


def erint_interop(pm, systems, with_cpu):
    """ERINT record interop on a water dimer (s/p/d, 58 shells): the writer
    (every canonical integral above 1e-14 in the reference's order), and the
    record-driven G build from the file and from records in host memory;
    with_cpu adds the oracle's single-thread scatter of the same records."""
    import tempfile

    system = systems.water_cluster(2, 1, 1, seed=7)
    pm.load_system(system)
    pm.set_options(tau=0.0, precision=0, timing=False)
    st = pm.EriStore()
    P = np.eye(system.norb)[np.tril_indices(system.norb)] * 0.1
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "MINTS.DAT")
        t0 = time.perf_counter()
        n = st.write_erints(path)
        t_write = time.perf_counter() - t0
        pm.rhf_gmatrix_erints(None, P, path=path)
        t0 = time.perf_counter()
        for _ in range(3):
            pm.rhf_gmatrix_erints(None, P, path=path)
        t_file = (time.perf_counter() - t0) / 3
        rec = np.fromfile(path, dtype=pm.ERINT_DTYPE)
    t0 = time.perf_counter()
    for _ in range(3):
        pm.rhf_gmatrix_erints(rec, P)
    t_mem = (time.perf_counter() - t0) / 3
    st.close()
    out = {"workload": f"water dimer, {system.nshell} shells, {system.norb} functions", "records": int(n),
           "write_s": t_write, "gmatrix_from_file_ms": t_file * 1e3, "gmatrix_from_records_ms": t_mem * 1e3,
           "records_per_s_from_records": n / t_mem}
    if with_cpu:
        import oracle

        ints = oracle.Basis(system).two_electron_ints()
        t0 = time.perf_counter()
        ints.rhf_gmatrix(P)
        out["cpu_scatter_ms"] = (time.perf_counter() - t0) * 1e3
    return out


def cpu_diag_ms(S, H, nocc):
    """One SCF diagonalisation step by the oracle's single-thread
    restatement: X = U s^-1/2, X^T F X, eigenvectors, C and P, i.e. two
    eigenproblems where an SCF iteration needs one (X is formed once per
    run); the CPU counterpart of the device SCF's work outside the G build."""
    import oracle

    t0 = time.time()
    oracle.diag_fock(S, H, nocc)
    return (time.time() - t0) * 1e3


def calibrate_copy(nbytes):
    """Three launches of a plain device-to-device copy of nbytes: the
    reference reading for the HBM traffic counters."""
    import torch

    n = max(1, nbytes // 8)
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 8, "launches": 3}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--serial-classes", action="store_true",
                    help="launch the class kernels of every build on one stream (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the G build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        return 2

    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)

    import torch
    import patmol_amd as pm
    from patmol_amd import systems

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    system = systems.water8()
    n, nocc = system.norb, system.nelec // 2
    pm.load_system(system)
    pm.set_options(tau=TAU, precision=0, timing=True, strict=False)
    S, T, H = pm.one_electron_ints()
    t0 = time.perf_counter()
    for _ in range(5):
        pm.one_electron_ints()
    ms_oneints = (time.perf_counter() - t0) * 1e3 / 5
    P = core_guess_density(S, H, n, nocc)
    log(f"workload {system.name}: {system.nshell} shells, {n} functions")
    store = pm.EriStore()
    st = store.stats()
    q36, f36 = store.class_stats()
    log(f"kept quartets {st['shell_quartets_kept']} of {st['shell_quartets_total']}, "
        f"algorithmic FLOPs/build {st['flops_per_build']:.4g}, work items {st['work_items']}")
    stream = torch.cuda.current_stream()
    dP = torch.tensor(P, dtype=torch.float64, device="cuda")
    dG = torch.zeros_like(dP)

    # FP64, per-class pass: classes serialised on one stream, each bracketed
    # by HIP events inside the library
    time_builds(torch, store, dP, dG, 1, 1, stream)
    tot_ms, cls_ms = store.last_timing()
    # FP64 timed pass: classes concurrent on side streams
    pm.set_options(timing=args.serial_classes)
    ms64 = time_builds(torch, store, dP, dG, args.steps, args.warmup, stream)
    G64 = dG.cpu().numpy().copy()
    E64 = elec_energy(P, H, G64, n)
    log(f"FP64: {ms64:.3f} ms/build (classes serialised: {tot_ms:.3f} ms), E_elec {E64:.10f}")
    # dominant class kernel (FP64): algorithmic FLOPs / its HIP-event duration
    k = int(np.argmax(cls_ms))
    achieved = f36[k] / (cls_ms[k] * 1e-3) / 1e12 if cls_ms[k] > 0 else 0.0
    names = ["ss", "ps", "pp", "ds", "dp", "dd"]
    dom = f"({names[k // 6]}|{names[k % 6]})"
    dom_kernel = store.class_kernel_name(k // 6, k % 6)
    traffic = dominant_traffic(dom_kernel)
    per_class = {f"({names[i // 6]}|{names[i % 6]})": {"ms": round(float(cls_ms[i]), 4), "quartets": int(q36[i]),
                                                      "tflops": round(f36[i] / (cls_ms[i] * 1e-3) / 1e12, 3)}
                 for i in range(36) if q36[i] > 0 and cls_ms[i] > 0}

    # mixed precision
    pm.set_options(precision=1, mixed_threshold=MIXED_THR)
    msmx = time_builds(torch, store, dP, dG, args.steps, args.warmup, stream)
    Gmx = dG.cpu().numpy().copy()
    Emx = elec_energy(P, H, Gmx, n)
    stm = store.stats()
    pm.set_options(precision=0)
    dE = abs(Emx - E64)
    log(f"mixed: {msmx:.3f} ms/build, |dE| {dE:.3e}, FP32 items {stm['work_items_fp32']}/{stm['work_items']}")

    # reference-compatible entry point: host P in, host G out (PCIe included)
    nh = max(1, min(args.steps, 5))
    store.rhf_gmatrix(P)
    t0 = time.perf_counter()
    for _ in range(nh):
        store.rhf_gmatrix(P)
    ms_host = (time.perf_counter() - t0) * 1e3 / nh
    log(f"host-array entry (rhf_form_gmatrix, PCIe included): {ms_host:.3f} ms/build")

    # the whole RHF iteration on the device (patmol_hip_rhf_scf), same
    # basis and screened store, core-Hamiltonian guess; a run makes
    # iterations + 1 updates (the first is not counted, rhfscf.c:85-88)
    scf, _, _ = pm.rhf_scf_device(store=store)
    nupd = scf["iterations"] + 1
    scf_dev = {"energy": scf["energy"], "iterations": scf["iterations"], "converged": scf["converged"],
               "ms_per_iteration": scf["total_ms"] / nupd, "gbuild_ms_per_iteration": scf["gbuild_ms"] / nupd,
               "other_ms_per_iteration": scf["other_ms"] / nupd, "purify_fallbacks": scf["purify_fallbacks"]}
    log(f"device SCF: E {scf['energy']:.10f} in {scf['iterations']} iterations, "
        f"{scf_dev['ms_per_iteration']:.2f} ms/iteration (G {scf_dev['gbuild_ms_per_iteration']:.2f}, "
        f"rest {scf_dev['other_ms_per_iteration']:.2f}), purification fallbacks {scf['purify_fallbacks']}")

    use_mixed = dE <= 1e-6 and stm["work_items_fp32"] > 0 and msmx < ms64
    ms = msmx if use_mixed else ms64
    kept = st["shell_quartets_kept"]
    value = kept / (ms * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if use_mixed else "f64",
        "data": "synthetic: seeded (H2O)8 geometry, even-tempered basis, core-Hamiltonian guess density",
        "config": {
            "workload": "(H2O)8, 2x2x2 lattice 2.9 A, seed 1001, even-tempered s/p/d (O d 0.3*3^k), "
                        "232 shells, 448 functions; one direct G build per step",
            "schwarz_tau": TAU,
            "mixed_threshold": MIXED_THR,
            "shell_quartets_kept": kept,
            "shell_quartets_total": st["shell_quartets_total"],
            "parallelism": "single GPU",
        },
        "s_per_iter_fp64": ms64 * 1e-3,
        "s_per_iter_mixed": msmx * 1e-3,
        "mixed_speedup": ms64 / msmx if msmx > 0 else None,
        "dE_mixed_vs_fp64": dE,
        "quartets_per_s_fp64": kept / (ms64 * 1e-3),
        "fp32_items": stm["work_items_fp32"],
        "ms_per_step_host_entry": ms_host,
        "one_electron_ints_ms": ms_oneints,
        "scf_device": scf_dev,
        "roofline": {
            "bound": "mfma",
            "kernel": f"jk_kernel {dom} FP64",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_fetch": traffic["fetch_bytes"] if traffic else None,
            "traffic_write": traffic["write_bytes"] if traffic else None,
            "kernel_symbol": dom_kernel,
            "build_tflops": st["flops_per_build"] / (ms64 * 1e-3) / 1e12,
            "build_frac": st["flops_per_build"] / (ms64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        },
        "per_class_fp64": per_class,
        "serialised_class_ms": tot_ms,
    }
    if not args.no_cpu:
        log("cpu baseline (oracle direct build, bounded sample) ...")
        out["cpu_baseline"] = cpu_baseline(system, P, kept)
        out["scf_device"]["cpu_diag_ms"] = cpu_diag_ms(S, H, nocc)
    store.close()
    out["erint_interop"] = erint_interop(pm, systems, not args.no_cpu)
    log(f"ERINT interop: {out['erint_interop']}")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
