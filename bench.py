#!/usr/bin/env python3
"""Benchmark of the SCC-DFTB hot path on one MI355X.

A step is one SCC iteration of the periodic 4096-atom diamond-like C/H/O supercell
(15769 orbitals, Gamma point, Fermi kT = 0.01 Ha): charge potential, H = H0 + 1/2 S(V+V),
unpack, generalized eigensolve, Fermi filling, density matrix, pack, Mulliken, energies and
Broyden mixing, everything of O(N^2) or more on the GPU with the inputs resident in HBM.
After the timed steps one force evaluation (energy-weighted density matrix + pair, gamma and
repulsive gradients) is timed as well.

The dominant kernel is the symmetric product of the Householder tridiagonalisation (HBM bound);
its launches are timed live with HIP events on the library's stream during the timed steps.
The CPU baseline is the numpy/LAPACK oracle timed on the host for SCC iterations of the 1000-atom
cell of the same recipe (3853 orbitals), scaled by N^3 (the dense eigensolve) to 15769 orbitals.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCC iter time + force evals/s @16k orbs: mixed vs FP64 GPU at ΔE<1e-6 Ha/atom"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(sample_cells=5, iters=2):
    """Oracle SCC iterations on the host, N^3-scaled to the benchmark size."""
    import numpy as np
    from dftbplus_amd import synth
    from oracle import dftb_ref as R
    geo, params, opts = synth.workload(f"diamond{sample_cells}")
    opts.max_scc_iter = iters
    calc = R.Calculator(params, geo, opts)
    n = int(calc.off[-1])
    t0 = time.perf_counter()
    calc.scc()
    dt = (time.perf_counter() - t0) / iters
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS", os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    return n, dt, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="diamond8")
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the SCC path of a single Gamma-point system runs on one "
              "device (see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from dftbplus_amd import api, synth
    L = api.lib()
    # every library call ends with a synchronisation of the library's HIP stream; torch is not
    # loaded here (its bundled HIP runtime would be a second runtime in the process)
    sync = lambda: None

    geo, params, opts = synth.workload(args.workload)
    t0 = time.perf_counter()
    dp = api.DftbPlus(params, geo, opts)
    setup_s = time.perf_counter() - t0
    n = dp.get_nr_orbitals()
    print(f"[bench] {args.workload}: {geo.natom} atoms, {n} orbitals, setup {setup_s:.2f}s",
          file=sys.stderr, flush=True)
    dp.scc_begin()
    for i in range(args.warmup):
        dp.scc_step()
        print(f"[bench] warmup step {i + 1} done", file=sys.stderr, flush=True)
    L.dftbhip_kernel_profile(1)
    sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        conv, err = dp.scc_step()
        print(f"[bench] step {i + 1}: charge error {err:.3e}", file=sys.stderr, flush=True)
    sync()
    elapsed = time.perf_counter() - t0
    import numpy as np
    st = np.zeros(3)
    L.dftbhip_kernel_stats(api._d(st))
    L.dftbhip_kernel_profile(0)
    phases = dp.timings()
    dp.scc_end()
    sync()
    t1 = time.perf_counter()
    dp.get_gradients()
    sync()
    force_s = time.perf_counter() - t1
    ms = 1e3 * elapsed / args.steps
    launches, ksec, kbytes = st
    achieved = kbytes / ksec / 1e9 if ksec > 0 else None
    avg_us = 1e6 * ksec / launches if launches else None
    rec = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "SCC iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic diamond-like C/H/O supercell, {geo.natom} atoms, {n} orbitals, "
                               f"Gamma point, Fermi kT=0.01 Ha ({args.workload})",
                   "precision": args.precision, "orbitals": n, "atoms": geo.natom},
        "force_evals_per_s": 1.0 / force_s,
        "force_eval_ms": 1e3 * force_s,
        "phase_s_per_step": {k: v / args.steps for k, v in phases.items()
                             if k not in ("setup", "last_iteration", "forces")},
        "roofline": {"kernel": "symv_kernel (tridiagonalisation, lower-triangle symmetric product)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "launches": int(launches), "avg_launch_us": avg_us,
                     "alg_bytes_per_launch": kbytes / launches if launches else None},
    }
    if not args.no_cpu_baseline:
        ns, dt, threads = cpu_baseline()
        scale = (n / ns) ** 3
        rec["cpu_baseline"] = {"value": 1.0 / (dt * scale), "unit": "SCC iterations/s",
                               "cores": threads, "kind": "port",
                               "sample": f"numpy/LAPACK oracle, 2 SCC iterations of the 1000-atom cell "
                                         f"({ns} orbitals, {dt:.2f} s/iteration) scaled by (N/{ns})^3 to N={n}"}
    print(json.dumps(rec), flush=True)
    dp.close()


if __name__ == "__main__":
    main()
