#!/usr/bin/env python3
"""Benchmark of the SCC-DFTB hot path on one MI355X.

A step is one SCC iteration of the periodic 4096-atom diamond-like C/H/O supercell
(15769 orbitals, Gamma point, Fermi kT = 0.01 Ha): charge potential, H = H0 + 1/2 S(V+V),
unpack, generalized eigensolve, Fermi filling, density matrix, pack, Mulliken, energies and
Broyden mixing, everything of O(N^2) or more on the GPU with the inputs resident in HBM.
A force evaluation is a converged SCC cycle from neutral charges (charge error below 1e-8 e)
followed by the energy-weighted density matrix and the pair, gamma and repulsive gradients.

Each leg runs one full SCC cycle from neutral charges on the same input. Inside the cycle, the
first W iterations are warmup, the next K iterations are the timed steps (with the SYMV launch
profile), and the cycle then continues untimed to convergence; the whole cycle plus the
gradients is the leg's force evaluation. Legs:
  fp64             the plain FP64 path;
  mixed_then_fp64  FP32 MFMA reduction, tridiagonalisation, back-transformation and density
                   matrices while the previous iteration's charge error is above 1e-3, plain
                   FP64 iterations after (the converged point is the FP64 one);
  mixed_tail       FP32 iterations above 1e-3, then one plain FP64 iteration that keeps all n
                   eigenvectors, and after it no eigensolve: two Ogita-Aishima steps refine the
                   previous iteration's FP64 eigenpairs against the new sparse H and S in
                   residual form (block-sparse products; X^T S X and X^T H X on their diagonal
                   tile band, the Rayleigh quotients, the in-band couplings and the residuals
                   H X - S X diag(lam) with the in-band part removed, all in FP64; the products
                   with the residuals and with the correction in FP16, FP32 sums);
  mixed_tail_fp32seed  as mixed_tail, seeded by FP32 eigenpairs refined by two steps (only
                   with --legs);
  mixed_tail32     as mixed_tail with X^T H X and X^T S X as FP64 GEMMs and the correction
                   product in FP32 (only with --legs);
  mixed_tail_early as mixed_tail with the switch at 3e-3 (only with --legs);
  mixed_fp32       FP32 iterations throughout (only with --legs; does not reach 1e-8 e).
`value` is the SCC iterations/s of the fp64 leg's timed steps; every leg reports its timed
rate, its converged cycle (iterations, seconds, force evaluations/s) and its energy, charge and
force differences from the fp64 leg at the converged point against |dE| < 1e-6 Ha/atom and
max |dF| < 1e-5 Ha/bohr.
Before the legs, the SYMV tile count is autotuned at the workload's size (FP64 and FP32; the
tables and the tuning time are reported, `--no-autotune` keeps the fixed thresholds).

The dominant kernel is the symmetric product of the Householder tridiagonalisation (HBM bound);
its launches are timed live with HIP events on the library's stream during the timed steps.
The CPU baseline is the numpy/LAPACK oracle timed on the host: the second SCC iteration of a
bounded sample (the 1000-atom cell of the same recipe, 3850 orbitals) scaled by N^3 (the dense
eigensolve dominates), with `--cpu-full` one SCC iteration of the benchmark system itself.
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
# FP64 symv_kernel HBM traffic per launch from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE)
# of `bench.py --mode sytrd --n 2048`, each scaled by the counter's reading for the copy kernel
# of the same run (FETCH_SIZE reads 0.5005x the copied bytes, WRITE_SIZE 1.000x):
# profiles/symv_hbm_counters_n2048.md
SYMV_TRAFFIC_N2048 = {"n": 2048, "alg_bytes_per_launch": 5611520, "hbm_read_bytes": 7212843,
                      "hbm_write_bytes": 183350, "traffic": 7396194, "traffic_over_alg": 1.318,
                      "fetch_scale": 1.998, "write_scale": 1.000}
LEGS = {"fp64": dict(precision=0),
        "mixed_then_fp64": dict(precision=1, refine_below=1e-3, refine_steps=0),
        "mixed_fp32": dict(precision=1, refine_below=0.0),
        "mixed_then_refined": dict(precision=1, refine_below=1e-3, refine_steps=1),
        "mixed_tail": dict(precision=1, refine_below=1e-3, refine_steps=0, tail_below=1e-3,
                           tail_steps=2),
        "mixed_tail_fp32seed": dict(precision=1, refine_below=1e-3, refine_steps=2,
                                    tail_below=1e-3, tail_steps=2),
        "mixed_tail32": dict(precision=1, refine_below=1e-3, refine_steps=0, tail_below=1e-3,
                             tail_steps=2, tuning={8: 1}),
        "mixed_tail_early": dict(precision=1, refine_below=3e-3, refine_steps=0,
                                 tail_below=3e-3, tail_steps=2)}
# library defaults of the tuning parameters a leg or mode may change (restored after it)
TUNING_DEFAULTS = {8: 2, 9: 1, 10: 1, 11: 0}
E_BOUND, F_BOUND = 1e-6, 1e-5


def host_cpu():
    """CPU model and the number of cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return model, ncpu


def blas_threads():
    """Threads OpenBLAS uses (threadpoolctl when present, else the environment)."""
    try:
        from threadpoolctl import threadpool_info
        n = [x["num_threads"] for x in threadpool_info() if x.get("internal_api") == "openblas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return int(os.environ.get("OPENBLAS_NUM_THREADS",
                              os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))


def cpu_iteration(workload, timing_gamma=False):
    """Two SCC iterations of the oracle on the host, each timed on its own (potential, H,
    dense generalized eigensolve, Fermi filling, density matrix, pack, Mulliken, energies and
    Broyden mixing, the steps of its SCC loop); returns the orbital count, the set-up time and
    the two iteration times (the second one is the baseline: the first one also pays one-off
    costs of the first LAPACK calls). timing_gamma: the atomic gamma matrix is the cluster 1/R
    instead of the Ewald sum, whose serial set-up takes minutes at 4096 atoms; an iteration
    costs the same (one product of the matrix with the charges)."""
    import numpy as np
    from dftbplus_amd import synth
    from oracle import dftb_ref as R
    geo, params, opts = synth.workload(workload)
    t0 = time.perf_counter()
    gamma = None
    if timing_gamma:
        gamma = R.invr_cluster(geo.coords)
        gamma[np.diag_indices(geo.natom)] = 0.4
    calc = R.Calculator(params, geo, opts, gamma=gamma)
    S = calc.dense(calc.s)
    setup = time.perf_counter() - t0
    n = int(calc.off[-1])
    nat = geo.natom
    qin = calc.q0.copy()
    mixer = R.BroydenMixer(nat, opts.mix_param, opts.omega0, opts.min_weight, opts.max_weight,
                           opts.weight_fac, opts.max_scc_iter + 1)
    times = []
    for _ in range(2):
        t1 = time.perf_counter()
        vorb, _ = calc.orbital_potential(qin - calc.q0)
        h = R.add_shift(calc.h0, calc.s, vorb, geo.species, calc.sp, calc.off)
        eig, C, f, ef, eband, ts, rho = calc._density(h, S)
        qorb = R.mulliken(calc.s, rho, geo.species, calc.sp, calc.off)
        qout = np.add.reduceat(qorb, calc.off[:-1])
        float(np.sum(R.mulliken(calc.h0, rho, geo.species, calc.sp, calc.off)))
        calc.orbital_potential(qout - calc.q0)
        qin = mixer.mix(qin, qout - qin)
        times.append(time.perf_counter() - t1)
    return n, setup, times[0], times[1]


def run_cycle_leg(api, geo, params, opts, leg, steps, warmup):
    """One SCC cycle from neutral charges: warmup iterations, the timed steps (with the SYMV
    launch profile), the rest of the cycle to convergence, then the gradients. Every library
    call ends with a synchronisation of the library's stream. If the cycle converges before the
    timed steps are done, the remaining steps repeat the converged iteration (unchanged input
    charges), so exactly `steps` iterations are timed."""
    import numpy as np
    L = api.lib()
    tag = f"[bench:{leg}]"
    kw = dict(LEGS[leg])
    tuning = kw.pop("tuning", {})
    for p, v in tuning.items():
        api.set_tuning(p, v)
    t0 = time.perf_counter()
    dp = api.DftbPlus(params, geo, opts, **kw)
    setup_s = time.perf_counter() - t0
    n = dp.get_nr_orbitals()
    print(f"{tag} {geo.natom} atoms, {n} orbitals, setup {setup_s:.2f}s", file=sys.stderr,
          flush=True)
    dp.scc_begin()
    it, conv, err = 0, False, float("inf")
    conv_it, conv_s = None, None
    timed_s, st, ph0, ph1 = None, np.zeros(3), None, None
    t_cycle = time.perf_counter()
    while True:
        if it == warmup:
            ph0 = dp.timings()
            L.dftbhip_kernel_profile(1)
            t_win = time.perf_counter()
        t_it = time.perf_counter()
        conv, err = dp.scc_step()
        it += 1
        now = time.perf_counter()
        extra = ""
        if kw.get("tail_below", 0.0) > 0.0:
            state = dp.scc_state()
            if state["oa_clipped_pairs"] or state["oa_far_pairs"]:
                extra = f" (pairs outside the first-order range {state['oa_clipped_pairs']}, " \
                        f"beyond the band {state['oa_far_pairs']})"
        print(f"{tag} iteration {it}: {now - t_it:.3f}s charge error {err:.3e}"
              + (" (timed)" if warmup < it <= warmup + steps else "") + extra, file=sys.stderr,
              flush=True)
        if it == warmup + steps:
            timed_s = now - t_win
            L.dftbhip_kernel_stats(api._d(st))
            L.dftbhip_kernel_profile(0)
            ph1 = dp.timings()
        if conv and conv_it is None:
            conv_it, conv_s = it, now - t_cycle
        if conv_it is not None and it >= warmup + steps:
            break
        if conv_it is None and it >= opts.max_scc_iter:
            break
    if timed_s is None:
        raise RuntimeError(f"{tag} the SCC cycle ended before the timed steps")
    dp.scc_end()
    energy, charges = dp.get_energy(), dp.get_gross_charges()
    t1 = time.perf_counter()
    grad = dp.get_gradients()
    grad_s = time.perf_counter() - t1
    print(f"{tag} timed {1e3 * timed_s / steps:.1f} ms/iteration; cycle "
          f"{'converged' if conv_it else 'NOT converged'} in {conv_it or it} iterations, "
          f"{conv_s or 0:.1f}s; gradients {1e3 * grad_s:.1f} ms", file=sys.stderr, flush=True)
    dp.close()
    for p in tuning:
        api.set_tuning(p, TUNING_DEFAULTS[p])
    launches, ksec, kbytes = st
    achieved = kbytes / ksec / 1e9 if ksec > 0 else None
    return {
        "n": n, "elapsed": timed_s, "setup_s": setup_s, "grad_s": grad_s,
        "converged": conv_it is not None, "iterations": conv_it or it,
        "scc_s": conv_s, "force_s": (conv_s + grad_s) if conv_s else None,
        "phases": {k: (ph1[k] - ph0[k]) / steps for k in ph1
                   if k not in ("setup", "last_iteration", "forces")},
        "energy": energy, "charges": charges, "grad": grad,
        "symv": {"launches": int(launches), "seconds": ksec, "bytes": kbytes,
                 "achieved": achieved},
    }


def mode_tail(args):
    """The mixed-precision tail of the cycle, for kernel traces of its iterations. Without
    --charges-in: runs the mixed_tail leg's cycle until its charge error is at most the tail
    threshold (1e-3) and saves the input charges of the next iteration to --charges-out (no
    tail iteration runs). With --charges-in: starts the mixed_tail leg's cycle from those
    charges (ReadInitialCharges): one FP32 iteration, the plain FP64 seed, then the tail
    iterations to convergence, each printed with its time."""
    import dataclasses
    import numpy as np
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload(args.workload)
    if args.mix:
        opts = dataclasses.replace(opts, mix_param=args.mix, max_scc_iter=300)
    kw = dict(LEGS["mixed_tail"])
    dp = api.DftbPlus(params, geo, opts, **kw)
    n = dp.get_nr_orbitals()
    if args.charges_in:
        dp.set_input_charges(np.load(args.charges_in))
    dp.scc_begin()
    times, conv, it = [], False, 0
    while not conv and it < opts.max_scc_iter:
        t0 = time.perf_counter()
        conv, err = dp.scc_step()
        it += 1
        times.append(time.perf_counter() - t0)
        st = dp.scc_state()
        print(f"[tail] iteration {it}: {times[-1]:.3f}s charge error {err:.3e}, pairs outside "
              f"the first-order range {st['oa_clipped_pairs']} (beyond the band: "
              f"{st['oa_far_pairs']})", file=sys.stderr, flush=True)
        if not args.charges_in and err <= kw["tail_below"]:
            # the FP32 iterations conserve the electron count only to their rounding (about
            # 0.3 e of 16179 here): rescaled to the count, as a charge file must hold it
            q = dp.get_input_charges()
            nel = sum(sum(params.species[int(z)].occ) for z in geo.species)
            np.save(args.charges_out, q * (nel / q.sum()))
            print(f"[tail] saved the input charges of iteration {it + 1} to {args.charges_out}",
                  file=sys.stderr, flush=True)
            break
    dp.scc_end()
    dp.close()
    print(json.dumps({"mode": "tail", "workload": args.workload, "orbitals": n,
                      "charges_in": args.charges_in, "iterations": it, "converged": conv,
                      "iteration_s": times}), flush=True)


def mode_cycle(args):
    """One converged SCC cycle per leg of --legs, without the FP64 reference leg or the CPU
    baseline (for kernel traces of a whole cycle, e.g. rocprofv3 -P over its tail); prints the
    per-iteration lines and one JSON summary line per leg."""
    import dataclasses
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload(args.workload)
    if args.mix:
        opts = dataclasses.replace(opts, mix_param=args.mix, max_scc_iter=300)
    import numpy as np
    res = {}
    for leg in [x for x in args.legs.split(",") if x]:
        r = run_cycle_leg(api, geo, params, opts, leg, args.steps, args.warmup)
        res[leg] = r
        first = next(iter(res.values()))
        print(json.dumps({"mode": "cycle", "workload": args.workload, "leg": leg,
                          "orbitals": r["n"], "converged": r["converged"],
                          "iterations": r["iterations"], "cycle_scc_s": r["scc_s"],
                          "energy": r["energy"],
                          "vs_first_leg": {
                              "dE_per_atom": abs(r["energy"] - first["energy"]) / geo.natom,
                              "dq_max": float(np.max(np.abs(r["charges"] - first["charges"]))),
                              "dF_max": float(np.max(np.abs(r["grad"] - first["grad"])))}}),
              flush=True)


def mode_tracesum(args):
    """Per-iteration kernel table of the mixed-precision tail from a rocprofv3 trace database
    (--trace, the rocpd SQLite output of --kernel-trace): the kernels of the last --steps
    tail iterations (from their first spmm_fold_t_kernel launch on, two launches per tail
    iteration; the cycle must end in the tail), by total time."""
    import collections
    import re
    import sqlite3
    con = sqlite3.connect(args.trace)
    rows = list(con.execute("select name, start, end, duration from kernels order by start"))
    marks = [r for r in rows if "spmm_fold_t_kernel" in r[0]]
    if not marks:
        raise SystemExit("no spmm_fold_t_kernel launch in the trace")
    nit = min(len(marks) // 2, max(args.steps, 1))
    win = [r for r in rows if r[1] >= marks[-2 * nit][1]]
    names = sorted({r[0] for r in win if r[0].startswith("_Z")})
    plain = {}
    if names:
        import subprocess
        try:   # the trace keeps some kernel names mangled
            out = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True,
                                 text=True, timeout=30).stdout.splitlines()
            plain = dict(zip(names, out)) if len(out) == len(names) else {}
        except (OSError, subprocess.SubprocessError):
            plain = {}
    agg = collections.OrderedDict()
    for name, st, en, du in win:
        name = plain.get(name, name)
        m = re.match(r"_ZN7dftbhip(\d+)", name)   # a name the demangler left: its identifier
        if m:
            name = "dftbhip::" + name[m.end():m.end() + int(m.group(1))]
        k = re.sub(r"^void ", "", re.sub(r"\(.*", "", name))[:64]
        a = agg.setdefault(k, [0, 0])
        a[0] += 1
        a[1] += du
    tot = sum(a[1] for a in agg.values())
    span = (win[-1][2] - win[0][1]) / 1e6
    print(f"tail iterations: {nit}; kernel time {tot / 1e6 / nit:.1f} ms and span "
          f"{span / nit:.1f} ms per iteration")
    print()
    print("| kernel | launches / iteration | ms / iteration | average µs |")
    print("|---|---:|---:|---:|")
    for k, (c, d) in sorted(agg.items(), key=lambda x: -x[1][1]):
        if d / nit < 1e4:   # below 0.01 ms per iteration
            continue
        print(f"| `{k}` | {c / nit:.1f} | {d / 1e6 / nit:.2f} | {d / c / 1e3:.1f} |")


def mode_cpu(args):
    """The CPU baseline alone: SCC iterations of the oracle on the host for --workload (at the
    full size of the 4096-atom cell, with the cluster 1/R gamma to skip the serial Ewald set-up)
    and, for comparison, the bounded sample --cpu-sample scaled by N^3."""
    model, ncpu = host_cpu()
    out = {"mode": "cpu", "cpu_model": model, "host_cpus": ncpu, "blas_threads": blas_threads()}
    for wl, tg in ((args.cpu_sample, False), (args.workload, True)):
        print(f"[cpu] {wl}: set-up and two SCC iterations", file=sys.stderr, flush=True)
        n, setup, first, second = cpu_iteration(wl, timing_gamma=tg)
        out[wl] = {"orbitals": n, "setup_s": setup, "first_s": first, "second_s": second,
                   "cluster_gamma": tg}
        print(f"[cpu] {wl}: {n} orbitals, set-up {setup:.1f}s, iterations {first:.2f}s "
              f"{second:.2f}s", file=sys.stderr, flush=True)
    ns, nf = out[args.cpu_sample]["orbitals"], out[args.workload]["orbitals"]
    out["sample_scaled_s"] = out[args.cpu_sample]["second_s"] * (nf / ns) ** 3
    print(json.dumps(out), flush=True)


def mode_sytrd(args):
    """Copy kernel (counter calibration: 8 n^2 bytes read and written) and the one-stage
    tridiagonalisation of a random symmetric n x n matrix."""
    import numpy as np
    from dftbplus_amd import api
    L = api.lib()
    n = args.n
    rng = np.random.default_rng(0)
    a = rng.standard_normal((n, n))
    a = a + a.T
    api.dlacpy(a)
    api.set_tuning(0, args.tiles)
    L.dftbhip_kernel_profile(1)
    t0 = time.perf_counter()
    api.dsytrd(a)
    dt = time.perf_counter() - t0
    st = np.zeros(3)
    L.dftbhip_kernel_stats(api._d(st))
    L.dftbhip_kernel_profile(0)
    launches, ksec, kbytes = st
    print(json.dumps({"mode": "sytrd", "n": n, "symv_tiles": args.tiles or "by size",
                      "seconds_incl_transfers": dt,
                      "symv_launches": int(launches), "symv_avg_us": 1e6 * ksec / launches,
                      "symv_alg_bytes_per_launch": kbytes / launches,
                      "symv_achieved_GBs": kbytes / ksec / 1e9,
                      "copy_kernel_bytes_each_way": 8 * n * n}), flush=True)


def mode_batch(args):
    """The molecule batch (SCC + forces of nmol molecules in one set of launches) against a loop
    of single-molecule runs of the plain path over every 128th molecule, scaled to nmol."""
    import numpy as np
    from dftbplus_amd import api, synth
    mols, params, opts = synth.workload(f"molecules{args.nmol}")
    t0 = time.perf_counter()
    b = api.MoleculeBatch(params, mols, opts)
    t1 = time.perf_counter()
    print(f"[bench:batch] {len(mols)} molecules set up in {t1 - t0:.2f}s", file=sys.stderr,
          flush=True)
    b.scc_begin()
    it = 0
    while True:
        ts = time.perf_counter()
        nactive = b.scc_step()
        it += 1
        if it % 10 == 1 or nactive == 0:
            print(f"[bench:batch] iteration {it}: {time.perf_counter() - ts:.3f}s, "
                  f"{nactive} molecules not converged", file=sys.stderr, flush=True)
        if nactive == 0 or it >= opts.max_scc_iter:
            break
    e = b.scc_end()
    t2 = time.perf_counter()
    g = b.gradients()
    t3 = time.perf_counter()
    print(f"[bench:batch] {len(mols)} molecules: setup {t1 - t0:.2f}s scc {t2 - t1:.2f}s "
          f"gradients {t3 - t2:.2f}s", file=sys.stderr, flush=True)
    sample = list(range(0, len(mols), 128))
    dev = 0.0
    t4 = time.perf_counter()
    for m in sample:
        dp = api.DftbPlus(params, mols[m], opts)
        em = dp.get_energy()
        gm = dp.get_gradients()
        dev = max(dev, abs(em - e[m]) / mols[m].natom, float(np.max(np.abs(gm - g[m]))))
        dp.close()
    loop_s = (time.perf_counter() - t4) * len(mols) / len(sample)
    batch_s = t3 - t0
    print(json.dumps({"mode": "batch", "molecules": len(mols),
                      "orbitals": int(sum(int(np.sum([params.species[x].norb for x in mm.species]))
                                          for mm in mols)),
                      "batch_setup_s": t1 - t0, "batch_scc_s": t2 - t1,
                      "batch_gradients_s": t3 - t2, "batch_total_s": batch_s,
                      "converged": int(np.sum(b.converged)),
                      "mean_scc_iterations": float(np.mean(b.n_iter)),
                      "loop_sample": len(sample), "loop_estimate_s": loop_s,
                      "speedup_vs_loop": loop_s / batch_s,
                      "sample_max_deviation": dev}), flush=True)
    b.close()


def mode_neighbours(args):
    """The neighbour list of the workload's geometry built on the device (with the download of
    the reference-layout arrays) and the geometry set-up of a calculator (neighbour list, pair
    and SK block lists, H0/S, gamma, overlap factorisation) with the list built on the device
    and on the host."""
    import numpy as np
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload(args.workload)
    cut = max(params.sk_cutoff(), params.rep_cutoff())
    api.neighbour_list(geo.coords, geo.latvecs, cut)   # first call: runtime and kernels loaded
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        call, i2c, neigh, _ = api.neighbour_list(geo.coords, geo.latvecs, cut)
    t_list = (time.perf_counter() - t0) / reps
    setup = {}
    try:
        for mode, name in ((1, "device"), (0, "host")):
            api.set_tuning(3, mode)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                dp = api.DftbPlus(params, geo, opts)
                ts.append(time.perf_counter() - t0)
                dp.close()
            setup[name] = min(ts)
    finally:
        api.set_tuning(3, 1)
    print(json.dumps({"mode": "neighbours", "workload": args.workload, "atoms": geo.natom,
                      "cutoff": cut, "images": int(len(i2c)),
                      "pairs": int(sum(len(v) for v in neigh)),
                      "device_list_s": t_list,
                      "geometry_setup_s": setup}), flush=True)


def mode_kernels(args):
    """One SCC iteration and one force evaluation of the workload, then an FP64 GEMM of size --n
    (NN and NT) through the C ABI, for kernel traces and counter runs; prints the sizes that fix
    the algorithmic bytes of the sparse kernels (SK blocks, sparse elements) and the flops of the
    density SYRK (orbitals, filled levels)."""
    import numpy as np
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload(args.workload)
    dp = api.DftbPlus(params, geo, opts)
    dp.scc_begin()
    dp.scc_step()
    dp.scc_end()
    dp.get_gradients()
    _, f = dp.get_eigenvalues()
    n = dp.get_nr_orbitals()
    dp.close()
    nocc = int(np.nonzero(f > np.finfo(float).eps)[0].max()) + 1 if np.any(f > 0) else 0
    norb = np.array([params.species[x].norb for x in geo.species])
    _, i2c, neigh, _ = api.neighbour_list(geo.coords, geo.latvecs, params.sk_cutoff())
    nblk = int(sum(len(v) for v in neigh))
    nnz = int(sum(int(norb[a]) * int(norb[i2c[neigh[a]]].sum()) for a in range(geo.natom)))
    rng = np.random.default_rng(0)
    a = rng.standard_normal((args.n, args.n))
    b = rng.standard_normal((args.n, args.n))
    # the GEMMs in both tile orders (launch grid, XCD-grouped): in the kernel trace the launches
    # follow this sequence
    for order in (0, 1):
        api.set_tuning(5, order)
        for ta, tb in (("N", "N"), ("N", "T"), ("T", "N")):
            t0 = time.perf_counter()
            api.dgemm(a, b, ta, tb)
            print(f"[kernels] gemm order {order} {ta}{tb} n={args.n}: "
                  f"{time.perf_counter() - t0:.3f}s with transfers", file=sys.stderr, flush=True)
    api.set_tuning(5, 1)
    print(json.dumps({"mode": "kernels", "workload": args.workload, "atoms": geo.natom,
                      "orbitals": n, "sk_blocks": nblk, "sparse_elements": nnz,
                      "filled_levels": nocc, "gemm_n": args.n}), flush=True)


def mode_kpoints(args):
    """SCC iterations and one force evaluation of a periodic workload with a 2x2x2 mesh:
    Gamma-centred (k in {0, 1/2}^3, the eight time-reversal-invariant points, whose H(k) is
    real) or shifted Monkhorst-Pack (k in {-1/4, 1/4}^3, complex). With --md-steps S: a cold
    converged SCC and forces, then S MD-like steps, each moving every coordinate by a uniform
    random amount in [-0.05, 0.05] bohr (seed 4) and running a converged SCC warm-started from
    the previous step's input charges plus the forces; once per leg of --legs (the same
    displacements), each leg's energies and forces compared with the first leg's."""
    import dataclasses
    import numpy as np
    from dftbplus_amd import api, synth
    geo, params, opts = synth.workload(args.workload)
    pts = (0.0, 0.5) if args.kmesh == "gc" else (-0.25, 0.25)
    mesh = np.array([[a, b, c] for a in pts for b in pts for c in pts])
    opts = dataclasses.replace(opts, kpoints=mesh)
    api.set_tuning(0, args.tiles)
    t0 = time.perf_counter()
    dp = api.DftbPlus(params, geo, opts)
    setup = time.perf_counter() - t0
    n = dp.get_nr_orbitals()
    out = {"mode": "kpoints", "workload": args.workload, "atoms": geo.natom, "orbitals": n,
           "kmesh": args.kmesh, "kpoints": len(mesh), "setup_s": setup}
    if args.md_steps <= 0:
        dp.scc_begin()
        times = []
        for i in range(args.steps):
            t1 = time.perf_counter()
            conv, err = dp.scc_step()
            times.append(time.perf_counter() - t1)
            print(f"[bench:kpoints] step {i + 1}: {times[-1]:.2f}s charge error {err:.3e}",
                  file=sys.stderr, flush=True)
        dp.scc_end()
        t2 = time.perf_counter()
        dp.get_gradients()
        out.update({"scc_iteration_s": times, "force_eval_s": time.perf_counter() - t2,
                    "phase_s": dp.timings()})
    else:
        dp.close()
        legs = [x for x in args.legs.split(",") if x]
        out["legs"] = {}
        first = None
        for leg in legs:
            kw = dict(LEGS[leg])
            tuning = kw.pop("tuning", {})
            for p_, v_ in tuning.items():
                api.set_tuning(p_, v_)
            t0 = time.perf_counter()
            dp = api.DftbPlus(params, geo, opts, **kw)
            rec = md_steps(api, dp, geo, args, time.perf_counter() - t0, leg, opts.max_scc_iter)
            dp.close()
            for p_ in tuning:
                api.set_tuning(p_, TUNING_DEFAULTS[p_])
            if first is None:
                first = rec
            rec["max_dE_per_atom_vs_first_leg"] = float(max(
                abs(a["energy"] - b["energy"]) for a, b in zip(rec["steps"], first["steps"])) /
                geo.natom)
            rec["max_dF_vs_first_leg"] = float(max(
                np.abs(a["_grad"] - b["_grad"]).max()
                for a, b in zip(rec["steps"], first["steps"])))
            out["legs"][leg] = rec
        for rec in out["legs"].values():
            for x in rec["steps"]:
                x.pop("_grad", None)
        print(json.dumps(out), flush=True)
        return
    print(json.dumps(out), flush=True)
    dp.close()


def md_steps(api, dp, geo, args, setup, leg, max_it):
    """--md-steps MD-like steps of one leg (see mode_kpoints), each SCC cycle stepped here with
    one line per iteration (at most max_it); the gradients of each step are kept under "_grad"
    for the comparison between legs."""
    import numpy as np
    dp.set_charge_restart(True)
    rng = np.random.default_rng(4)
    coords = geo.coords.copy()
    steps = []
    for step in range(args.md_steps + 1):
        if step > 0:
            coords = coords + rng.uniform(-0.05, 0.05, coords.shape)
            t_geo = time.perf_counter()
            dp.set_geometry(coords, geo.latvecs)
            geo_s = time.perf_counter() - t_geo
        else:
            geo_s = setup
        t1 = time.perf_counter()
        dp.scc_begin()
        it, conv = 0, False
        while not conv and it < max_it:
            t_it = time.perf_counter()
            conv, err = dp.scc_step()
            it += 1
            print(f"[bench:kpoints:{leg}] MD step {step} iteration {it}: "
                  f"{time.perf_counter() - t_it:.2f}s charge error {err:.3e}", file=sys.stderr,
                  flush=True)
        dp.scc_end()
        e = dp.get_energy()
        scc_s = time.perf_counter() - t1
        t2 = time.perf_counter()
        g = dp.get_gradients()
        grad_s = time.perf_counter() - t2
        rec = {"step": step, "warm": step > 0, "iterations": dp.n_iter,
               "converged": dp.converged, "geometry_s": geo_s, "scc_s": scc_s,
               "gradients_s": grad_s, "energy": e, "max_force": float(np.abs(g).max()),
               "_grad": g}
        steps.append(rec)
        print(f"[bench:kpoints:{leg}] MD step {step}: {dp.n_iter} iterations, SCC "
              f"{scc_s:.2f}s, forces {grad_s:.2f}s, converged {dp.converged}",
              file=sys.stderr, flush=True)
    warm = steps[1:]
    per_step = [x["geometry_s"] + x["scc_s"] + x["gradients_s"] for x in warm]
    return {"md_steps": args.md_steps, "displacement_bohr": 0.05, "steps": steps,
            "cold_force_eval_s": steps[0]["scc_s"] + steps[0]["gradients_s"],
            "warm_step_s_mean": float(np.mean(per_step)),
            "warm_force_evals_per_s": float(1.0 / np.mean(per_step)),
            "warm_iterations_mean": float(np.mean([x["iterations"] for x in warm])),
            "all_converged": all(x["converged"] for x in steps),
            "phase_s_last_step": dp.timings()}


def mode_autograd(args):
    """The differentiation form at the water-box size: energy and forces through the
    torch.autograd.Function (analytic adjoint backward, library in a worker process) in FP64
    and mixed precision with refinement; central finite differences (h = 1e-4 bohr) of the FP64
    energy on a few coordinates; the mixed gradients against the FP64 ones."""
    import numpy as np
    import torch
    from dftbplus_amd import synth, torch_api
    geo, params, opts = synth.workload(args.workload)
    out = {"mode": "autograd", "workload": args.workload, "atoms": geo.natom}
    grads = {}
    # mixed: FP32 bulk with two refinement steps every iteration (this box's charges move too
    # slowly under its mixing for the FP32-only iterations to reach a switch point)
    for name, kw in (("fp64", {}), ("mixed_refined", dict(precision=1, refine_below=float("inf"),
                                                          refine_steps=2))):
        calc = torch_api.Calculator(params, geo, opts, force_mode=1, **kw)
        x = torch.tensor(geo.coords, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        try:
            e = torch_api.energy(x, calc)
        except RuntimeError as err:   # an SCC cycle that does not converge is reported as such
            out[name] = {"error": str(err), "seconds": time.perf_counter() - t0}
            print(f"[bench:autograd] {name}: {err}", file=sys.stderr, flush=True)
            calc.close()
            continue
        t1 = time.perf_counter()
        e.backward()
        t2 = time.perf_counter()
        grads[name] = x.grad.numpy().copy()
        rec = {"energy": float(e.item()), "scc_s": t1 - t0, "backward_s": t2 - t1,
               "force_evals_per_s_incl_scc": 1.0 / (t2 - t0), "backward_per_s": 1.0 / (t2 - t1)}
        print(f"[bench:autograd] {name}: SCC {t1 - t0:.2f}s backward {t2 - t1:.3f}s",
              file=sys.stderr, flush=True)
        if name == "fp64":
            h = 1e-4
            rng = np.random.default_rng(48)
            idx = rng.choice(3 * geo.natom, size=args.fd_coords, replace=False)
            c0 = geo.coords.reshape(-1).copy()
            err = 0.0
            for k in idx:
                c = c0.copy()
                c[k] += h
                ep = calc.energy(c)
                c[k] -= 2 * h
                em = calc.energy(c)
                dk = abs((ep - em) / (2 * h) - grads[name].reshape(-1)[k])
                err = max(err, dk)
                print(f"[bench:autograd] FD coordinate {k}: |FD - adjoint| {dk:.2e}",
                      file=sys.stderr, flush=True)
            rec["fd_coords"] = int(args.fd_coords)
            rec["fd_max_err"] = err
        calc.close()
        out[name] = rec
    if "fp64" in grads and "mixed_refined" in grads:
        out["mixed_vs_fp64"] = {
            "dE_per_atom": abs(out["mixed_refined"]["energy"] - out["fp64"]["energy"]) / geo.natom,
            "dF_max": float(np.max(np.abs(grads["mixed_refined"] - grads["fp64"])))}
    print(json.dumps(out), flush=True)


def _fixed_tiles(m):
    """The tile count of the fixed thresholds (the library's default without a tuned table)."""
    return 8 if m >= 7000 else (4 if m >= 2500 else 2)


def autotune_symv(api, geo, params, legs):
    """Tunes the SYMV tile count at the workload's size for FP64 and, when a mixed leg runs, FP32
    (installed in the library for the legs that follow); returns the tables, the tuning time and
    the SYMV time per tridiagonalisation that the table saves over the fixed thresholds on the
    tuning run."""
    n = int(sum(params.species[s].norb for s in geo.species))
    out = {"bucket_rows": 256}
    t0 = time.perf_counter()
    for name, prec in (("fp64", 0), ("fp32", 1)):
        if prec == 1 and all(p == "fp64" for p in legs):
            continue
        rows, secs = api.autotune(n, prec)
        opts = (2, 4, 8, 16)
        fixed = sum(secs[opts.index(_fixed_tiles(256 * b)), b] for b in range(secs.shape[1]))
        best = float(secs.min(axis=0).sum())
        out[name] = {"table": rows, "symv_s_tuned": best, "symv_s_fixed_thresholds": float(fixed)}
        print(f"[bench] autotune {name}: {rows} (SYMV {best:.3f} s vs {fixed:.3f} s with the "
              f"fixed thresholds)", file=sys.stderr, flush=True)
    out["seconds"] = time.perf_counter() - t0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="diamond8")
    ap.add_argument("--legs", default="fp64,mixed_tail",
                    help="comma-separated subset of " + ",".join(LEGS))
    ap.add_argument("--mix", type=float, default=0.05,
                    help="Broyden mixing parameter (default 0.05: at 0.2 the charge of the "
                         "4096-atom cell sloshes and the cycle does not converge in 100 "
                         "iterations; 0 takes the workload's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", default="diamond5",
                    help="workload of the bounded CPU-baseline sample")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU baseline on the benchmark workload itself")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the fixed SYMV tile thresholds instead of autotuning them")
    ap.add_argument("--mode", default="scc",
                    choices=["scc", "sytrd", "batch", "kpoints", "autograd", "neighbours",
                             "kernels", "cpu", "tail", "cycle", "tracesum"],
                    help="scc: the headline SCC benchmark (default); sytrd: the tridiagonalisation "
                         "alone on a random matrix (SYMV counter runs); batch: the molecule batch "
                         "against a sampled loop of single-molecule runs; kpoints: SCC iterations "
                         "and a force evaluation with a k-point mesh; autograd: energy and forces "
                         "through torch.autograd with finite-difference checks; neighbours: the "
                         "neighbour list on the device and the geometry set-up with it and "
                         "with the host-built list; kernels: one SCC iteration, forces and a "
                         "GEMM of size --n with the sizes that fix their bytes and flops (for "
                         "kernel traces and counter runs); cpu: the CPU baseline alone, at the "
                         "full size and on the bounded sample; tail: the mixed-precision "
                         "tail from saved charges (--charges-in; without it: save them at the "
                         "tail threshold to --charges-out), for kernel traces; "
                         "cycle: one converged cycle per leg of --legs, no reference leg "
                         "(kernel traces over a window of the cycle); tracesum: the tail's "
                         "per-iteration kernel table from a trace database (--trace)")
    ap.add_argument("--trace", default="", help="--mode tracesum: rocpd trace database")
    ap.add_argument("--charges-in", default="",
                    help="--mode tail: start the tail from these input charges (.npy)")
    ap.add_argument("--charges-out", default="tail_charges.npy",
                    help="--mode tail without --charges-in: where to save the charges")
    ap.add_argument("--n", type=int, default=8192,
                    help="matrix size of --mode sytrd and of the GEMM of --mode kernels")
    ap.add_argument("--nmol", type=int, default=4096, help="molecules of --mode batch")
    ap.add_argument("--fd-coords", type=int, default=4,
                    help="coordinates checked by finite differences in --mode autograd")
    ap.add_argument("--kmesh", default="gc", choices=["gc", "mp"],
                    help="2x2x2 mesh of --mode kpoints: Gamma-centred (gc) or shifted (mp)")
    ap.add_argument("--md-steps", type=int, default=0,
                    help="MD-like steps of --mode kpoints (0: timed SCC iterations only)")
    ap.add_argument("--tiles", type=int, default=0,
                    help="SYMV row tiles per workgroup for --mode sytrd and kpoints (0: by "
                         "trailing size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the SCC path of a single Gamma-point system runs on one "
              "device and no multi-GPU sharding is built; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from dftbplus_amd import api, synth
    # torch is not loaded here: its bundled HIP runtime would be a second runtime in the process
    if args.mode != "scc":
        {"sytrd": mode_sytrd, "batch": mode_batch, "kpoints": mode_kpoints,
         "neighbours": mode_neighbours, "kernels": mode_kernels,
         "autograd": mode_autograd, "cpu": mode_cpu, "tail": mode_tail,
         "cycle": mode_cycle, "tracesum": mode_tracesum}[args.mode](args)
        return
    import dataclasses
    geo, params, opts = synth.workload(args.workload)
    if args.mix:
        opts = dataclasses.replace(opts, mix_param=args.mix, max_scc_iter=300)
    legs = [x for x in args.legs.split(",") if x]
    if "fp64" not in legs:
        legs.insert(0, "fp64")   # the reference leg of the error comparison
    tuned = None if args.no_autotune else autotune_symv(api, geo, params, legs)
    res = {p: run_cycle_leg(api, geo, params, opts, p, args.steps, args.warmup) for p in legs}
    nat = geo.natom
    ref = res["fp64"]
    for p, r in res.items():
        r["error"] = {"dE_per_atom": abs(r["energy"] - ref["energy"]) / nat,
                      "dq_max": float(np.max(np.abs(r["charges"] - ref["charges"]))),
                      "dF_max": float(np.max(np.abs(r["grad"] - ref["grad"])))}
        r["meets_bound"] = (r["converged"] and r["error"]["dE_per_atom"] < E_BOUND
                            and r["error"]["dF_max"] < F_BOUND)
    r = ref
    n = r["n"]
    sv = r["symv"]
    achieved = sv["achieved"]
    ok = [p for p in res if res[p]["meets_bound"] and res[p]["force_s"]]
    best = min(ok, key=lambda p: res[p]["force_s"]) if ok else None
    rec = {
        "metric": METRIC,
        "value": args.steps / r["elapsed"],
        "unit": "SCC iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic diamond-like C/H/O supercell, {geo.natom} atoms, "
                               f"{n} orbitals, Gamma point, Fermi kT=0.01 Ha ({args.workload})",
                   "leg": "fp64", "orbitals": n, "atoms": geo.natom,
                   "broyden_mix": opts.mix_param, "scc_tol": opts.scc_tol},
        "force_evals_per_s": (1.0 / r["force_s"]) if r["force_s"] else None,
        "force_eval_s": r["force_s"],
        "scc_iterations_to_converge": r["iterations"],
        "best_leg": best,
        "best_force_evals_per_s": (1.0 / res[best]["force_s"]) if best else None,
        "phase_s_per_step": r["phases"],
        "roofline": {"kernel": "symv_kernel (tridiagonalisation, lower-triangle symmetric product, "
                               "FP64)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     # the counter runs could not be taken at this size (the profiler's counter
                     # mode ends in a host-side segfault from n = 4096 up); their n = 2048
                     # figures follow
                     "traffic": None,
                     "traffic_n2048": SYMV_TRAFFIC_N2048,
                     "launches": sv["launches"],
                     "avg_launch_us": 1e6 * sv["seconds"] / sv["launches"] if sv["launches"] else None,
                     "alg_bytes_per_launch": sv["bytes"] / sv["launches"] if sv["launches"] else None},
        "legs": {p: {"timed_scc_iter_per_s": args.steps / x["elapsed"],
                     "timed_ms_per_step": 1e3 * x["elapsed"] / args.steps,
                     "converged": x["converged"], "cycle_iterations": x["iterations"],
                     "cycle_scc_s": x["scc_s"], "gradients_s": x["grad_s"],
                     "force_eval_s": x["force_s"],
                     "force_evals_per_s": (1.0 / x["force_s"]) if x["force_s"] else None,
                     "cycle_speedup_vs_fp64": (ref["force_s"] / x["force_s"])
                     if (x["force_s"] and ref["force_s"]) else None,
                     "error_vs_fp64_at_convergence": x["error"], "meets_bound": x["meets_bound"],
                     "phase_s_per_step": x["phases"],
                     "symv_achieved_GBs": x["symv"]["achieved"]} for p, x in res.items()},
    }
    rec["autotune"] = tuned
    if not args.no_cpu_baseline:
        model, ncpu = host_cpu()
        wl = args.workload if args.cpu_full else args.cpu_sample
        ns, setup, first, second = cpu_iteration(wl, timing_gamma=args.cpu_full)
        scale = (n / ns) ** 3
        rec["cpu_baseline"] = {
            "value": 1.0 / (second * scale), "unit": "SCC iterations/s", "cores": blas_threads(),
            "kind": "port", "cpu_model": model, "host_cpus": ncpu,
            "sample": (f"numpy/LAPACK oracle (scipy eigh = LAPACK dsygvd on OpenBLAS), second SCC "
                       f"iteration of {wl} ({ns} orbitals): {second:.2f} s"
                       + (f", scaled by (N/{ns})^3 to N={n}" if ns != n else "")
                       + f"; first iteration {first:.2f} s, set-up {setup:.1f} s")}
        # the full-size measurement (bench.py --mode cpu, two oracle iterations of the
        # 15769-orbital cell on the same host, about two minutes) where it is on record
        full = os.path.join(ROOT, "profiles", f"cpu_baseline_{args.workload}.json")
        if ns != n and os.path.exists(full):
            with open(full) as fh:
                fr = json.load(fh).get(args.workload)
            if fr and fr.get("orbitals") == n:
                rec["cpu_baseline"]["full_size_measured"] = {
                    "value": 1.0 / fr["second_s"], "unit": "SCC iterations/s",
                    "iteration_s": fr["second_s"], "source": os.path.relpath(full, ROOT)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
