"""Throughput benchmark of the MI355X PM3 engine: batched SCF with the learned initial guess versus
the plain superposition guess.

A step is one batched solve of the whole workload: every molecule of the batch is taken from its
initial guess to a converged RHF PM3 density (density convergence 1e-8, DIIS on) -- for the
learned leg the pair features and the bf16 network run first, then the persistent batched SCF
kernel -- with the inputs already resident in HBM.  ``value`` is converged, validated solves per
second of the learned leg over the timed steps; ``plain`` holds the same for the superposition
guess, timed the same way.  A learned solve counts only if it converged and its heat of
formation is within 1e-5 kcal/mol of the plain GPU solve of the same molecule (and the CPU sample
within 1e-5 of the CPU restatement).

With no ``--workload`` the line's headline is mixed4096 (learned vs plain) and ``configs`` holds the
other configurations, each measured after it with the built-in tuning table: water32 (one 32-water
cluster), water512 (SCF + gradient of one 512-water cluster) and conformers2048 (SCF + gradient,
learned guess), each with its roofline and CPU baseline.

Workloads (``--workload``):
  mixed4096      4096 molecules: alkanes CnH2n+2 (n in [1,24]) and water clusters (2-16 waters),
                 every atom jittered by sigma 0.05 A (default; the batched distribution the
                 solves/sec metric is quoted on)
  water32        one 32-water cluster (96 atoms, 192 AOs)
  conformers2048 2048 jittered conformers (sigma 0.1 A) of C12H26 + 20 waters (98 atoms, 194 AOs)
  water512       one 512-water cluster in a 25 A cube (1536 atoms, 3072 AOs): SCF plus analytic
                 gradient on the multi-kernel large-molecule path; a step is one solve
  optimize512    512 alkanes CnH2n+2 (n = 1..24 in turn), every atom jittered by 0.08 A, L-BFGS
                 geometry optimisation of all of them (pm3_optimize_batch); a step is the whole batch
  dc             divide-and-conquer SCF (run_dc, default DcOptions) of 64-, 128-, 256- and 512-water
                 clusters; a step is one solve of each, value the 512-water cluster's solves/s

The CPU baseline is the C++ restatement of the reference CPU path (oracle/) on a bounded sample of
the same workload, one solve per host thread (SCF then gradient with ``--gradient``).  The same
sample checks |dHf| <= 1e-5 kcal/mol and, with ``--gradient``, max |dg| <= 2.295e-7 eV/Bohr.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "pm3-rs_amd"))

import pm3amd as G  # noqa: E402
from pm3amd import workloads as W  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
NET_FLOP_PER_PAIR = 2.0 * (96 * 128 + 128 * 128 + 128 * 16)  # forward pass per directed pair
# HBM bytes per SCF launch (the 512-, 256- and 128-thread kernels of one solve, summed) measured
# with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a 512 MB copy: x2.0 / x1.0) on the default
# workload, per leg.  A recorded value, not measured by this run: the JSON line carries it with its
# source (the profile summary and the build it was taken on).
MEASURED_TRAFFIC = {("mixed4096", "learned"): 8.98e10, ("mixed4096", "plain"): 1.237e11,
                    ("water512", "tile"): 9.97e8}  # HBM bytes per SCF launch / per purification step
TRAFFIC_SOURCE = ("recorded, not measured in this run: profiles/scf_batch_kernel_traffic_mixed4096.md "
                  "(rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of bench.py --steps 1)")
HBM_PEAK_GBS = 8000.0
HOF_TOL = 1.0e-5
GRAD_TOL = 2.295e-7  # eV/Bohr, max |dg| against the CPU path
# Final heats of formation of two optimisations of one molecule: both stop at max |g| < 1e-3
# eV/Bohr, and on soft modes (methyl torsions, intermolecular modes) two stopping points can lie
# a few 1e-3 kcal/mol apart once the trajectories have parted.  The reference's own optimiser test
# accepts an optimised heat of formation within 0.005 kcal/mol (src/optimizer.rs:384).
OPT_HOF_TOL = 5.0e-3


def cpu_info(threads):
    """The host the CPU baseline ran on: CPU model, logical cores, threads used."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads": threads}


def build_workload(name, seed):
    if name == "mixed4096":
        mols = W.mixed_batch(4096, seed)
    elif name == "water32":
        mols = [W.water_cluster(32, 10.0, seed)]
    elif name == "conformers2048":
        mols = W.solvated_conformers(2048, seed)
    else:
        raise SystemExit(f"unknown workload {name}")
    return [(np.asarray(z, dtype=np.int32), W.to_bohr(x)) for z, x in mols]


def algorithmic_flops(naos, results):
    """Per molecule and SCF iteration: (10/3) n^3 eigensolve + 2 n^3 for F P + n^2 n_occ for the
    density (the fixed per-unit convention); summed over the iterations each molecule ran."""
    tot = 0.0
    for n, r in zip(naos, results):
        n = float(n)
        tot += r.iterations * ((10.0 / 3.0) * n ** 3 + 2.0 * n ** 3 + n * n * r.n_occ)
    return tot


def large_cpu_baseline(z, x, iterations, threads):
    """The CPU path of one water512 solve, priced per SCF iteration: the restatement's Fock build at
    full size (orc_fock minus its H_core, which the reference builds once) plus its dense eigensolve,
    timed at n = 768 and scaled by (3072 / 768)^3 (a full-size eigensolve takes minutes on one
    thread), times the GPU run's iteration count."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    n = O.nao(z)
    P = np.zeros((n, n))
    t0 = time.perf_counter()
    O.hcore(z, x)
    t1 = time.perf_counter()
    O.fock(z, x, P)
    t2 = time.perf_counter()
    m = 768
    A = np.random.default_rng(0).standard_normal((m, m))
    A = A + A.T
    t3 = time.perf_counter()
    O.eigen(A)
    t4 = time.perf_counter()
    fock_s = max((t2 - t1) - (t1 - t0), 0.0)
    eig_s = (t4 - t3) * (n / m) ** 3
    per_iter = fock_s + eig_s
    return {"value": 1.0 / (per_iter * iterations), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one SCF iteration priced from the C++ restatement: Fock build at {n} AOs "
                      f"({fock_s:.2f} s) + dense eigensolve timed at {m} and scaled by ({n}/{m})^3 "
                      f"({eig_s:.1f} s), times {iterations} iterations; one thread (the restatement "
                      f"does not split a solve across threads); gradient not included",
            **cpu_info(1)}


def run_large(args, steps, warmup, tune, cpu=True):
    """One large molecule (SCF + analytic gradient per step) on the host-driven multi-kernel path."""
    z, x = W.water_cluster(512, 25.0, args.seed)
    mol = G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))
    opts = G.Pm3Options(p_tol=1.0e-8, e_tol=1.0e-8)
    tuning = apply_tuning(args, [mol], opts, True, reps=1, tune=tune)
    batch = G.Batch([mol], options=opts, want_gradient=True, want_density=True)
    for _ in range(warmup):
        batch.solve()
    kms, lts = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.solve()
        kms.append(batch.last_kernel_ms())
        lts.append(batch.large_timing())
    t1 = time.perf_counter()
    st, res = batch.results()
    r = res[0]
    g = batch.gradients()[0]
    ms_per_step = (t1 - t0) * 1e3 / steps
    P = r.density
    n = P.shape[0]
    idem = float(np.abs((P / 2) @ (P / 2) - P / 2).max())
    # dominant kernel: the purification steps on the FP64 matrix cores (the lower-triangle tile
    # kernel with its update fused, or rocBLAS dgemm products when the tuning table says so); beside
    # it the Fock build against HBM with its algorithmic bytes 8 (sum_pairs p_a p_b + 3 n^2), p = 10 / 1
    prod_ms = sum(t["product_ms"] for t in lts)
    prod_flop = sum(t["product_flop"] for t in lts)
    nprod = sum(t["products"] for t in lts)
    achieved = prod_flop / (prod_ms * 1e-3) / 1e12 if prod_ms > 0 else 0.0
    split = G.tuning_get()["large_split"]
    kname = ("lg_sp2_sym_kernel (purification step: X X on the lower-triangle 64 x 64 tiles of an n = %d "
             "square, FP64 MFMA, 2 X - X X and the trace sums fused into its store)" % n if split == 1 else
             "purification product (rocBLAS dgemm: the %s of an n = %d square)"
             % ("2 x 2 block-lower part" if split == 2 and n % 2 == 0 else
                "block-lower part" if split > 2 else "whole", n))
    zz = np.asarray(z)
    nh, nH = int(np.sum(zz > 2)), int(np.sum(zz <= 2))
    fock_bytes = 8.0 * (100 * nh * (nh - 1) / 2 + 10 * nh * nH + nH * (nH - 1) / 2 + 3 * n * n)
    fock_ms = sum(t["fock_ms"] for t in lts)
    nfock = sum(t["fock_builds"] for t in lts)
    fock_gbs = fock_bytes * nfock / (fock_ms * 1e-3) / 1e9 if fock_ms > 0 else 0.0
    out = {
        "metric": "converged PM3 SCF + analytic gradient solves/sec (one 512-water cluster, 3072 AOs)",
        "value": (1.0 if st[0] == 0 else 0.0) / (ms_per_step * 1e-3),
        "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded geometry)",
        "config": {"workload": "water512", "molecules": 1, "atoms": len(z), "nao": n, "seed": args.seed,
                   "p_tol": 1e-8, "guess": "sad", "gradient": True,
                   "parallelism": "grid-wide kernels per SCF step; purified densities by %s; frontier levels by seeded Lanczos (rocBLAS gemv)"
                   % ("the lower-triangle tile kernel" if split == 1 else "rocBLAS dgemm")},
        "iterations": r.iterations, "converged": int(st[0] == 0), "heat_of_formation_kcal": r.heat_of_formation_kcal,
        "max_abs_gradient_ev_per_bohr": float(np.abs(g).max()),
        "checks": {"idempotency_max_abs": idem, "trace_minus_electrons": float(np.trace(P) - 2 * r.n_occ),
                   "gradient_sum_max_abs": float(np.abs(g.sum(axis=0)).max())},
        "device_ms_per_step": float(np.mean(kms)),
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": MEASURED_TRAFFIC.get(("water512", "tile")) if split == 1 and n == 3072 else None,
                     "traffic_source": ("recorded, not measured in this run: profiles/large_path_water512.md (rocprofv3 "
                                        "--pmc FETCH_SIZE and WRITE_SIZE runs of bench.py --workload water512)")
                     if split == 1 and n == 3072 else "not measured for this kernel",
                     "launch_ms": prod_ms / max(nprod, 1), "launches_per_step": nprod / steps,
                     "flop_per_launch": prod_flop / max(nprod, 1),
                     "flop_note": "achieved counts the FLOP issued (2 x 64 x 64 x the slab-rounded AO count per "
                                  "tile of the lower triangle); achieved_n3 counts n^3 per step, the product's "
                                  "lower triangle",
                     "achieved_n3": float(n) ** 3 / (prod_ms / max(nprod, 1) * 1e-3) / 1e12 if prod_ms > 0 else 0.0},
        "fock_roofline": {"bound": "hbm", "kernel": "lg_fock_pairs + lg_fock_jsum + lg_fock_diag",
                          "achieved": fock_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fock_gbs / HBM_PEAK_GBS,
                          "bytes_per_build": fock_bytes, "build_ms": fock_ms / max(nfock, 1),
                          "builds_per_step": nfock / steps},
        "cpu_baseline": large_cpu_baseline(np.asarray(z, np.int32), W.to_bohr(x), r.iterations, 1) if cpu else None,
        "tuning": tuning,
    }
    return out


def bench_dc(args):
    """Divide-and-conquer SCF (run_dc, src/dc/scf.rs:74-160) of seeded water clusters at the
    32-waters-in-10-A density, default DcOptions, on the device: one step is one run_dc per cluster
    size (64, 128, 256 and 512 waters); value is the 512-water cluster's solves/s.  The CPU baseline
    is the oracle's restatement of run_dc on the 64-water cluster, one thread."""
    sizes = [(64, 12.6), (128, 15.9), (256, 20.0), (512, 25.0)]
    opts = G.Pm3Options(max_scf=400)
    out_sizes = {}
    for k, box in sizes:
        z, x = W.water_cluster(k, box, args.seed)
        m = G.Molecule(np.asarray(z, np.int32), W.to_bohr(x))
        for _ in range(args.warmup):
            G.run_dc(m, options=opts)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = G.run_dc(m, options=opts)
        t1 = time.perf_counter()
        out_sizes[f"water{k}"] = {"nao": 6 * k, "ms_per_solve": (t1 - t0) * 1e3 / args.steps,
                                  "iterations": r.iterations, "subsystems": r.n_subsystems,
                                  "largest_subsystem": r.largest_subsystem, "heat_of_formation_kcal":
                                  r.heat_of_formation_kcal, "converged": bool(r.converged)}
    big = out_sizes["water512"]
    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        z, x = W.water_cluster(64, 12.6, args.seed)
        c0 = time.perf_counter()
        st, o = O.run_dc(np.asarray(z, np.int32), W.to_bohr(x), 0.0, 1, O.options(max_scf=400))
        c1 = time.perf_counter()
        small = out_sizes["water64"]
        cpu = {"value": 1.0 / (c1 - c0), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": "the 64-water cluster (384 AOs) by the oracle's restatement of run_dc, one thread; "
                         f"the GPU takes {small['ms_per_solve']:.0f} ms for the same solve",
               "gpu_speedup_on_sample": (c1 - c0) * 1e3 / small["ms_per_solve"],
               "max_abs_dhf_vs_gpu_kcal": abs(o["hof_kcal"] - small["heat_of_formation_kcal"]) if st == 0 else None,
               **cpu_info(1)}
    return {
        "metric": "divide-and-conquer PM3 SCF solves/sec (one 512-water cluster, 3072 AOs, default DcOptions)",
        "value": (1.0 if big["converged"] else 0.0) / (big["ms_per_solve"] * 1e-3), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": big["ms_per_solve"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded geometries)",
        "config": {"workload": "dc", "clusters": [f"water{k}" for k, _ in sizes], "seed": args.seed,
                   "parallelism": "subsystem eigensolves batched (rocSOLVER dsyevd_strided_batched), one global "
                                  "chemical potential by device bisection, Yang-Lee assembly and density DIIS kernels"},
        "sizes": out_sizes,
        "roofline": None,
        "roofline_note": "the time goes to the batched subsystem eigensolves (rocSOLVER, one call per size "
                         "bucket); the kernel trace is in profiles/dc_water.md",
        "cpu_baseline": cpu,
    }


def bench_optimize(args):
    """L-BFGS geometry optimisations (pm3_optimize_batch): a step optimises the whole batch from
    the same starting geometries; every round of a step is one batched SCF + gradient solve."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(args.seed)
    zx = []
    for k in range(512):
        z, x = W.alkane(1 + k % 24)
        zx.append((np.asarray(z, np.int32), W.to_bohr(W.jitter(np.asarray(x), 0.08, rng))))
    mols = [G.Molecule(z, x) for z, x in zx]
    # SCF to p_tol 1e-10 by default, so that final geometries are checkable against the CPU path
    # (--p-tol 1e-7 is the reference's Pm3Options::default(), whose SCF noise parts trajectories
    # along soft torsions by more than the check's 1e-3 kcal/mol)
    p_tol = args.p_tol or 1.0e-10
    e_tol = 1.0e-12 if p_tol < 1.0e-8 else 1.0e-8
    opts = G.Pm3Options(p_tol=p_tol, e_tol=e_tol)
    opt = G.OptOptions()
    tuning = apply_tuning(args, mols, opts, True, reps=1)
    for _ in range(args.warmup):
        G.optimize_batch(mols, options=opts, opt=opt, trajectory=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, rs = G.optimize_batch(mols, options=opts, opt=opt, trajectory=False)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # the same batch with every round's SCF started from the molecule's previous density
    wopt = G.OptOptions(warm_start=True)
    G.optimize_batch(mols, options=opts, opt=wopt, trajectory=False)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        wst, wrs = G.optimize_batch(mols, options=opts, opt=wopt, trajectory=False)
    w1 = time.perf_counter()
    w_ms = (w1 - w0) * 1e3 / args.steps
    wok = np.array([s == 0 and r.converged for s, r in zip(wst, wrs)])
    ok = np.array([s == 0 and r.converged for s, r in zip(st, rs)])
    cpu, checked, max_dhf, max_dx, same = None, 0, None, None, 0
    valid = int(ok.sum())
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        n = min(args.cpu_sample, 64, len(mols))
        stride = max(1, len(mols) // n)
        idx = list(range(0, len(mols), stride))[:n]
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        oo = O.options(p_tol=p_tol, e_tol=e_tol)
        c0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            cres = list(ex.map(lambda i: O.optimize(zx[i][0], zx[i][1], 0.0, oo), idx))
        c1 = time.perf_counter()
        dh = [abs(rs[i].heat_of_formation_kcal - c["hof_kcal"]) for i, c in zip(idx, cres) if ok[i] and c["converged"]]
        wdh = [abs(wrs[i].heat_of_formation_kcal - c["hof_kcal"]) for i, c in zip(idx, cres) if wok[i] and c["converged"]]
        dx = [float(np.abs(rs[i].molecule.positions - c["xyz"]).max()) for i, c in zip(idx, cres)
              if ok[i] and c["converged"]]
        checked = len(dh)
        max_dhf = max(dh) if dh else None
        max_dx = max(dx) if dx else None
        same = sum(1 for i, c in zip(idx, cres) if ok[i] and c["converged"] and rs[i].iterations == c["iterations"])
        if max_dhf is not None and max_dhf > OPT_HOF_TOL:
            valid = 0
        cpu = {"value": len(idx) / (c1 - c0), "unit": "optimisations/s", "cores": threads, "kind": "port",
               "sample": f"every {stride}th molecule of the same workload ({len(idx)} molecules): the oracle's "
                         f"restatement of the reference optimiser (energy-only trials, gradient at accepted "
                         f"points), one optimisation per host thread"}
    out = {
        "metric": "converged L-BFGS geometry optimisations/sec (512 jittered alkanes C1-C24)",
        "value": valid / (ms_per_step * 1e-3), "unit": "optimisations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded geometries, jittered by 0.08 A)",
        "config": {"workload": "optimize512", "molecules": len(mols), "seed": args.seed, "p_tol": p_tol,
                   "e_tol": e_tol, "gtol": opt.gtol, "history": opt.history,
                   "parallelism": "one batched SCF + gradient launch per L-BFGS round over all running molecules"},
        "converged": int(ok.sum()),
        "mean_iterations": float(np.mean([r.iterations for r in rs])),
        "max_iterations": int(max(r.iterations for r in rs)),
        "mean_evaluations": float(np.mean([r.evaluations for r in rs])),
        "max_evaluations": int(max(r.evaluations for r in rs)),
        "hof_checked": checked, "max_abs_dhf_kcal": max_dhf, "max_abs_dxyz_bohr": max_dx,
        "same_iterations_as_cpu": same if args.cpu_sample > 0 else None, "hof_tolerance_kcal": OPT_HOF_TOL,
        "roofline": None,
        "roofline_note": "each round is one scf_batch_kernel launch over the running molecules; that kernel's "
                         "roofline is measured on mixed4096",
        "warm_start": {"value": (int(wok.sum()) if args.cpu_sample <= 0 or (wdh and max(wdh) <= OPT_HOF_TOL) else 0)
                       / (w_ms * 1e-3), "ms_per_step": w_ms, "converged": int(wok.sum()),
                       "mean_iterations": float(np.mean([r.iterations for r in wrs])),
                       "max_evaluations": int(max(r.evaluations for r in wrs)),
                       "max_abs_dhf_kcal": (max(wdh) if args.cpu_sample > 0 and wdh else None),
                       "note": "PM3_BATCH_WARM_START: each round's SCF starts from the molecule's previous "
                               "density (the reference starts every SCF from its guess)"},
        "cpu_baseline": cpu,
        "tuning": tuning,
    }
    print(json.dumps(out))


def apply_tuning(args, molecules, options, want_gradient, reps=2, tune=None):
    """Set the process-wide tuning table as --tune asks (or `tune`); returns its record."""
    tune = tune or args.tune
    if tune == "off":
        G.tuning_set(G.tuning_default())
        return {"source": "built-in (the table pm3_autotune keeps on this workload, profiles/autotune_tables.md)",
                "table": G.tuning_get()}
    if tune != "auto":
        G.tuning_load(tune)
        return {"source": tune, "table": G.tuning_get()}
    t0 = time.perf_counter()
    table, rep = G.autotune(molecules, options=options, want_gradient=want_gradient, reps=reps)
    return {"source": "pm3_autotune on this workload (untimed, before the legs)", "table": table, "report": rep,
            "seconds": time.perf_counter() - t0}


def run_batched(args, workload, gradient, steps, warmup, cpu_sample, tune, learned=True):
    """The batched workloads: plain and learned legs over the same device-resident batch."""
    # The gradient at the converged density inherits the density's convergence error, so the
    # gradient runs converge to 1e-10 (energy 1e-12), the tolerance of the gradient parity tests.
    p_tol = args.p_tol or (1.0e-10 if gradient else 1.0e-8)
    e_tol = 1.0e-12 if p_tol < 1.0e-8 else 1.0e-8
    mols = build_workload(workload, args.seed)
    molecules = [G.Molecule(z, x) for z, x in mols]
    model = None
    if not args.plain_only and learned:
        import pm3amd.guess as GS
        path = args.guess_model or GS.DEFAULT_MODEL
        if os.path.exists(path):
            model = GS.GuessModel.load(path)

    tune_opts = (G.Pm3Options(p_tol=p_tol, e_tol=e_tol, guess="learned", guess_model=model) if model is not None
                 else G.Pm3Options(p_tol=p_tol, e_tol=e_tol))
    tuning = apply_tuning(args, molecules, tune_opts, gradient, tune=tune)

    def leg(opts):
        batch = G.Batch(molecules, options=opts, want_gradient=gradient)
        for _ in range(warmup):
            batch.solve()
        kms, pms, gms, nms = [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.solve()
            kms.append(batch.last_kernel_ms())
            pms.append(batch.last_pair_ms())
            gms.append(batch.last_guess_ms())
            nms.append(batch.guess_timing()[1])
        t1 = time.perf_counter()
        st, rs = batch.results()
        leg.network_ms = float(np.mean(nms))
        leg.pairs = batch.guess_timing()[2]
        leg.large = batch.large_timing()["fock_builds"] > 0  # molecules on the multi-kernel path
        return batch, st, rs, (t1 - t0) * 1e3 / steps, float(np.mean(kms)), float(np.mean(pms)), float(np.mean(gms))

    pbatch, pstatus, pres, p_ms, p_kms, p_pms, _ = leg(G.Pm3Options(p_tol=p_tol, e_tol=e_tol))
    if model is not None:
        batch, status, res, ms_per_step, k_ms, pair_ms, guess_ms = leg(
            G.Pm3Options(p_tol=p_tol, e_tol=e_tol, guess="learned", guess_model=model))
    else:
        batch, status, res, ms_per_step, k_ms, pair_ms, guess_ms = pbatch, pstatus, pres, p_ms, p_kms, p_pms, 0.0
    converged = int(np.sum(status == 0))
    # every learned solve against the plain GPU solve of the same molecule
    ok_vs_plain = (status == 0) & (pstatus == 0) & np.array(
        [abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) <= HOF_TOL for a, b in zip(res, pres)])
    max_dhf_plain = float(max(abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) for a, b in zip(res, pres)))

    naos = [int(np.sum(np.where(z <= 2, 1, 4))) for z, _ in mols]
    flops = algorithmic_flops(naos, res)
    achieved = flops / (k_ms * 1e-3) / 1e12

    cpu = None
    grad_check = None
    max_dhf = None
    checked = 0
    hof_ok = True
    if cpu_sample > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        stride = max(1, len(mols) // cpu_sample)
        idx = list(range(0, len(mols), stride))[:cpu_sample]
        sample = [mols[i] for i in idx]
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        c0 = time.perf_counter()
        if gradient:
            # the CPU path does the step's whole work: SCF, then closed_form_gradient
            _, hof_cpu, _, st_cpu, g_cpu = O.gradient_batch(sample, O.options(p_tol=p_tol, e_tol=e_tol),
                                                            nthreads=threads)
        else:
            _, hof_cpu, _, st_cpu = O.run_batch(sample, O.options(p_tol=p_tol, e_tol=e_tol), nthreads=threads)
        c1 = time.perf_counter()
        if gradient:
            gpu_g = batch.gradients()
            both = [k for k, i in enumerate(idx) if st_cpu[k] == 0 and status[i] == 0]
            gmax = max((float(np.abs(gpu_g[idx[k]] - g_cpu[k]).max()) for k in both), default=None)
            grad_check = {"molecules": len(both), "max_abs_dgrad_ev_per_bohr": gmax}
        dh = np.array([res[i].heat_of_formation_kcal - hof_cpu[k] for k, i in enumerate(idx)
                       if st_cpu[k] == 0 and status[i] == 0] +
                      [pres[i].heat_of_formation_kcal - hof_cpu[k] for k, i in enumerate(idx)
                       if st_cpu[k] == 0 and pstatus[i] == 0])
        checked = int(dh.size)
        max_dhf = float(np.abs(dh).max()) if dh.size else None
        hof_ok = bool(dh.size == 0 or np.abs(dh).max() <= HOF_TOL)
        if grad_check is not None and grad_check["max_abs_dgrad_ev_per_bohr"] is not None:
            grad_check["within_tolerance"] = grad_check["max_abs_dgrad_ev_per_bohr"] <= GRAD_TOL
            hof_ok = hof_ok and grad_check["within_tolerance"]
        # the other variant: one solve at a time (the restatement runs one solve on one thread, so
        # this is the single-solve rate; the faster of the two is the baseline)
        nseq = min(len(sample), max(1, threads // 2))
        s0 = time.perf_counter()
        if gradient:
            O.gradient_batch(sample[:nseq], O.options(p_tol=p_tol, e_tol=e_tol), nthreads=1)
        else:
            O.run_batch(sample[:nseq], O.options(p_tol=p_tol, e_tol=e_tol), nthreads=1)
        s1 = time.perf_counter()
        per_thread = len(sample) / (c1 - c0)
        one_at_a_time = nseq / (s1 - s0)
        what = "SCF + gradient" if gradient else "solve"
        cpu = {"value": max(per_thread, one_at_a_time), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"every {stride}th molecule of the same workload ({len(sample)} molecules), C++ "
                         f"restatement of the reference CPU path, one {what} per thread on {threads} threads "
                         f"({per_thread:.3g}/s); one {what} at a time on one thread over the first {nseq} "
                         f"({one_at_a_time:.3g}/s); value is the faster",
               "one_per_thread": per_thread, "one_at_a_time": one_at_a_time, **cpu_info(threads)}

    valid = int(np.sum(ok_vs_plain)) if hof_ok else 0
    value = valid / (ms_per_step * 1e-3)
    plain_valid = int(np.sum(pstatus == 0)) if hof_ok else 0
    plain_value = plain_valid / (p_ms * 1e-3)
    used = np.array([r.guess_used for r in res])
    out = {
        "metric": ("converged PM3 SCF + analytic gradient solves/sec (conformer-gradients/sec), learned-guess vs "
                   "plain, at |dHf|<=1e-5 kcal/mol and max|dg|<=2.295e-7 eV/Bohr" if gradient else
                   "converged PM3 SCF solves/sec, learned-guess vs plain, at |dHf|<=1e-5 kcal/mol"),
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometries)",
        "config": {"workload": workload, "molecules": len(mols), "seed": args.seed, "p_tol": p_tol, "e_tol": e_tol,
                   "guess": "learned" if model is not None else "sad",
                   "parallelism": ("one molecule over the whole GPU (multi-kernel path: fused purification-step, "
                                   "Fock, DIIS and Lanczos kernels; superposition guess)" if leg.large else
                                   "one workgroup per molecule, persistent launch")},
        "converged": converged,
        "valid_vs_plain": int(np.sum(ok_vs_plain)),
        "max_abs_dhf_vs_plain_kcal": max_dhf_plain,
        "hof_checked": checked,
        "max_abs_dhf_kcal": max_dhf,
        "mean_iterations": float(np.mean([r.iterations for r in res])),
        "plain": {"value": plain_value, "ms_per_step": p_ms, "kernel_ms": p_kms, "pair_kernel_ms": p_pms,
                  "mean_iterations": float(np.mean([r.iterations for r in pres])), "converged": plain_valid},
        "speedup_vs_plain": value / plain_value if plain_value > 0 else None,
        "learned": None if model is None else {
            "guess_kernels_ms": guess_ms, "network_kernel_ms": leg.network_ms, "directed_pairs": leg.pairs,
            "network_roofline": {"bound": "mfma", "kernel": "mlp_fwd_kernel", "unit": "TFLOP/s", "peak": BF16_PEAK_TFLOPS,
                                 "achieved": NET_FLOP_PER_PAIR * leg.pairs / (leg.network_ms * 1e-3) / 1e12,
                                 "frac": NET_FLOP_PER_PAIR * leg.pairs / (leg.network_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS},
            "kept_fraction": float(np.mean(used == 1)),
            "fallback_rate": float(np.mean(used == 2)), "not_applicable": float(np.mean(used == 0)),
            "model_steps": model.steps},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": MEASURED_TRAFFIC.get((workload, "learned" if model is not None else "plain"))
                     if args.seed == 2026 else None,
                     "traffic_source": TRAFFIC_SOURCE if (workload, "learned") in MEASURED_TRAFFIC and args.seed == 2026
                     else "not measured for this workload",
                     "kernel": ("the multi-kernel path's solve (its kernels and host round trips between "
                                "two events)" if leg.large else "scf_batch_kernel"),
                     "kernel_ms": k_ms, "algorithmic_flop": flops, "pair_kernel_ms": pair_ms},
        "cpu_baseline": cpu,
        "tuning": tuning,
    }
    if gradient:
        out["gradient_check"] = grad_check
        out["config"]["gradient"] = True
    if os.environ.get("PM3_PHASE_TIMERS") == "1":
        ticks = batch.phase_ticks()
        counts = ticks.pop("counts")
        tot = sum(ticks.values()) or 1
        out["phase_share"] = {k: round(v / tot, 4) for k, v in ticks.items()}
        out["phase_counts"] = counts
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed4096")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-sample", type=int, default=512, help="molecules in the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--gradient", action="store_true", help="also compute analytic gradients in every step")
    ap.add_argument("--p-tol", type=float, default=0.0,
                    help="density convergence (default 1e-8; 1e-10 with --gradient, where the gradient check needs it)")
    ap.add_argument("--copy-calibration-mb", type=float, default=0.0,
                    help="profiling helper: run only a plain device copy of this size (for counter scaling)")
    ap.add_argument("--guess-model", default="", help="learned-guess network file (default: the packaged one)")
    ap.add_argument("--plain-only", action="store_true", help="time only the superposition-guess leg")
    ap.add_argument("--tune", default="auto",
                    help="launch-shape / tile table: 'auto' runs pm3_autotune on the workload first (untimed), "
                         "'off' keeps the built-in table, anything else is a table file to load")
    ap.add_argument("--configs", default="all",
                    help="with the default workload: the other configurations measured after the headline, as "
                         "sub-records of the same line ('all', 'none', or a comma list of water32, water512, "
                         "conformers2048)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding a batch over several GPUs is "
                                   "designed but not launched", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "water512":
        print(json.dumps(run_large(args, args.steps, args.warmup, args.tune)))
        return
    if args.workload == "optimize512":
        return bench_optimize(args)
    if args.workload == "dc":
        print(json.dumps(bench_dc(args)))
        return

    if args.copy_calibration_mb > 0:
        import torch
        n = int(args.copy_calibration_mb * 1e6 / 8)
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 8 * n}))
        return

    out = run_batched(args, args.workload, args.gradient, args.steps, args.warmup, args.cpu_sample, args.tune)
    if args.workload == "mixed4096" and args.configs != "none" and not args.gradient:
        # The other configurations, each timed on its own after the headline with the built-in tuning
        # table (what the tuner keeps on them); sub-records, not part of the headline value.
        want = {"water32", "water512", "conformers2048"} if args.configs == "all" else set(args.configs.split(","))
        sub = {"C2_C3_mixed4096": "this line's headline (value: learned leg; plain: the plain leg)"}
        if "water32" in want:
            sub["C0_water32"] = run_batched(args, "water32", False, 10, 2, 1, "off")
        if "water512" in want:
            sub["C1_water512"] = run_large(args, 1, 1, "off")
        if "conformers2048" in want:
            sub["C4_conformers2048"] = run_batched(args, "conformers2048", True, 2, 1, 64, "off")
        out["configs"] = sub
    print(json.dumps(out))


if __name__ == "__main__":
    main()
