"""Train the learned-guess network on the GPU from densities this engine converges.

    python -m pm3amd.train_guess --out pm3amd/data/guess_hco.bin --report report.json

Training data: seeded molecules of the batched workload distribution (alkanes CnH2n+2, n in
[1, 24], and 2-16-water clusters, atoms jittered by 0.05 A) plus jittered conformers
(0.1 A) of C12H26 solvated by 20 waters, drawn from seeds disjoint from the benchmark and test
seeds (mixed: 1_000_000 + chunk, conformers: 2_000_000 + system; held-out: 3_000_000 + k).  Each
chunk is solved by the batched SCF (RHF, p_tol = e_tol = 1e-8) with density output, and its
converged densities go straight into the device-resident training set.  Training is Adam on the
per-atom mean squared block error, all on the device; progress is printed as it goes.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Batch, Molecule, Pm3Options
from . import guess as GS
from . import workloads as W

MIXED_SEED0 = 1_000_000
CONF_SEED0 = 2_000_000
HELD_SEED0 = 3_000_000


def _log(msg):
    print(msg, flush=True)


def _molecules(pairs):
    return [Molecule(np.asarray(z, np.int32), W.to_bohr(x)) for z, x in pairs]


def add_solved(data, mols, opts):
    b = Batch(mols, options=opts, want_density=True)
    b.solve()
    n = data.add(b)
    b.close()
    return n


def build(data, n_mixed, n_conf_systems, conf_per_system, opts, chunk=4096, seed_mixed=MIXED_SEED0,
          seed_conf=CONF_SEED0):
    t0 = time.perf_counter()
    added = 0
    for c in range((n_mixed + chunk - 1) // chunk):
        cnt = min(chunk, n_mixed - c * chunk)
        added += add_solved(data, _molecules(W.mixed_batch(cnt, seed_mixed + c)), opts)
        _log(f"data: mixed chunk {c} -> {added} molecules ({time.perf_counter() - t0:.1f} s)")
    for s in range(n_conf_systems):
        added += add_solved(data, _molecules(W.solvated_conformers(conf_per_system, seed_conf + s)), opts)
        if s % 4 == 3 or s == n_conf_systems - 1:
            _log(f"data: conformer system {s} -> {added} molecules ({time.perf_counter() - t0:.1f} s)")
    return added, time.perf_counter() - t0


def iteration_gain(model, count, seed, opts):
    mols = _molecules(W.mixed_batch(count, seed))
    out = {}
    for name, o in (("sad", Pm3Options(p_tol=opts.p_tol, e_tol=opts.e_tol)),
                    ("learned", Pm3Options(p_tol=opts.p_tol, e_tol=opts.e_tol, guess="learned", guess_model=model))):
        b = Batch(mols, options=o)
        b.solve()
        t0 = time.perf_counter()
        b.solve()
        dt = time.perf_counter() - t0
        st, res = b.results()
        out[name] = {"mean_iterations": float(np.mean([r.iterations for r in res])), "converged": int(np.sum(st == 0)),
                     "solves_per_s": float(count / dt), "hof": [r.heat_of_formation_kcal for r in res],
                     "kept": float(np.mean([r.guess_used == 1 for r in res]))}
        b.close()
    out["max_abs_dhf_kcal"] = float(np.max(np.abs(np.array(out["sad"]["hof"]) - np.array(out["learned"]["hof"]))))
    for k in ("sad", "learned"):
        out[k].pop("hof")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default="")
    ap.add_argument("--mixed", type=int, default=180_224, help="molecules of the mixed distribution")
    ap.add_argument("--conf-systems", type=int, default=40)
    ap.add_argument("--conf-per-system", type=int, default=512)
    ap.add_argument("--held-mixed", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--segment", type=int, default=2000, help="steps between held-out evaluations")
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5, help="end of the cosine step-size schedule")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default="", help="start from this model file instead of random weights")
    ap.add_argument("--gain-molecules", type=int, default=2048)
    args = ap.parse_args(argv)

    opts = Pm3Options(p_tol=1e-8, e_tol=1e-8)
    data = GS.GuessData()
    n_train, t_data = build(data, args.mixed, args.conf_systems, args.conf_per_system, opts)
    held = GS.GuessData()
    n_held, _ = build(held, args.held_mixed, 1, 512, opts, seed_mixed=HELD_SEED0, seed_conf=HELD_SEED0 + 500)
    _log(f"data: train {data.stats()} held-out {held.stats()}")

    model = GS.GuessModel.load(args.init) if args.init else GS.GuessModel.create(args.seed)
    curve = []
    t_train = 0.0
    done = 0
    while done < args.steps:
        k = min(args.segment, args.steps - done)
        losses, sec = GS.train(model, data, k, args.batch, args.lr, seed=args.seed * 1000 + done,
                               lr_final=args.lr_final, decay_steps=args.steps)
        t_train += sec
        done += k
        h = GS.evaluate(model, held)
        curve.append({"step": done, "train_loss_mean": float(np.mean(losses)), "held_out_loss": h})
        _log(f"train: step {done} train {np.mean(losses):.5f} held-out {h:.5f} ({t_train:.1f} s)")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    gain = iteration_gain(model, args.gain_molecules, HELD_SEED0 + 7, opts)
    _log(f"gain: {json.dumps(gain)}")
    report = {"train": data.stats(), "held_out": held.stats(), "train_molecules": n_train, "data_seconds": t_data,
              "train_seconds": t_train, "steps": done, "batch_molecules": args.batch, "lr": args.lr,
              "lr_final": args.lr_final,
              "curve": curve, "gain": gain}
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)
    print(json.dumps({k: report[k] for k in ("train", "held_out", "train_seconds", "steps")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
