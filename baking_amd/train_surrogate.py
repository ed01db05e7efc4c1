"""Train the return-map surrogate on rows the exact return map emits during a workload's run.

    python -m baking_amd.train_surrogate [--workload loaf_256|batched_48] [--rows 4000000] [--steps 2000]
                                         [--stride 50] [--epochs 20]

Runs the workload -- the 256^3 loaf (0.72M particles, dt 5e-5, 2000 substeps)
or the 48 batched doughs (1.0M particles, dt 1.25e-5, 1000 substeps) -- with
the exact coupled return map, emitting (inputs, exact outputs) rows for every
yielding particle on every ``--stride``-th substep, up to ``--rows`` rows plus a
held-out tail.  The network's inputs carry no time step, and the return map
depends on it, so each workload's network is trained on its own rows.  Collection stops at the first substep that raises a device
flag, so every row comes from a state the simulator accepts.  Rows are appended in substep order, so the tail comes from
the latest emitting substeps and serves as a hold-out later in the bake than
anything trained on.  Trains with the library's FP32 kernels and Adam on the
GPU and writes the model (``--out``) and its held-out error table
(``--report``, JSON).
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

from . import scenes, surrogate


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--rows", type=int, default=4_000_000, help="training rows (cfg: 4M)")
    ap.add_argument("--holdout-rows", type=int, default=300_000, help="rows of the latest emitting substeps kept out")
    ap.add_argument("--workload", choices=("loaf_256", "batched_48"), default="loaf_256")
    ap.add_argument("--steps", type=int, default=None, help="substeps to run (default: the workload's own)")
    ap.add_argument("--stride", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="model file (default: the workload's packaged model)")
    ap.add_argument("--report", default=None, help="JSON file for the held-out error table")
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    if args.workload == "loaf_256":
        spec = scenes.loaf_256()
        sim = scenes.build_simulator(spec)
        steps, first = args.steps or spec.steps, sim
    else:
        from .physics.batch import BatchSimulator
        specs = scenes.batched_doughs()
        sim = BatchSimulator([scenes.build_simulator(s) for s in specs])
        spec, steps, first = specs[0], args.steps or 1000, sim.doughs[0]
    out = args.out or str(surrogate.default_model(args.workload))
    rows = surrogate.collect_samples(sim, steps, capacity=args.rows + args.holdout_rows, stride=args.stride,
                                     log=lambda m: print(m, flush=True))
    t1 = time.perf_counter()
    n_hold = min(args.holdout_rows, rows.shape[0] // 10)
    train, later = rows[:rows.shape[0] - n_hold], rows[rows.shape[0] - n_hold:]
    print(f"collected {rows.shape[0]} rows in {t1 - t0:.1f} s; training on {train.shape[0]}, later hold-out {n_hold}", flush=True)
    meta = dict(workload=args.workload, signature=surrogate.material_signature(first), steps=steps, stride=args.stride)
    model, shuffled = surrogate.ReturnMapSurrogate.fit(train, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
                                                       log=lambda m: print(m, flush=True), meta=meta)
    later_table = model.evaluate(later)
    model.meta["holdout_later_substeps"] = later_table
    t2 = time.perf_counter()
    model.meta["train_seconds"] = t2 - t1
    model.save(out)
    report = dict(rows_collected=int(rows.shape[0]), training_rows=int(train.shape[0]), train_seconds=t2 - t1,
                  collect_seconds=t1 - t0, train_mse=model.meta["train_mse"], holdout_shuffled=shuffled,
                  holdout_later_substeps=later_table, meta={k: v for k, v in meta.items()})
    print(json.dumps(report, indent=1))
    if args.report:
        Path(args.report).parent.mkdir(parents=True, exist_ok=True)
        Path(args.report).write_text(json.dumps(report, indent=1))


if __name__ == "__main__":
    main()
