"""Write ``sphere_small_1000.npz``: the oracle's 1000-substep run of the small sphere scene.

The scene: sphere r = 0.1 at (0.5, 0.5, 0.3), 64^3 grid, ~8.8k jittered particles (seed 0), Ding dough with
pore gas and heat/water transport, oven 453.15 K, dt = 1e-4 (``baking_amd.scenes.sphere_small``).  The FP32
oracle (the reference's precision) and the FP64 oracle (every float promoted) each run 1000 substeps; every
10th substep the acceptance aggregates of ``OracleSimulator.Measure`` (reference
experiments/recording.py:46-112) are recorded, and after the last substep the fields of 512 seeded particles.

Run from the repository root: ``python tests/golden/make_sphere_small_1000.py`` (about 90 s).  The fixture is
data the oracle wrote; ``tests/test_oracle_acceptance.py`` re-derives its first substeps to pin it.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from baking_amd import scenes  # noqa: E402
from tests.harness import oracle_fields, oracle_from_spec  # noqa: E402

STEPS = 1000
EVERY = 10
SAMPLE = 512
KEYS = ("particle_volume_m3", "temperature_mean_k", "water_mass_kg", "gas_total_mol", "water_lost_kg", "heat_exchange_j",
        "thermal_energy_j")
FIELDS = ("x", "v", "F", "C", "Fe", "Fp", "tau", "porosity", "moles", "pressure", "T", "W", "s", "losses", "exchanges",
          "masses", "yield_stress", "viscosity", "density", "gelled")


def sample_indices(n):
    return np.sort(np.random.default_rng(1000).choice(n, SAMPLE, replace=False))


def run(precision, steps=STEPS, every=EVERY, log=None):
    spec = scenes.sphere_small()
    o = oracle_from_spec(spec, precision)
    rows = [[o.Measure()[k] for k in KEYS]]
    for step in range(1, steps + 1):
        o.Step()
        if step % every == 0:
            rows.append([o.Measure()[k] for k in KEYS])
            if log and step % 100 == 0:
                log(f"{precision} substep {step}")
    return o, np.asarray(rows, dtype=np.float64)


def main():
    out = {"keys": np.array(KEYS), "steps": np.int64(STEPS), "every": np.int64(EVERY)}
    for precision in ("f32", "f64"):
        o, rows = run(precision, log=lambda m: print(m, flush=True))
        out[f"aggregates_{precision}"] = rows
        idx = sample_indices(o.n)
        out["sample"] = idx.astype(np.int32)
        fields = oracle_fields(o)
        for name in FIELDS:
            out[f"{precision}_{name}"] = np.asarray(fields[name])[idx]
    path = Path(__file__).with_name("sphere_small_1000.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, path.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
