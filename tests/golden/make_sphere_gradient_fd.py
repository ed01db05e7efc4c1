"""Write ``sphere_gradient_fd.npz``: FP64 central finite differences of the differentiable workload's loss.

The workload (``baking_amd.scenes.sphere_gradient``): the small sphere scene for 256 substeps; loss
L = (y_c - y*)^2 + (s_mean - s*)^2 of the final state (mass-weighted centroid height, mean saturation).
Parameters: each particle's initial porosity theta (its gas moles p0 theta / (R T0) with it) and the oven
temperature.  For 32 seeded particles and for the oven temperature, the FP64 oracle runs 256 substeps from
theta_i +/- 1e-3 (the oven temperature +/- 1e-3 K) and records (L+ - L-) / 2e-3.

Run from the repository root: ``python tests/golden/make_sphere_gradient_fd.py`` (66 oracle runs of 256
substeps on all host cores, a few minutes).
"""
import sys
from dataclasses import replace
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

EPS = 1e-3
COORDS = 32


def loss_of(task):
    from baking_amd import scenes
    from tests.harness import O, oracle_from_spec
    kind, index, delta, precision = task
    spec = scenes.sphere_gradient()
    c = spec.ding
    n = len(spec.samples.particle_positions)
    theta = np.full(n, c.porosity)
    if kind == "oven":
        spec.transport = replace(spec.transport, oven_temperature_k=spec.transport.oven_temperature_k + delta)
    elif kind == "theta":
        theta[index] += delta
    spec.porosities = theta
    spec.gas_moles = c.gas_initial_pressure_pa * theta / (O.GAS_CONSTANT * spec.transport.initial_temperature_k)
    o = oracle_from_spec(spec, precision)
    for _ in range(spec.steps):
        o.Step()
    return float(scenes.gradient_loss(spec, o.state.x, o.state.transport["W"], o.state.F))


def coordinates(n):
    return np.sort(np.random.default_rng(4).choice(n, COORDS, replace=False))


def main():
    from baking_amd import scenes
    spec = scenes.sphere_gradient()
    idx = coordinates(len(spec.samples.particle_positions))
    tasks = [("base", -1, 0.0, "f64"), ("base", -1, 0.0, "f32"), ("oven", -1, EPS, "f64"), ("oven", -1, -EPS, "f64")]
    tasks += [("theta", int(i), d, "f64") for i in idx for d in (EPS, -EPS)]
    with Pool(8) as pool:
        losses = pool.map(loss_of, tasks, chunksize=1)
    base64, base32, op, om = losses[:4]
    th = np.asarray(losses[4:]).reshape(COORDS, 2)
    out = dict(indices=idx.astype(np.int32), eps=np.float64(EPS), loss_f64=np.float64(base64), loss_f32=np.float64(base32),
               fd_oven=np.float64((op - om) / (2 * EPS)), fd_theta=(th[:, 0] - th[:, 1]) / (2 * EPS),
               centroid_target=np.float64(spec.notes["centroid_target"]),
               saturation_target=np.float64(spec.notes["saturation_target"]))
    path = Path(__file__).with_name("sphere_gradient_fd.npz")
    np.savez(path, **out)
    print("loss FP64", base64, "FP32", base32, "dL/dT_oven", out["fd_oven"], "dL/dtheta", out["fd_theta"])
    print("wrote", path)


if __name__ == "__main__":
    main()
