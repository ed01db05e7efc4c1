#!/usr/bin/env python3
"""Benchmark: particle-substeps/s of the fused MLS-MPM baking substep on one MI355X.

Workloads (synthetic seeded inputs, Ding + gas + heat/water transport, FP32
state with the reference's FP64 transport; one "step" is one substep of every
particle):

* ``batched_48`` (default, the headline): see below.
* ``loaf_256``: the 0.30 x 0.15 x 0.12 loaf on a 256^3 grid, ~0.72M
  particles, one simulator, with the packaged return-map network attached.
* ``rest_loaf``: ``loaf_256`` resting before baking under the automatic
  settling controller.
* ``sphere_small``: the small sphere scene, one simulator.
* ``adjoint_loaf``: gradients (``baking_amd.adjoint.rollout``) through 64
  substeps of ``loaf_256`` (Ding + pore gas + heat/water transport): forward
  with checkpoints every 16 substeps, then the backward (segment recompute,
  forward rerun and adjoint kernels per substep).  Prints its own JSON line
  (gradient particle-substeps/s), not the headline.
* ``batched_48``: 48 independent doughs of 3k-60k particles (~1.0M) on their
  own 64^3 / 128^3 grids, advanced as one batch (one launch per stage over
  every dough, one HIP graph per substep, autotuned launch shapes).  Also
  times the plain comparator: the same doughs one at a time through the
  single-dough substep, launched kernel by kernel without a graph.

Prints ONE JSON line with the throughput, a roofline entry for the dominant
kernel (HIP events on the launch stream, algorithmic bytes at the
reference's precisions) and a CPU baseline (the serial C++ oracle on a
bounded sample of the same workload, on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
FP64_PEAK_TFLOPS = 78.6   # MI355X vector/matrix FP64, published
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32, published
# VALU issue: 256 CUs x 4 SIMDs, one 64-lane wavefront instruction per SIMD every 4 cycles at 2.4 GHz
VALU_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4

# Algorithmic bytes of each stage at the reference's precisions: (per particle, per active stencil node).
# Particle terms count each array the stage must read once and write once; node terms the grid arrays it
# must read or write once per touched node.  Sort/tiles/clear move no algorithmic bytes (pure overhead).
STAGE_BYTES = {
    # x, F, V0, capacity, T, W (f64), exposed+contact (f64) -> 7 f64 scatters per node
    "pass_a": (12 + 36 + 4 + 8 + 8 + 8 + 16, 56),
    # reads V, C, E, W, heat, loss, exchange (f64) ; writes dT, dW (f64)
    "diffuse": (0, 56 + 16),
    # reads x, F, V0, dry, T, W, losses, exchanges, moles, tau_solid, porosity; writes T, W, losses, exchanges,
    # masses, density, T_mat, moles; gathers dT, dW, W_node
    "pass_b1": (12 + 36 + 4 + 8 + 32 + 8 + 36 + 4 + 32 + 4 + 4 + 4 + 8, 24),
    # reads x, v, C, tau_total, mass, V0; scatters m, p
    "pass_b2": (12 + 12 + 36 + 36 + 4 + 4, 16),
    # reads m, p ; writes v
    "grid": (0, 16 + 12),
    # reads x, F, Fe_old, porosity, gelled, T_mat, moles, W, V0 ; writes x, v, C, F, Fe, Fp, tau, residual,
    # porosity, porosity residual, T_mat, gelled, yield, viscosity, pressure, saturation; gathers v
    "pass_c": (12 + 36 + 36 + 4 + 4 + 4 + 8 + 8 + 4 + 96 + 108 + 8 + 16 + 8 + 8, 12),
    # settling controller of a resting dough: reads v, C, the reference Fp and the new Fp
    "settling": (12 + 36 + 36 + 36, 0),
}
# Deferred coupled-porosity return maps, per queued particle: reads Fe_old, porosity, gelled, F, C, T_mat,
# moles, W, V0; writes Fe, Fp, tau, residual, porosity residual, porosity, T_mat, gelled, yield, viscosity,
# pressure, saturation.
DEFERRED_BYTES = 36 + 4 + 4 + 36 + 36 + 4 + 8 + 8 + 4 + 108 + 8 + 4 + 4 + 4 + 4 + 4 + 8 + 8
SUBSTEP_BYTES = (552, 208)  # whole substep: particle state read + written once, grid work per active node
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 matrix peak, published
MLP_MACS = 7 * 64 + 3 * 64 * 64 + 64 * 4  # multiply-accumulates of the surrogate network per row


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed substeps (default: 1000 batched_48, 2000 loaf_256, 1000 sphere_small)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="batched_48",
                    choices=["batched_48", "loaf_256", "sphere_small", "adjoint_loaf", "rest_loaf"])
    ap.add_argument("--no-surrogate", action="store_true", help="exact return map only")
    ap.add_argument("--model", default=None, help="surrogate model file (default: the workload's packaged one)")
    ap.add_argument("--plain-steps", type=int, default=20, help="batched_48: substeps per dough of the plain comparator")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-baseline-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--graph-steps", type=int, default=None, help="substeps per graph launch in timed runs (default: the simulator's; the autotuner starts there)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def build_scene(name):
    from baking_amd import scenes
    return scenes.loaf_256() if name == "loaf_256" else scenes.sphere_small()


def host_cpu():
    """The host this runs on: CPU model (lscpu), nproc, and the threads this process may use (its share of the
    box: at most 16)."""
    import os
    import subprocess
    model = None
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, check=False).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, check=False).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    return {"lscpu_model": model, "nproc": nproc, "threads_used": threads}


def oracle_builds():
    """The CPU baseline's oracle builds, compiled here for this host: serial -O3 -march=native and OpenMP.  Falls back
    to the checker build (serial -O2) when the compiler is unavailable."""
    from oracle import oracle as O
    try:
        return {"serial": O.build_timing(False), "openmp": O.build_timing(True),
                "flags": "g++ -O3 -march=native -ffp-contract=off (OpenMP build: -fopenmp, atomic scatters)"}
    except Exception as error:  # noqa: BLE001 -- report the fallback instead of failing the bench
        return {"serial": None, "openmp": None, "flags": f"checker build g++ -O2 (timing build failed: {error})"}


def _time_oracle(specs, steps, library, threads):
    """particle-substeps/s of the oracle on `specs` (each: 1 untimed substep, then `steps` timed), and seconds."""
    from oracle import oracle as O
    from tests.harness import oracle_from_spec
    if library is not None:
        O.lib(library).oracle_set_threads(threads)
    total_p, total_s = 0, 0.0
    for spec in specs:
        sim = oracle_from_spec(spec, library=library)
        sim.Step()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.Step()
        total_s += time.perf_counter() - t0
        total_p += sim.n
        del sim
    return total_p * steps / total_s, total_s


def cpu_baseline(spec, steps, label=None):
    """The oracle (CPU restatement of the reference kernels) on `steps` substeps of `spec`, timed on this host with
    the serial -O3 build and the OpenMP build on every thread of this process's share."""
    host, builds = host_cpu(), oracle_builds()
    serial, t_serial = _time_oracle([spec], steps, builds["serial"], 1)
    out = {"value": serial, "unit": "particle-substeps/s", "cores": 1, "kind": "port", "build": builds["flags"],
           "host": host, "serial": {"value": serial, "cores": 1, "seconds": t_serial}}
    if builds["openmp"] is not None:
        par, t_par = _time_oracle([spec], steps, builds["openmp"], host["threads_used"])
        out["openmp"] = {"value": par, "cores": host["threads_used"], "seconds": t_par}
        out["value"], out["cores"] = par, host["threads_used"]
    out["sample"] = (label or f"{steps} substeps (after 1 untimed) of the full {spec.name} scene") + \
        f"; value = the {'OpenMP' if builds['openmp'] is not None else 'serial'} build"
    return out


def check_state_timing(spec, sim, with_host, repeats=5):
    """Wall milliseconds of MpmSimulator.CheckState (one bk_check_state pass and a 15-double copy), median of
    `repeats`; with_host also times the reference's host CheckState on the same state (state copies to the host
    and the numpy checks of the oracle), the computation the device pass replaces."""
    import torch
    times = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = sim.CheckState()
        times.append(time.perf_counter() - t0)
    out = {"gpu_ms": float(np.median(times)) * 1e3, "estimated_cfl": stats["estimated_cfl"], "host_ms": None}
    if with_host:
        from tests.harness import load_gpu_state, oracle_from_spec
        orc = oracle_from_spec(spec)
        host = []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            load_gpu_state(orc, sim).check_state()
            host.append(time.perf_counter() - t0)
        out["host_ms"] = float(np.median(host)) * 1e3
    return out


def counter_entry(workload, kernel, particles, active_nodes, deferred):
    """The committed rocprofv3 counter figures of `kernel` (profiles/<workload>_counters.json: HBM bytes from the
    sized read and write requests, VALU instructions and FLOPs, per unit of work) and this launch's units;
    (None, 0) if the kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", f"{workload}_counters.json")
    if not os.path.exists(path):
        return None, 0
    with open(path) as stream:
        entry = json.load(stream)["per_kernel"].get(f"k_{kernel}")
    if entry is None:
        return None, 0
    units = {"particle": particles, "queued particle": deferred, "active stencil node": active_nodes}[entry["unit"]]
    return entry, units


def compute_roofline(entry, units, ms):
    """fp64, fp32 and VALU-issue rates of one launch: the profiled per-unit counts times this launch's units over
    its live duration.  FLOPs count every lane of each issued instruction (SQ_INSTS_VALU_FLOPS_* x 64)."""
    if entry is None or ms <= 0:
        return None
    sec = ms * 1e-3
    fp64 = entry["fp64_flops_per_unit"] * units / sec / 1e12
    fp32 = entry["fp32_flops_per_unit"] * units / sec / 1e12
    valu = entry["valu_instr_per_unit"] * units / sec / 1e9
    return {"fp64": {"achieved": fp64, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fp64 / FP64_PEAK_TFLOPS},
            "fp32": {"achieved": fp32, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fp32 / FP32_PEAK_TFLOPS},
            "valu_issue": {"achieved": valu, "peak": VALU_ISSUE_PEAK_G, "unit": "G wavefront-instructions/s",
                           "frac": valu / VALU_ISSUE_PEAK_G}}


def stage_profile(lib_launch, workspace, steps):
    """Per-stage milliseconds per substep (HIP events around each stage on the launch stream) and the
    workspace counters {tiles, active node blocks, deferred return maps} of the last substep."""
    import ctypes
    from baking_amd import _native
    lib = _native.load()
    _native.check(lib.bk_profile_begin(workspace, steps), "bk_profile_begin")
    for _ in range(steps):
        lib_launch()
    nst = lib.bk_profile_stage_count()
    ms = (ctypes.c_float * nst)()
    rec = ctypes.c_int32(0)
    _native.check(lib.bk_profile_end(workspace, ctypes.cast(ms, ctypes.c_void_p), ctypes.cast(ctypes.pointer(rec), ctypes.c_void_p)),
                  "bk_profile_end")
    stages = {lib.bk_profile_stage_name(k).decode(): ms[k] / max(rec.value, 1) for k in range(nst)}
    counts = (ctypes.c_int32 * 3)()
    _native.check(lib.bk_workspace_counts(workspace, ctypes.cast(counts, ctypes.c_void_p), _native.stream_ptr()),
                  "bk_workspace_counts")
    return stages, [int(c) for c in counts]


def roofline(stages, n, active_nodes, deferred, substep_ms, workload):
    """Roofline entry of the dominant stage: algorithmic bytes per launch over its event-timed duration."""
    dominant = max(stages, key=lambda k: stages[k])
    if dominant == "pass_c_deferred":
        dom_bytes = DEFERRED_BYTES * deferred
    else:
        per_p, per_node = STAGE_BYTES.get(dominant, (0, 0))
        dom_bytes = per_p * n + per_node * active_nodes
    dom_ms = stages[dominant]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    substep_bytes = SUBSTEP_BYTES[0] * n + SUBSTEP_BYTES[1] * active_nodes
    substep_gbs = substep_bytes / (substep_ms * 1e-3) / 1e9
    entry, units = counter_entry(workload, dominant, n, active_nodes, deferred)
    roof = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": entry["hbm_bytes_per_unit"] * units if entry else None,
            "algorithmic_bytes": dom_bytes, "avg_ms": dom_ms, "stages_ms": stages,
            "substep_achieved": substep_gbs, "substep_frac": substep_gbs / HBM_PEAK_GBS,
            "compute": compute_roofline(entry, units, dom_ms)}
    return roof, dominant, dom_bytes, dom_ms


def _timed(batch, steps, stream):
    """Milliseconds of `steps` graph replays of the batch substep, pipelined (each substep's transport pre-pass and
    P2G set-up beside the previous one's deferred return maps) and ending with the last one's deferred pass (HIP
    events on the launch stream)."""
    import torch
    batch._ensure_graphs()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    batch._launch_substeps(steps, stream)
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop)


def _relative_errors(ref, got, start):
    """Max over doughs of |got - ref| / |ref| for the final aggregates, and of the same for their increments."""
    out = {}
    for key in ("particle_volume_m3", "temperature_mean_k", "water_mass_kg", "gas_total_mol"):
        r, g, s0 = ref[key], got[key], start[key]
        out[key] = float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-300)))
        inc = np.abs(r - s0)
        out[key + "_increment"] = float(np.max(np.abs(g - r) / np.maximum(inc, 1e-300))) if np.any(inc > 0) else 0.0
    return out


def run_batched(args):
    """The batched workload: plain one-at-a-time comparator, then exact-batched and surrogate+batched from one state."""
    import torch
    from baking_amd import scenes, surrogate
    from baking_amd.autotune import autotune
    from baking_amd.measure import measure
    from baking_amd.physics.batch import BatchSimulator

    steps = args.steps or 1000
    t_build = time.perf_counter()
    specs = scenes.batched_doughs()
    sims = [scenes.build_simulator(s) for s in specs]
    build_s = time.perf_counter() - t_build
    counts = [s.particle_data.particle_count for s in sims]
    n = int(sum(counts))
    stream = torch.cuda.current_stream()

    # plain comparator: each dough alone, single-dough substep launched kernel by kernel, exact return map
    plain_ms = 0.0
    for s in sims:
        s.use_graphs = False
        for _ in range(2):
            s._launch_substep()
            s._accept()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.plain_steps):
            s._launch_substep()
            s._accept()
        stop.record(stream)
        stop.synchronize()
        plain_ms += start.elapsed_time(stop)
    plain_rate = n * args.plain_steps / (plain_ms * 1e-3)
    # the same doughs alone, each substep one HIP graph replay
    plain_graph_ms = 0.0
    for s in sims:
        s.use_graphs = True
        for _ in range(2):
            s._launch_substep()
            s._accept()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.plain_steps):
            s._launch_substep()
            s._accept()
        stop.record(stream)
        stop.synchronize()
        plain_graph_ms += start.elapsed_time(stop)
        s._destroy_graphs()
    plain_graph_rate = n * args.plain_steps / (plain_graph_ms * 1e-3)

    batch = BatchSimulator(sims)
    if args.graph_steps is not None:
        batch.graph_steps = args.graph_steps
    model = None
    if not args.no_surrogate:
        model = surrogate.ReturnMapSurrogate.load(args.model or surrogate.default_model("batched_48"))
        surrogate.attach(batch, model)
    tuned = None if args.no_autotune else autotune(batch)
    if model is not None:
        surrogate.detach(batch)
    for _ in range(args.warmup):
        batch._launch()
    torch.cuda.synchronize()
    snapshot = batch._save_for_tuning()
    start_aggr = measure(batch)

    # exact return map, batched
    surrogate.stats(batch, reset=True)
    exact_ms = _timed(batch, steps, stream)
    exact_aggr, exact_stats = measure(batch), surrogate.stats(batch, reset=True)
    flags_exact = batch.ReadFlags()
    result_method = "batched (exact return map)"
    value_ms, stats_used, aggr_used = exact_ms, exact_stats, exact_aggr
    accuracy = guided = None
    if model is not None:
        # the network guiding every solve to the exact root, then standing in for it (the measured method)
        batch._restore_after_tuning(snapshot)
        surrogate.attach(batch, model, mode="guided")
        guided_ms = _timed(batch, steps, stream)
        g_aggr, g_stats = measure(batch), surrogate.stats(batch, reset=True)
        guided = {"value": n * steps / (guided_ms * 1e-3), "ms_per_step": guided_ms / steps,
                  "evaluations_per_map": g_stats["evaluations"] / max(g_stats["return_maps"], 1),
                  "accuracy_vs_exact": _relative_errors(exact_aggr, g_aggr, start_aggr)}
        batch._restore_after_tuning(snapshot)
        surrogate.attach(batch, model, mode="standin")
        value_ms = _timed(batch, steps, stream)
        aggr_used, stats_used = measure(batch), surrogate.stats(batch, reset=True)
        accuracy = _relative_errors(exact_aggr, aggr_used, start_aggr)
        result_method = "surrogate+batched"
    flags = batch.ReadFlags()
    del snapshot

    stages, wcounts = stage_profile(lib_launch=lambda: batch._launch(use_graph=False), workspace=batch._workspace,
                                    steps=args.profile_steps)
    active_nodes = int((batch.grid["m"] > 0).sum().item())
    substep_ms = value_ms / steps
    roof, _, _, _ = roofline(stages, n, active_nodes, wcounts[2], substep_ms, "batched_48")
    if "surrogate" in stages and stages["surrogate"] > 0 and wcounts[2] > 0:
        flops = 2.0 * MLP_MACS * wcounts[2]
        tf = flops / (stages["surrogate"] * 1e-3) / 1e12
        roof["surrogate_mfma"] = {"bound": "mfma", "achieved": tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": tf / BF16_PEAK_TFLOPS, "rows": wcounts[2], "avg_ms": stages["surrogate"]}
    residuals = batch._states[batch._cur].get("residual")
    yield_fraction = float((residuals > 0).float().mean().item()) if residuals is not None else 0.0
    value = n * steps / (value_ms * 1e-3)
    exact_rate = n * steps / (exact_ms * 1e-3)
    per_map = lambda st: st["evaluations"] / st["return_maps"] if st["return_maps"] else 0.0  # noqa: E731
    result = {
        "metric": "particle-substeps/s on 1 MI355X, plain vs surrogate+batched at <=1e-3 rel err",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": substep_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_48", "method": result_method, "hip_graph": True, "doughs": len(sims),
                   "particles": n, "particles_per_dough": [min(counts), max(counts)],
                   "grids": sorted(set(s.simulation_config.grid_resolution for s in sims)),
                   "dt": sims[0].simulation_config.time_step,
                   "active_stencil_nodes": active_nodes, "yield_fraction": yield_fraction, "deferred_return_maps": wcounts[2],
                   "tiles": wcounts[0], "active_node_blocks": wcounts[1], "tuning": dict(batch.Tuning(), graph_steps=batch.graph_steps),
                   "autotune": None if tuned is None else {"start": tuned["start"], "best_ms": tuned["ms"],
                                                           "trials": len(tuned["trials"])},
                   "return_map_stats": {"exact": exact_stats, "used": stats_used,
                                        "evaluations_per_map": {"exact": per_map(exact_stats), "used": per_map(stats_used)}},
                   "wall_build_s": build_s,
                   "flags_raised": {"exact": [[int(d), int(k)] for d, k in zip(*np.nonzero(flags_exact))],
                                    "used": [[int(d), int(k)] for d, k in zip(*np.nonzero(flags))]}},
        "plain": {"value": plain_rate, "unit": "particle-substeps/s", "steps_per_dough": args.plain_steps,
                  "method": "each dough alone, single-dough substep launched kernel by kernel (no graph), exact return map",
                  "ms_total": plain_ms,
                  "graph": {"value": plain_graph_rate, "ms_total": plain_graph_ms,
                            "method": "each dough alone, single-dough substep as one HIP graph replay, exact return map"}},
        "speedup_vs_plain": value / plain_rate,
        "speedup_vs_plain_graph": value / plain_graph_rate,
        "batched_exact": {"value": exact_rate, "ms_per_step": exact_ms / steps},
        "surrogate_gain_vs_batched_exact": value / exact_rate,
        "surrogate_guided": guided,
        "accuracy_vs_exact": accuracy,
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_batched(specs, args.cpu_baseline_steps)
    print(json.dumps(result))


def cpu_baseline_batched(specs, steps, every=4):
    """The oracle on every `every`-th dough of the batch (a bounded sample of the workload), serial -O3 and OpenMP
    builds compiled for this host; plus the small sphere scene (8.8k particles) for its full 1000 substeps."""
    from baking_amd import scenes
    host, builds = host_cpu(), oracle_builds()
    sample = specs[::every]
    serial, t_serial = _time_oracle(sample, steps, builds["serial"], 1)
    n = sum(len(s.samples.particle_positions) for s in sample)
    out = {"value": serial, "unit": "particle-substeps/s", "cores": 1, "kind": "port", "build": builds["flags"], "host": host,
           "serial": {"value": serial, "cores": 1, "seconds": t_serial}}
    if builds["openmp"] is not None:
        par, t_par = _time_oracle(sample, steps, builds["openmp"], host["threads_used"])
        out["openmp"] = {"value": par, "cores": host["threads_used"], "seconds": t_par}
        out["value"], out["cores"] = par, host["threads_used"]
    out["sample"] = (f"{steps} substeps (after 1 untimed) of every {every}th dough of the batch ({len(sample)} doughs, {n} "
                     f"particles, each on its own dense grid); value = the {'OpenMP' if builds['openmp'] is not None else 'serial'} build")
    sphere = scenes.sphere_small()
    full = {"substeps": 1000, "particles": len(sphere.samples.particle_positions)}
    v, t = _time_oracle([sphere], 1000, builds["serial"], 1)
    full["serial"] = {"value": v, "seconds": t, "cores": 1}
    if builds["openmp"] is not None:
        v, t = _time_oracle([sphere], 1000, builds["openmp"], host["threads_used"])
        full["openmp"] = {"value": v, "seconds": t, "cores": host["threads_used"]}
    out["sphere_small_1000"] = full
    return out


def copy_probe():
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 4, "repeats": 5}))


# Adjoint substep (k_adj_g2p, k_adj_grid, k_adj_p2g), algorithmic bytes per particle: the forward state it
# reads (in: x, F, Fe, porosity, gelled, T_mat, C, v, tau, moles, mass, V0; out: F, C, T, moles) and the FP64
# adjoint state read and written (x, v, C, F, Fe, tau, porosity, moles); per active node: the seven FP64 grid
# adjoints written and read once and the forward grid (m, p, v) read.
ADJOINT_PARTICLE_BYTES = (12 + 36 + 36 + 4 + 4 + 4 + 36 + 12 + 36 + 8 + 4 + 4) + (36 + 36 + 16 + 8) + 2 * (24 + 24 + 72 + 72 + 72 + 72 + 8 + 8)
ADJOINT_NODE_BYTES = 2 * 56 + 28


def run_adjoint(args):
    """Gradient throughput: rollout forward (checkpointed) + backward on the loaf."""
    import ctypes
    import torch
    from baking_amd import _native, adjoint
    from baking_amd.scenes import build_simulator, loaf_256
    spec = loaf_256()
    sim = build_simulator(spec)
    n = sim.particle_data.particle_count
    steps = args.steps or 64
    for _ in range(args.warmup):
        sim._launch_substep()
        sim._accept()
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    w = torch.as_tensor(rng.normal(size=(n, 3)), device="cuda")
    # one untimed gradient first: the allocator's pools and the reverse binding index are then in place
    warm = adjoint.rollout(sim, 16, {"v": sim.current_state.particle_velocities.detach().clone().requires_grad_(True)})
    (warm["x"].double() * w).sum().backward()
    del warm
    torch.cuda.synchronize()
    v0 = sim.current_state.particle_velocities.detach().clone().requires_grad_(True)
    eta = sim.material.viscosity_parameter.detach().clone().requires_grad_(True)
    h = sim.transport.heat_transfer_parameter.detach().clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = adjoint.rollout(sim, steps, {"v": v0}, viscosity_parameter=eta, heat_transfer_parameter=h)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = (out["x"].double() * w).sum() + out["T"].mean()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    # the adjoint kernels alone: repeated calls on one substep's (in, out, grid)
    lib = _native.load()
    sim._launch_substep()  # the forward substep whose grid and P2G tiles the adjoint uses
    torch.cuda.synchronize()
    g = sim.grid_state
    active_nodes = int((g.tiled_masses > 0).sum().item())
    adj = {f: torch.as_tensor(rng.normal(size=n * adjoint._WIDTH[f]) * 1e-3, device="cuda") for f in adjoint._fields_of(sim)}
    bk_adj = _native.BkAdjoint()
    for f, t in adj.items():
        setattr(bk_adj, f, t.data_ptr())
    for f, t in adjoint._binding_reverse(sim).items():
        setattr(bk_adj, f, t.data_ptr())
    scratch = torch.zeros(int(lib.bk_adjoint_scratch_size(ctypes.byref(sim._params))), dtype=torch.float64, device="cuda")
    param = torch.zeros(4, dtype=torch.float64, device="cuda")  # viscosity, gas production, heat transfer, oven temperature
    st_in, st_out = sim._state_struct(sim.current_state), sim._state_struct(sim.next_state)
    stream = torch.cuda.current_stream()

    def call():
        _native.check(lib.bk_adjoint_substep(ctypes.byref(sim._params), ctypes.byref(sim._scene), ctypes.byref(st_in),
                                             ctypes.byref(st_out), ctypes.byref(sim._grid), sim._workspace, ctypes.byref(bk_adj),
                                             _native.ptr(scratch), _native.ptr(param), _native.stream_ptr()), "bk_adjoint_substep")
    call()
    reps = 20
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(reps):
        call()
    stop.record(stream)
    torch.cuda.synchronize()
    adj_ms = start.elapsed_time(stop) / reps
    adj_bytes = n * ADJOINT_PARTICLE_BYTES + active_nodes * ADJOINT_NODE_BYTES
    result = {
        "metric": "gradient particle-substeps/s on 1 MI355X (forward with checkpoints + backward)",
        "value": n * steps / (fwd + bwd),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": (fwd + bwd) / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 adjoint",
        "data": "synthetic",
        "config": {"workload": spec.name, "particles": n, "grid": spec.simulation.grid_resolution, "dt": spec.simulation.time_step,
                   "checkpoint_every": 16, "active_stencil_nodes": active_nodes,
                   "gradient_of": "sum(w * x_final) + mean(T_final), w ~ N(0,1), w.r.t. v0, the viscosity and the "
                                  "heat-transfer parameters"},
        "forward_ms_per_substep": fwd / steps * 1e3,
        "backward_ms_per_substep": bwd / steps * 1e3,
        "adjoint_kernels": {"ms_per_substep": adj_ms, "algorithmic_bytes": adj_bytes,
                            "achieved_GBs": adj_bytes / (adj_ms * 1e-3) / 1e9, "peak_GBs": HBM_PEAK_GBS,
                            "frac": adj_bytes / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "viscosity_gradient": float(eta.grad),
        "heat_transfer_gradient": float(h.grad),
        "v0_gradient_norm": float(v0.grad.norm()),
    }
    print(json.dumps(result))


def run_rest(args):
    """The loaf resting before baking (automatic settling with examples/oven_baking.txt:66-71): substep
    throughput with the device-side settling controller, and the controller's own cost."""
    import torch
    from dataclasses import replace
    from baking_amd.scenes import build_simulator, loaf_256
    spec = loaf_256()
    spec.transport = replace(spec.transport, settling_duration_s=5.0, settling_auto=True, settling_min_duration_s=0.1,
                             settling_hold_duration_s=0.1, settling_speed_tolerance=0.001, settling_rate_tolerance=0.01)
    sim = build_simulator(spec)
    n = sim.particle_data.particle_count
    # the rest cannot end before min_duration_s (2000 substeps at dt 5e-5): every timed substep rests
    steps = args.steps or 1000
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        sim._launch_substep()
        sim._accept()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        sim._launch_substep()
        sim._accept()
    stop.record(stream)
    torch.cuda.synchronize()
    gpu_ms = start.elapsed_time(stop)
    sim._pull_settling()
    resting = sim.IsSettling
    stages, counts = stage_profile(lib_launch=lambda: (sim._launch_substep(use_graph=False), sim._accept()),
                                   workspace=sim._workspace, steps=args.profile_steps)
    sim._pull_settling()
    settle_ms = stages.get("settling", 0.0)
    settle_bytes = STAGE_BYTES["settling"][0] * n
    result = {
        "metric": "resting particle-substeps/s on 1 MI355X (settling controller on the device)",
        "value": n * steps / (gpu_ms * 1e-3),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "rest_loaf", "particles": n, "grid": spec.simulation.grid_resolution,
                   "dt": spec.simulation.time_step, "settling": "auto, examples/oven_baking.txt:66-71",
                   "still_resting": resting, "quiet_steps": sim.settling_quiet_steps,
                   "diagnostics": sim.settling_diagnostics, "tiles": int(counts[0])},
        "settling_controller": {"kernels": ["k_settling_update", "k_settling_finish"], "ms_per_substep": settle_ms,
                                "algorithmic_bytes": settle_bytes,
                                "achieved_GBs": settle_bytes / (settle_ms * 1e-3) / 1e9 if settle_ms > 0 else 0.0,
                                "peak_GBs": HBM_PEAK_GBS,
                                "frac": settle_bytes / (settle_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if settle_ms > 0 else 0.0,
                                "share_of_substep": settle_ms / (gpu_ms / steps)},
        "stages_ms": stages,
    }
    print(json.dumps(result))


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every configuration here fits one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from baking_amd import _native
    from baking_amd.scenes import build_simulator

    if args.workload == "batched_48":
        return run_batched(args)
    if args.workload == "adjoint_loaf":
        return run_adjoint(args)
    if args.workload == "rest_loaf":
        return run_rest(args)
    from baking_amd import surrogate
    from baking_amd.measure import measure
    spec = build_scene(args.workload)
    # loaf_256: the scene's 2000 substeps
    steps = args.steps or (spec.steps if spec.name == "loaf_256" else 1000)
    sim = build_simulator(spec)
    sim.use_graphs = not args.no_graph
    if args.graph_steps is not None:
        sim.graph_steps = args.graph_steps
    n = sim.particle_data.particle_count
    stream = torch.cuda.current_stream()
    model = None
    if not args.no_surrogate and sim.current_state.material_state is not None and sim.material.ding_config.gas_enabled:
        model = surrogate.ReturnMapSurrogate.load(args.model or surrogate.DEFAULT_MODEL)
        if surrogate.attach(sim, model, strict=False) == 0:
            model = None
    tuned = None
    if not args.no_autotune and sim.use_graphs:
        from baking_amd.autotune import autotune
        tuned = autotune(sim)
    if model is not None:
        surrogate.detach(sim)

    for _ in range(args.warmup):
        sim._launch_substep()
        sim._accept()
    torch.cuda.synchronize()
    flags_before = sim._read_flags()
    snapshot = sim._save_for_tuning()
    start_aggr = measure(sim)

    def timed():
        if sim.use_graphs:  # capture the graphs the timed run replays outside the timed region
            sim._launch_substeps(max(2, sim.graph_steps), stream)
        sim._restore_after_tuning(snapshot)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record(stream)
        sim._launch_substeps(steps, stream)  # as graphs: pipelined, ending with the last substep's deferred pass
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop), time.perf_counter() - t0

    surrogate.stats(sim, reset=True)
    exact_ms, wall = timed()
    exact_aggr, exact_stats = measure(sim), surrogate.stats(sim, reset=True)
    gpu_ms, aggr_used, stats_used, accuracy, method = exact_ms, exact_aggr, exact_stats, None, "plain (exact return map)"
    if model is not None:
        surrogate.attach(sim, model, strict=False)
        surrogate.stats(sim, reset=True)
        gpu_ms, wall = timed()
        aggr_used, stats_used = measure(sim), surrogate.stats(sim, reset=True)
        accuracy = _relative_errors(exact_aggr, aggr_used, start_aggr)
        method = "plain + surrogate-started return map"
    flags = sim._read_flags()
    del snapshot

    stages, counts = stage_profile(lib_launch=lambda: (sim._launch_substep(use_graph=False), sim._accept()),
                                   workspace=sim._workspace, steps=args.profile_steps)
    deferred = int(counts[2])
    active_nodes = int((sim.grid_state.tiled_masses > 0).sum().item())
    substep_ms = gpu_ms / steps
    roof, dominant, dom_bytes, dom_ms = roofline(stages, n, active_nodes, deferred, substep_ms, spec.name)
    yield_fraction = float((sim.current_state.material_state.return_residuals > 0).float().mean().item()) \
        if sim.current_state.material_state is not None else 0.0
    per_map = lambda st: st["evaluations"] / st["return_maps"] if st["return_maps"] else 0.0  # noqa: E731

    value = n * steps / (gpu_ms * 1e-3)
    result = {
        "metric": "particle-substeps/s on 1 MI355X, plain vs surrogate+batched at <=1e-3 rel err",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": substep_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": spec.name, "method": method, "hip_graph": sim.use_graphs, "particles": n,
                   "grid": spec.simulation.grid_resolution, "dt": spec.simulation.time_step, "active_stencil_nodes": active_nodes,
                   "yield_fraction": yield_fraction, "deferred_return_maps": deferred,
                   "tiles": int(counts[0]), "active_node_blocks": int(counts[1]), "tuning": dict(sim.Tuning(), graph_steps=sim.graph_steps),
                   "autotune": None if tuned is None else {"start": tuned["start"], "best_ms": tuned["ms"],
                                                           "trials": len(tuned["trials"])},
                   "return_map_stats": {"exact": exact_stats, "used": stats_used,
                                        "evaluations_per_map": {"exact": per_map(exact_stats), "used": per_map(stats_used)}},
                   "wall_s": wall, "flags": flags, "flags_before_timing": flags_before},
        "exact": {"value": n * steps / (exact_ms * 1e-3), "ms_per_step": exact_ms / steps},
        "accuracy_vs_exact": accuracy,
        "roofline": roof,
        "cpu_baseline": None,
    }
    result["check_state"] = check_state_timing(spec, sim, with_host=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(spec, args.cpu_baseline_steps)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
