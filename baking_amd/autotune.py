"""Launch-shape autotuner of the fused substep (no reference counterpart).

Times the substep of a live simulator (``MpmSimulator`` or ``BatchSimulator``)
under candidate ``bk_tuning`` values and keeps the fastest:

* ``tile_chunk``     -- most particles per P2G tile: small tiles spread a
  dense particle block over several workgroups (more parallelism, more flush
  atomics); large tiles keep one workgroup per block.
* ``particle_block`` -- workgroup size of the per-particle passes.
* ``grid_cap``       -- workgroups of the grid-stride tile and node-block passes.
* ``deferred_wgs``   -- workgroups of the deferred return-map pass where it runs beside the next substep in a
  pipelined run of substeps (0: runs are not pipelined).  A short queue finishes within the next substep's first
  stages however it is spread; a long one slows them unless it is held to fewer wavefront slots.
* ``graph_steps``    -- substeps per HIP-graph launch in a run of substeps (the simulator's ``graph_steps``):
  more per launch saves the gap between launches, but the pipelined deferred passes inside one graph may then be
  scheduled less well.

Coordinate descent: each parameter in turn is swept with the others held,
twice over.  Every trial replays the captured HIP graph for ``steps``
substeps (after ``warmup`` untimed ones) between HIP events on the launch
stream.  The simulator's particle state, substep count and flags are
restored afterwards and its grid is zeroed (each substep rebuilds the grid
from the particles), so tuning leaves the simulation where it was.
"""
from __future__ import annotations

import torch

CANDIDATES = {
    "tile_chunk": (32, 64, 128, 256, 512, 4096),
    "particle_block": (64, 128, 256),
    "grid_cap": (512, 1024, 2048, 4096, 8192),
    "deferred_wgs": (0, 128, 256, 512, 1024),
    "graph_steps": (1, 2, 4, 8),
}


def _tuning(sim):
    """The simulator's bk_tuning values and its substeps per graph launch."""
    return dict(sim.Tuning(), graph_steps=sim.graph_steps)


def _set_tuning(sim, tuning):
    tuning = dict(tuning)
    sim.graph_steps = tuning.pop("graph_steps")
    sim.SetTuning(**tuning)


def time_substeps(sim, steps, warmup):
    """Milliseconds per substep of the simulator's current tuning (graph replay, HIP events)."""
    stream = torch.cuda.current_stream()
    sim._launch_substeps(warmup)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    sim._launch_substeps(steps)
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / steps


def autotune(sim, steps=10, warmup=3, rounds=2, candidates=None, log=None):
    """Pick the fastest launch shapes for ``sim`` and apply them.

    Returns {"best": tuning, "ms": ms per substep, "start": (tuning, ms), "trials": [(tuning, ms), ...]}.
    """
    candidates = candidates or CANDIDATES
    saved = sim._save_for_tuning()
    try:
        best = _tuning(sim)
        _set_tuning(sim, best)
        best_ms = time_substeps(sim, steps, warmup)
        trials = [(dict(best), best_ms)]
        start = (dict(best), best_ms)
        for _ in range(rounds):
            improved = False
            for name, values in candidates.items():
                for value in values:
                    if value == best[name]:
                        continue
                    trial = dict(best, **{name: value})
                    _set_tuning(sim, trial)
                    ms = time_substeps(sim, steps, warmup)
                    trials.append((trial, ms))
                    if log:
                        log(f"autotune {trial}: {ms:.4f} ms/substep")
                    if ms < best_ms:
                        best, best_ms, improved = trial, ms, True
            if not improved:
                break
        _set_tuning(sim, best)
    finally:
        sim._restore_after_tuning(saved)
    return {"best": best, "ms": best_ms, "start": start, "trials": trials}
