"""GPU parity of the batched substep: every dough of a batch evolves as it does alone.

Each dough of a small uneven batch (the batched workload's generator at
smaller half-extents) is checked after several substeps against the FP32 and
FP64 oracle runs of that dough alone, with the three-way criterion of the
single-dough tests, and against the same dough run alone on the GPU.  The
batch runs with small P2G tiles, so blocks are split across workgroups.
Error flags stay with their dough.
"""
import numpy as np
import pytest

from tests.harness import check_fields, check_same, gpu_fields, oracle_from_spec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import scenes  # noqa: E402
from baking_amd.physics.batch import BatchSimulator  # noqa: E402


def _small_batch(count=4, seed=5):
    return scenes.batched_doughs(count=count, seed=seed, half_extent=(0.02, 0.035), max_particles=6000)


@pytest.mark.parametrize("use_graphs", [True, False])
def test_batch_matches_each_dough_alone(use_graphs):
    specs = _small_batch()
    steps = 8
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs],
                           tuning=dict(tile_chunk=48, particle_block=128, grid_cap=512))
    batch.use_graphs = use_graphs
    print("batch particles", [int(n) for n in np.diff(batch.particle_offsets)], "tuning", batch.Tuning(), flush=True)
    batch.Advance(steps)
    for d, spec in enumerate(specs):
        alone = scenes.build_simulator(spec)
        alone.Advance(steps)
        o32, o64 = oracle_from_spec(spec, "f32"), oracle_from_spec(spec, "f64")
        for _ in range(steps):
            o32.Step()
            o64.Step()
        got = gpu_fields(batch.doughs[d])
        check_fields(f"{spec.name} batch vs oracle", got, o32, o64)
        check_fields(f"{spec.name} alone vs oracle", gpu_fields(alone), o32, o64)
        check_same(f"{spec.name}: batch vs the dough alone", got, gpu_fields(alone), o32, o64)
        assert batch.doughs[d].step_count == steps
    assert batch.ReadFlags().max() == 0


def test_batch_tuning_does_not_change_results():
    specs = _small_batch(count=3, seed=6)
    runs = []
    for tuning in (dict(tile_chunk=4096, particle_block=256, grid_cap=4096), dict(tile_chunk=32, particle_block=64, grid_cap=96)):
        batch = BatchSimulator([scenes.build_simulator(s) for s in specs], tuning=tuning)
        batch.Advance(5)
        runs.append([gpu_fields(s) for s in batch.doughs])
    for d, spec in enumerate(specs):
        o32, o64 = oracle_from_spec(spec, "f32"), oracle_from_spec(spec, "f64")
        for _ in range(5):
            o32.Step()
            o64.Step()
        check_same(f"{spec.name}: tuning changed results", runs[1][d], runs[0][d], o32, o64)


def test_batch_flags_stay_with_their_dough():
    """A dough pushed off its grid raises the stencil flag in its own flag block only."""
    specs = _small_batch(count=3, seed=7)
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs])
    x = batch.doughs[1].current_state.particle_positions
    x[:8, 0] = 0.999  # base node res - 1: the stencil leaves the grid (guarded in every kernel)
    with pytest.raises(RuntimeError, match="doughs"):
        batch.Advance(1)
    flags = batch.ReadFlags()
    assert flags[1, 1] == 1
    assert flags[0].max() == 0 and flags[2].max() == 0
    with pytest.raises(RuntimeError, match="stencil"):
        batch.CheckState(1)
    assert batch.CheckState(0)["step_count"] == 1
