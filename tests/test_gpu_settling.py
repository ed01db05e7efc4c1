"""GPU parity of the rest before baking (reference tests/test_settling.py, tests/test_auto_settling.py).

Each test replays one of the reference's settling tests on the same scene and
configuration (the transport cube of tests/test_transport.py:12-20 with pore
gas) and asserts what it asserts.  The same trajectory also runs through the
FP32 and FP64 oracle: the settling decisions (substep of the transition, quiet
substeps, reason) must agree exactly, the diagnostics to FP32 rounding, and
the fields with the three-way criterion of the other parity tests.  The rest
itself runs on the device (the settling controller), so these tests also pin
that a resting ``Advance`` needs no per-substep host decision.
"""
from dataclasses import replace

import numpy as np
import pytest

from tests.harness import check_fields, field_errors, field_scales, gpu_fields, oracle_from_spec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import adjoint, scenes  # noqa: E402
from baking_amd.launch import launch  # noqa: E402
from baking_amd.physics.batch import BatchSimulator  # noqa: E402
from baking_amd.physics.quasistatic import SettlingMetrics  # noqa: E402
from oracle import oracle as O  # noqa: E402

RTOL = 1.0e-5
T_REST = 293.15  # TransportConfig.initial_temperature_k


def _check_fields(tag, sim, o32, o64):
    check_fields(tag, gpu_fields(sim), o32, o64)


def _same_decisions(tag, sim, orc):
    assert (sim.IsSettling, sim.settling_reason, sim.settling_quiet_steps, sim.settling_steps, sim.baking_steps) == \
        (orc.IsSettling, orc.settling_reason, orc.settling_quiet_steps, orc.settling_steps, orc.baking_steps), tag
    assert sim.settling_reference_step == orc.settling_reference_step, tag
    got = np.array(list(sim.settling_diagnostics.values()))
    want = np.array(list(orc.settling_diagnostics.values()))
    # speed, strain rate, plastic rate.  Rounding noise is absolute: the strain rate of a rigidly moving dough is
    # |v| 4/dx times FP32 rounding (~1e-5 1/s here; 1e-14 in FP64), and the plastic rate differences near-identity
    # FP32 gradients (~1e-7 / (window |Fp|)); both bounds sit 100x below the reference's 0.01 1/s tolerance
    for name, g, w, atol in zip(sim.settling_diagnostics, got, want, (1e-7, 1e-4, 1e-3)):
        np.testing.assert_allclose(g, w, rtol=1e-3, atol=atol, err_msg=f"{tag}: {name}")


def _both(spec):
    return scenes.build_simulator(spec), oracle_from_spec(spec, "f32"), oracle_from_spec(spec, "f64")


def _f(t):
    return t.detach().cpu().numpy()


def test_settling_metrics_kernel():
    """tests/test_auto_settling.py:66-76 (closed form: |sym(0.1 I)| = 0.1 sqrt 3, |0.01 I| / (0.01 |I|) = 1), then
    random states against the oracle."""
    dev = "cuda"
    eye = torch.eye(3, device=dev)[None]
    metrics = torch.zeros(3, device=dev)
    launch(SettlingMetrics, 1, inputs=[torch.zeros(1, 3, device=dev), eye * 0.1, eye.clone(), eye * 1.01, metrics, 0.01])
    values = _f(metrics)
    assert values[0] == 0 and values[1] > 0.1 and values[2] > 0.9
    np.testing.assert_allclose(values, [0.0, 0.1 * np.sqrt(3.0), 1.0], rtol=2e-6)
    rng = np.random.default_rng(11)
    n = 20000
    v = rng.normal(size=(n, 3)).astype(np.float32)
    C = rng.normal(size=(n, 3, 3)).astype(np.float32)
    old = (np.eye(3) + 0.1 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    new = (old + 1e-3 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    want = np.zeros(3, np.float32)
    O.Kernels("f32").SettlingMetrics(n, v, C, old, new, want, 2.5e-4)
    got = torch.zeros(3, device=dev)
    t = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    launch(SettlingMetrics, n, inputs=[t(v), t(C), t(old), t(new), got, 2.5e-4])
    np.testing.assert_allclose(_f(got), want, rtol=1e-6)


def test_rest_is_isothermal_mass_conserving_and_falls_then_heats():
    """tests/test_settling.py:15-35: 4 resting substeps under gravity change no water, gas or temperature; the
    fifth bakes."""
    spec = scenes.demo_resting()
    sim, o32, o64 = _both(spec)
    s0 = sim.current_state
    water = _f(s0.transport_state.water_masses).copy()
    gas = _f(s0.material_state.gas_moles).copy()
    y = _f(s0.particle_positions)[:, 1].mean()
    assert sim.IsSettling and sim.settling_steps == 4 and sim.settling_reason == "pending"
    sim.Advance(4)
    for o in (o32, o64):
        for _ in range(4):
            o.Step()
    s = sim.current_state
    np.testing.assert_array_equal(_f(s.transport_state.water_masses), water)
    np.testing.assert_array_equal(_f(s.material_state.gas_moles), gas)
    np.testing.assert_allclose(_f(s.transport_state.temperatures), T_REST, atol=1e-12)
    np.testing.assert_array_equal(_f(s.transport_state.water_losses), 0.0)
    np.testing.assert_array_equal(_f(s.transport_state.heat_exchanges), 0.0)
    assert _f(s.particle_positions)[:, 1].mean() < y
    assert _f(s.particle_velocities)[:, 1].mean() < 0
    assert not sim.IsSettling and sim.settling_reason == "fixed_duration" and sim.baking_time == 0
    _same_decisions("after the rest", sim, o32)
    _check_fields("demo_resting 4 substeps", sim, o32, o64)
    vy = _f(s.particle_velocities)[:, 1].mean()
    sim.Step()
    for o in (o32, o64):
        o.Step()
    s = sim.current_state
    assert sim.baking_time == sim.simulation_config.time_step
    assert _f(s.transport_state.temperatures).max() > T_REST
    assert _f(s.material_state.gas_moles).min() > gas.min()
    assert _f(s.particle_velocities)[:, 1].mean() < vy
    _check_fields("demo_resting 5 substeps", sim, o32, o64)
    sim.Advance(3)
    for o in (o32, o64):
        for _ in range(3):
            o.Step()
    _same_decisions("baking", sim, o32)
    _check_fields("demo_resting 8 substeps", sim, o32, o64)


def test_zero_rest_preserves_legacy_and_duration_rounding():
    """tests/test_settling.py:58-64: the rest follows transport.config, rounded up to whole substeps."""
    sim = scenes.build_simulator(scenes.demo_resting())
    sim.transport.config = replace(sim.transport.config, settling_duration_s=0.00041)
    assert sim.settling_steps == 5
    sim.transport.config = replace(sim.transport.config, settling_duration_s=0.0)
    assert sim.settling_steps == 0
    sim.Step()
    assert _f(sim.current_state.transport_state.temperatures).max() > T_REST


def test_quiet_confirmation_then_baking_and_reset():
    """tests/test_auto_settling.py:19-49 without the recorder: quiet from substep 3, baking after substep 5."""
    spec = scenes.demo_auto_settling()
    sim, o32, o64 = _both(spec)
    for k in range(4):
        sim.Step()
        for o in (o32, o64):
            o.Step()
        assert sim.IsSettling and sim.baking_time == 0
        np.testing.assert_allclose(_f(sim.current_state.transport_state.temperatures), T_REST, atol=1e-12)
        _same_decisions(f"substep {k + 1}", sim, o32)
    sim.Step()
    for o in (o32, o64):
        o.Step()
    assert not sim.IsSettling and sim.settling_reason == "quasi_static"
    assert sim.settling_steps == 5 and sim.baking_time == 0
    _same_decisions("transition", sim, o32)
    sim.Step()
    for o in (o32, o64):
        o.Step()
    assert sim.baking_steps == 1 and sim.baking_time == sim.simulation_config.time_step
    assert _f(sim.current_state.transport_state.temperatures).max() > T_REST
    _check_fields("demo_auto_settling 6 substeps", sim, o32, o64)
    stepped = gpu_fields(sim)
    sim.Reset()
    assert sim.IsSettling and sim.settling_quiet_steps == 0 and sim.baking_steps == 0
    # the same 6 substeps in one Advance: the controller decides on the device
    sim.Advance(6)
    assert not sim.IsSettling and sim.settling_steps == 5 and sim.baking_steps == 1
    again = gpu_fields(sim)
    errs = field_errors(again, stepped, floors=field_scales(stepped, o32))
    assert max(errs.values()) <= RTOL, errs


def test_moving_object_times_out_without_heat():
    """tests/test_auto_settling.py:52-57: a falling dough never settles; the rest times out at substep 20."""
    spec = scenes.demo_auto_settling(gravity=(0.0, -9.81, 0.0))
    sim, o32, o64 = _both(spec)
    with pytest.raises(RuntimeError, match="baking was not started"):
        sim.Advance(20)
    assert sim.IsSettling and sim.settling_reason == "timeout" and sim.baking_time == 0
    assert sim.step_count == 20
    np.testing.assert_allclose(_f(sim.current_state.transport_state.temperatures), T_REST, atol=1e-12)
    for o in (o32, o64):
        with pytest.raises(RuntimeError, match="baking was not started"):
            for _ in range(20):
                o.Step()
        assert o.step_count == 20
    _same_decisions("timeout", sim, o32)
    assert sim.settling_diagnostics["max_speed_m_s"] > 1e-4
    _check_fields("falling dough 20 substeps", sim, o32, o64)


def test_brief_quiet_and_deformation_cannot_trigger_transition():
    """tests/test_auto_settling.py:60-65: a kick after one quiet substep restarts the count."""
    sim = scenes.build_simulator(scenes.demo_auto_settling())
    sim.Advance(3)
    assert sim.settling_quiet_steps == 1
    sim.current_state.particle_velocities.copy_(torch.tensor([0.01, 0.0, 0.0], device="cuda").expand_as(
        sim.current_state.particle_velocities))
    sim.Step()
    assert sim.IsSettling and sim.settling_quiet_steps == 0


def _mean_temperature_gradient(spec, steps=8, h=20.0):
    sim = scenes.build_simulator(spec)
    param = torch.tensor([h], device="cuda", requires_grad=True)
    out = adjoint.rollout(sim, steps, heat_transfer_parameter=param)
    out["T"].mean().backward()
    return float(param.grad), sim


def _mean_temperature(spec, h, steps=8):
    sim = scenes.build_simulator(spec)
    sim.transport.heat_transfer_parameter.fill_(h)
    sim.Advance(steps, check=False)
    return float(sim.current_state.transport_state.temperatures.mean())


def test_gradient_across_rest_baking_boundary():
    """tests/test_settling.py:38-55: d mean(T) / d h over 4 resting and 4 baking substeps."""
    spec = scenes.demo_resting()
    grad, sim = _mean_temperature_gradient(spec)
    assert sim.settling_steps == 4 and sim.baking_steps == 4
    finite = (_mean_temperature(spec, 22.0) - _mean_temperature(spec, 18.0)) / 4.0

    def oracle_loss(value):
        o = oracle_from_spec(spec, "f64")
        o.heat_param[0] = value
        for _ in range(8):
            o.Step()
        return float(o.state.transport["T"].mean())

    fd64 = (oracle_loss(20.5) - oracle_loss(19.5)) / 1.0
    print(f"gradient {grad:.9g}, GPU FD (eps 2) {finite:.9g}, FP64 oracle FD (eps 0.5) {fd64:.9g}")
    assert grad > 0
    np.testing.assert_allclose(grad, finite, rtol=0.003, atol=1e-10)
    np.testing.assert_allclose(grad, fd64, rtol=1e-3)


def test_gradient_with_unchanged_stopping_step():
    """tests/test_auto_settling.py:79-98: the automatic rest stops after substep 5 at h = 18, 20 and 22."""
    spec = scenes.demo_auto_settling()
    grad, sim = _mean_temperature_gradient(spec)
    assert sim.settling_steps == 5
    finite = (_mean_temperature(spec, 22.0) - _mean_temperature(spec, 18.0)) / 4.0
    print(f"gradient {grad:.9g}, GPU FD (eps 2) {finite:.9g}")
    assert grad > 0
    np.testing.assert_allclose(grad, finite, rtol=0.003, atol=1e-10)


def test_batch_rests_each_dough_as_alone():
    """A batch of a fixed rest, an automatic rest and a dough without rest: each rests and bakes as it does alone."""
    specs = [scenes.demo_resting(), scenes.demo_auto_settling(), scenes.demo_transport(water_stress=10000.0)]
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs])
    batch.Advance(3)
    batch.Advance(5)
    for d, spec in enumerate(specs):
        dough = batch.doughs[d]
        alone, o32, o64 = _both(spec)
        alone.Advance(8)
        for o in (o32, o64):
            for _ in range(8):
                o.Step()
        assert (dough.IsSettling, dough.settling_reason, dough.settling_steps, dough.baking_steps) == \
            (alone.IsSettling, alone.settling_reason, alone.settling_steps, alone.baking_steps), spec.name
        _same_decisions(f"{spec.name} in the batch", dough, o32)
        _check_fields(f"{spec.name} in the batch", dough, o32, o64)
    assert [batch.doughs[d].settling_reason for d in range(3)] == ["fixed_duration", "quasi_static", "disabled"]
    assert not batch._table_rests  # once every dough bakes, the substep no longer runs the controller


def test_batch_names_the_dough_whose_rest_times_out():
    specs = [scenes.demo_auto_settling(), scenes.demo_auto_settling(gravity=(0.0, -9.81, 0.0))]
    batch = BatchSimulator([scenes.build_simulator(s) for s in specs])
    with pytest.raises(RuntimeError, match=r"doughs \[1\]"):
        batch.Advance(25)
    assert batch.step_count == 20
    assert batch.doughs[0].settling_reason == "quasi_static" and batch.doughs[1].settling_reason == "timeout"
