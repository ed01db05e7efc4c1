"""Pins of the oracle's settling restatement (reference physics/quasistatic.py:11-21, physics/mpm.py:57-109).

The oracle's SettlingMetrics is checked against closed forms, and its
UpdateSettling against what the reference's own settling tests assert
(tests/test_settling.py, tests/test_auto_settling.py) on the same scenes:
the substep of the transition, the quiet-substep count, the reason, the
timeout, and a rest that changes no water, gas or temperature.
"""
import numpy as np
import pytest

from baking_amd import scenes
from oracle import oracle as O
from tests.harness import oracle_from_spec

T_REST = 293.15


def _metrics(v, C, old, new, elapsed):
    out = np.zeros(3, np.float32)
    O.Kernels("f32").SettlingMetrics(len(v), np.ascontiguousarray(v, np.float32), np.ascontiguousarray(C, np.float32),
                                     np.ascontiguousarray(old, np.float32), np.ascontiguousarray(new, np.float32), out, elapsed)
    return out


def test_settling_metrics_closed_forms():
    eye = np.eye(3)[None]
    # tests/test_auto_settling.py:66-76
    m = _metrics(np.zeros((1, 3)), 0.1 * eye, eye, 1.01 * eye, 0.01)
    assert m[0] == 0 and m[1] > 0.1 and m[2] > 0.9
    np.testing.assert_allclose(m, [0.0, 0.1 * np.sqrt(3.0), 1.0], rtol=2e-6)
    # speed is |v|; a spin (antisymmetric C) has no strain rate; the plastic rate is relative to |Fp_ref|
    spin = np.array([[[0.0, 2.0, 0.0], [-2.0, 0.0, 0.0], [0.0, 0.0, 0.0]]])
    m = _metrics(np.array([[3.0, 4.0, 0.0]]), spin, 2.0 * eye, 2.0 * eye + np.diag([0.0, 0.0, 0.3])[None], 0.5)
    np.testing.assert_allclose(m, [5.0, 0.0, 0.3 / (0.5 * 2.0 * np.sqrt(3.0))], rtol=1e-6)
    # the maxima are taken per metric over the particles, each from a different particle
    v = np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 0]])
    C = np.stack([np.zeros((3, 3)), np.diag([1.0, 2.0, 2.0]), np.zeros((3, 3))])
    old = np.stack([np.eye(3)] * 3)
    new = old + np.stack([np.zeros((3, 3)), np.zeros((3, 3)), np.diag([0.0, 0.0, 0.1])])
    np.testing.assert_allclose(_metrics(v, C, old, new, 1.0), [1.0, 3.0, 0.1 / np.sqrt(3.0)], rtol=1e-6)


def test_fixed_rest_is_isothermal_and_mass_conserving_then_heats():
    """tests/test_settling.py:15-35 and 58-64 on the oracle."""
    o = oracle_from_spec(scenes.demo_resting())
    assert o.IsSettling and o.settling_steps == 4 and o.settling_reason == "pending"
    water, gas, y = o.state.transport["W"].copy(), o.state.ding["moles"].copy(), o.state.x[:, 1].mean()
    for _ in range(4):
        o.Step()
    t = o.state.transport
    np.testing.assert_array_equal(t["W"], water)
    np.testing.assert_array_equal(o.state.ding["moles"], gas)
    np.testing.assert_allclose(t["T"], T_REST, atol=1e-12)
    assert not t["losses"].any() and not t["exchanges"].any()
    assert o.state.x[:, 1].mean() < y and o.state.v[:, 1].mean() < 0
    assert not o.IsSettling and o.settling_reason == "fixed_duration" and o.baking_time == 0
    o.Step()
    assert o.baking_time == o.dt and o.state.transport["T"].max() > T_REST and o.state.ding["moles"].min() > gas.min()
    z = oracle_from_spec(scenes.demo_resting(settling_duration_s=0.00041))
    assert z.settling_steps == 5


def test_auto_rest_confirms_quiet_then_bakes():
    """tests/test_auto_settling.py:19-49: quiet from substep 3 (min 0.2 ms), baking after 3 quiet substeps (0.3 ms)."""
    o = oracle_from_spec(scenes.demo_auto_settling())
    quiet = []
    for _ in range(5):
        o.Step()
        quiet.append(o.settling_quiet_steps)
        assert o.IsSettling == (o.step_count < 5)
    assert quiet == [0, 0, 1, 2, 3]
    assert o.settling_reason == "quasi_static" and o.settling_steps == 5 and o.baking_time == 0
    assert o.settling_reference_step == 3  # the first plastic-rate window closed after 0.3 ms
    np.testing.assert_allclose(o.state.transport["T"], T_REST, atol=1e-12)
    o.Step()
    assert o.baking_steps == 1 and o.state.transport["T"].max() > T_REST


def test_falling_dough_times_out_and_a_kick_resets_the_count():
    """tests/test_auto_settling.py:52-65."""
    o = oracle_from_spec(scenes.demo_auto_settling(gravity=(0.0, -9.81, 0.0)))
    with pytest.raises(RuntimeError, match="baking was not started"):
        for _ in range(25):
            o.Step()
    assert o.step_count == 20 and o.IsSettling and o.settling_reason == "timeout" and o.baking_time == 0
    assert o.settling_diagnostics["max_speed_m_s"] > 1e-4
    np.testing.assert_allclose(o.state.transport["T"], T_REST, atol=1e-12)
    k = oracle_from_spec(scenes.demo_auto_settling())
    for _ in range(3):
        k.Step()
    assert k.settling_quiet_steps == 1
    k.state.v[...] = np.array([0.01, 0.0, 0.0], np.float32)
    k.Step()
    assert k.IsSettling and k.settling_quiet_steps == 0


def test_falling_dough_metrics_in_fp64_are_free_fall():
    """In FP64 the falling cube translates rigidly: max speed = g t exactly as far as rounding goes, strain and
    plastic rates at rounding level (the FP32 run's ~1e-5 1/s strain rate is its rounding noise)."""
    spec = scenes.demo_auto_settling(gravity=(0.0, -9.81, 0.0))
    o = oracle_from_spec(spec, "f64")
    with pytest.raises(RuntimeError, match="baking was not started"):
        for _ in range(25):
            o.Step()
    d = o.settling_diagnostics
    np.testing.assert_allclose(d["max_speed_m_s"], 9.81 * 20 * spec.simulation.time_step, rtol=1e-9)
    assert d["max_strain_rate_per_s"] < 1e-9 and d["max_plastic_rate_per_s"] < 1e-9
