"""Oracle pins for the Ding return map: the reference tests' closed forms, invariants and known answers.

Restates reference tests/test_ding.py and tests/test_ding_thermal.py against
the CPU oracle, plus the SVD against LAPACK.
"""
import numpy as np
import pytest

from baking_amd.config import DingConfig, MaterialConfig
from tests.harness import O, OraclePoint


def _rotation(angle, axis=2):
    c, s = np.cos(angle), np.sin(angle)
    R = np.eye(3)
    i, j = [a for a in range(3) if a != axis]
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


@pytest.mark.parametrize("seed", range(6))
def test_svd_matches_lapack(seed):
    rng = np.random.default_rng(seed)
    cases = [np.eye(3), np.diag([1.3, 1.3, 0.7]), _rotation(0.4) @ np.diag([2.0, 2.0, 2.0]),
             _rotation(0.3, 0) @ np.diag([1.1, 0.9, 0.9]) @ _rotation(-1.1, 1).T]
    F = rng.normal(size=(3, 3)) * 0.3 + np.eye(3)
    if np.linalg.det(F) < 0:
        F[:, 0] *= -1
    cases.append(F)
    for A in cases:
        U, S, V = O.svd3(A)
        np.testing.assert_allclose(U @ np.diag(S) @ V.T, A, atol=1e-13)
        np.testing.assert_allclose(np.sort(S)[::-1], np.linalg.svd(A, compute_uv=False), rtol=1e-13)
        np.testing.assert_allclose(U.T @ U, np.eye(3), atol=1e-13)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-13)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0


def test_elastic_limit_objectivity_and_porosity():
    """Reference tests/test_ding.py:59-78."""
    mat = MaterialConfig()
    pt = OraclePoint(mat, DingConfig(yield_stress=1e9))
    F = np.diag([1.1, 0.95, 1.02])
    s = pt.step(F)
    eps = np.log(np.diag(F))
    expected = np.diag(2 * mat.shear_modulus * eps + mat.lame_lambda * eps.sum())
    np.testing.assert_allclose(s["tau"][0], expected, rtol=2e-5, atol=0.02)
    np.testing.assert_allclose(s["Fp"][0], np.eye(3), atol=1e-6)
    R = _rotation(0.5)
    s = pt.step(R @ F)
    np.testing.assert_allclose(s["tau"][0], R @ expected @ R.T, atol=0.05)
    for porosity in (0.0, 0.2):
        pt = OraclePoint(mat, DingConfig(yield_stress=100.0, porosity=porosity))
        s = pt.step(np.eye(3) * 0.9)
        plastic_volume = np.linalg.det(s["Fp"][0])
        if porosity == 0:
            np.testing.assert_allclose(plastic_volume, 1, atol=1e-5)
        else:
            assert plastic_volume < 0.999


@pytest.mark.parametrize("stretch", [1.25, 0.8])
def test_load_hold_unload(stretch):
    """Reference tests/test_ding.py:32-56."""
    mat = MaterialConfig()
    pt = OraclePoint(mat, DingConfig(yield_stress=500.0, viscosity=1e10))
    F = np.diag([stretch, 1.0, 1.0])
    s = pt.step(F)
    tau0 = s["tau"][0].copy()
    energies = []
    for _ in range(40):
        s = pt.step(F)
        eps = np.log(np.linalg.svd(s["Fe"][0], compute_uv=False))
        energies.append(mat.shear_modulus * np.dot(eps, eps) + 0.5 * mat.lame_lambda * eps.sum() ** 2)
    tau = s["tau"][0]
    dev = lambda t: t - np.eye(3) * np.trace(t) / 3  # noqa: E731
    assert tau0[0, 0] * (stretch - 1) > 0
    assert np.linalg.norm(dev(tau)) < np.linalg.norm(dev(tau0))
    assert np.max(np.diff(energies)) < 0.01
    np.testing.assert_allclose(np.linalg.det(s["Fp"][0]), 1.0, atol=2e-5)
    np.testing.assert_allclose(s["Fe"][0] @ s["Fp"][0], F, atol=1e-5)
    assert s["residual"].max() < 1e-5
    residual_gradient = s["Fp"][0].copy()
    s = pt.step(residual_gradient)
    np.testing.assert_allclose(s["tau"], 0, atol=0.5)
    assert np.linalg.norm(residual_gradient - np.eye(3)) > 0.01


def _stationarity(mat, ding, tau, trial_tau, a, dt):
    mu, lam = mat.shear_modulus, mat.lame_lambda
    compliance = np.linalg.inv(2 * mu * np.eye(3) + lam * np.ones((3, 3)))
    mean = tau.mean()
    dev = tau - mean
    f = 1.5 * np.dot(dev, dev) + a * mean ** 2 - ding.yield_stress ** 2
    grad = 3 * dev + (2 * a / 3) * mean * np.ones(3)
    return compliance @ (tau - trial_tau) + dt / ding.viscosity * max(f, 0) * grad


def test_local_stationarity_and_viscosity_dependence():
    """Reference tests/test_ding.py:174-193: the return map's discrete optimality condition."""
    mat = MaterialConfig()
    norms = []
    for visc in (1e9, 1e11):
        ding = DingConfig(yield_stress=300.0, viscosity=visc, porosity=0.2)
        pt = OraclePoint(mat, ding)
        F = np.diag([1.2, 0.9, 1.04])
        tau = np.diag(pt.step(F)["tau"][0]).astype(float)
        trial = (2 * mat.shear_modulus * np.eye(3) + mat.lame_lambda * np.ones((3, 3))) @ np.log(np.diag(F))
        a = 9.0 / (4.0 * np.log(1 / 0.2) ** 2 * (3.5 * 0.2 + 1.0))
        np.testing.assert_allclose(_stationarity(mat, ding, tau, trial, a, 0.01), 0, atol=2e-6)
        norms.append(np.linalg.norm(tau - tau.mean()))
    assert norms[1] > norms[0]


def test_load_rate_dependence():
    """Reference tests/test_ding.py:196-205."""
    stresses = []
    for total in (0.02, 0.2):
        pt = OraclePoint(MaterialConfig(), DingConfig(yield_stress=500.0, viscosity=1e10))
        for i in range(20):
            st = np.exp(np.log(1.2) * (i + 1) / 20)
            s = pt.step(np.diag([st, 1.0, 1.0]), total / 20)
        tau = s["tau"][0]
        stresses.append(np.linalg.norm(tau - np.trace(tau) / 3 * np.eye(3)))
    assert stresses[0] > stresses[1]


def test_temperature_and_porosity_parameter_values():
    """Reference tests/test_ding_thermal.py:61-80: table values at knots, between and outside."""
    cfg = DingConfig(temperature_enabled=True)
    for T, yf, vf in ((260.0, 1.5, 1.5), (278.15, 1.5, 1.5), (293.15, 1.0, 1.0), (303.15, 0.6, 0.6), (313.15, 0.2, 0.2),
                      (333.15, 1.1, 0.125), (353.15, 2.0, 0.05), (400.0, 2.0, 0.05)):
        pt = OraclePoint(MaterialConfig(), cfg, initial_temperature_k=T)
        s = pt.step(np.eye(3))
        np.testing.assert_allclose(s["yield_stress"], cfg.yield_stress * yf, rtol=2e-6)
        np.testing.assert_allclose(s["viscosity"], cfg.viscosity * vf, rtol=2e-6)
        np.testing.assert_allclose(s["T"], T, atol=2e-5)
        assert abs(O.temperature_factor(T, cfg.temperature_knots_k, cfg.yield_factors) - yf) < 2e-6 * yf
    porosities = np.array([0.0, 0.01, 0.1, 0.2, 0.5, 0.999], dtype=np.float32)
    got = np.array([O.porosity_coefficient(n) for n in porosities])
    expected = [0.0] + [9 / (4 * np.log(1 / v) ** 2 * (3.5 * v + 1)) for v in porosities[1:]]
    np.testing.assert_allclose(got, expected, rtol=5e-5)
    assert np.all(np.diff(got) > 0)


def test_initial_porosity_and_elastic_versus_plastic_volume():
    """Reference tests/test_ding_thermal.py:83-98: porosity = det Fp - 1 + n0; elastic volume change leaves it."""
    for porosity in (0.0, 0.2, 0.5):
        pt = OraclePoint(MaterialConfig(), DingConfig(yield_stress=1e9, porosity=porosity))
        s = pt.step(np.diag([1.1, 0.95, 1.02]))
        np.testing.assert_allclose(s["porosity"], porosity, atol=1e-6)
    for stretch in (0.98, 1.02):
        pt = OraclePoint(MaterialConfig(), DingConfig(yield_stress=100.0, porosity=0.2))
        for _ in range(5):
            s = pt.step(np.eye(3) * stretch)
            np.testing.assert_allclose(s["porosity"], np.linalg.det(s["Fp"]) - 1 + 0.2, atol=2e-7)
        assert (s["porosity"][0] - 0.2) * (stretch - 1) > 0


def test_updated_porosity_defines_next_return_map():
    """Reference tests/test_ding_thermal.py:101-121."""
    cfg = DingConfig(yield_stress=100.0, porosity=0.2)
    mat = MaterialConfig()
    pt = OraclePoint(mat, cfg)
    F = np.eye(3) * 0.98
    s = pt.step(F)
    old_tau = s["tau"][0].copy()
    old_n = float(s["porosity"][0])
    s = pt.step(F)
    tau = np.linalg.eigvalsh(s["tau"][0]).astype(float)
    trial = np.linalg.eigvalsh(old_tau).astype(float)
    a = 9 / (4 * np.log(1 / old_n) ** 2 * (3.5 * old_n + 1))
    np.testing.assert_allclose(_stationarity(mat, cfg, tau, trial, a, 0.01), 0, atol=2e-6)
    assert abs(old_n - cfg.porosity) > 1e-4


def test_gelatinization_latch_persists_on_cooling():
    """Reference tests/test_ding_thermal.py:228-234."""
    pt = OraclePoint(MaterialConfig(), DingConfig(temperature_enabled=True))
    for T, expected_yield, gelled in ((303.15, 1200., 0), (313.15, 400., 0), (353.15, 4000., 1), (293.15, 4000., 1)):
        pt.s["T"][...] = T
        s = pt.step(np.eye(3))
        np.testing.assert_allclose(s["yield_stress"], expected_yield, rtol=2e-6)
        assert s["gelled"][0] == gelled


@pytest.mark.parametrize("porosity,stretch", [(0.95, 1.2), (0.2, 0.7)])
def test_porosity_leaving_domain_sets_flag(porosity, stretch):
    """Reference tests/test_ding_thermal.py:153-157 ('porosity left' error = flag 0)."""
    pt = OraclePoint(MaterialConfig(), DingConfig(porosity=porosity, yield_stress=100.0, viscosity=1e4))
    pt.step(np.eye(3) * stretch)
    assert pt.flags[0] == 1


def test_near_closure_porosity_known_answer():
    """Reference tests/test_gas.py:208-222 known answer."""
    from_value = 3.725290298461914e-7
    n0 = np.float32(0.2)
    pv = 1.0 - float(n0) + float(np.float32(from_value))
    n = O.solve_porosity(4.115, 29.464, 0.001052, 0.003390, -0.0230067, pv, float(n0))
    pred = O.porosity_prediction(float(np.float32(n)), 4.115, 29.464, 0.001052, 0.003390, -0.0230067, pv, float(n0))
    assert 0 < n < 3.725290298461914e-7
    np.testing.assert_allclose(n, pred, atol=1e-12)


def test_overstress_newton_solves_the_scalar_equation():
    """The 32-step f32 Newton and <=48-step f64 Newton roots satisfy gamma + 1 = q/(1+d g)^2 + p/(1+m g)^2."""
    for q, p, d, m in ((4.0, 0.5, 0.01, 0.002), (1.5, 0.0, 1.0, 0.0), (30.0, 10.0, 3e-3, 1e-3)):
        g = O.solve_overstress(q, p, d, m)
        assert abs(g + 1 - q / (1 + d * g) ** 2 - p / (1 + m * g) ** 2) < 1e-4 * max(1, g)
        g64 = O.lib().oracle_overstress64(q, p, d, m)
        assert abs(g64 + 1 - q / (1 + d * g64) ** 2 - p / (1 + m * g64) ** 2) < 1e-11 * max(1, g64)
    assert O.solve_overstress(0.5, 0.3, 1.0, 1.0) == 0.0
