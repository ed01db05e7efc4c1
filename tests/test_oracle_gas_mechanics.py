"""Oracle pins for gas, P2G/G2P and the driver: the reference tests' closed forms and invariants.

Restates reference tests/test_gas.py and tests/test_mpm.py against the CPU
oracle, plus APIC exactness and momentum conservation of the transfers.
"""
from dataclasses import replace

import numpy as np
import pytest

from baking_amd.config import DingConfig, MaterialConfig, SimulationConfig, TransportConfig
from baking_amd.geometry.mesh_loader import CreateDemoMesh
from baking_amd.geometry.particle_sampling import SampleParticles
from tests.harness import O, OraclePoint


def test_ideal_gas_units_and_scaling():
    """Reference tests/test_gas.py:46-51."""
    got = [O.gas_pressure(0.2, T, J, n) for T, J, n in ((300., 1., .2), (600., 1., .2), (300., 2., .2), (300., 1., .1))]
    np.testing.assert_allclose(np.array(got) / (O.GAS_CONSTANT * 300), [1, 2, 0.5, 2], rtol=2e-7)


def _gas_stress(ding, F, moles, T=293.15, porosity=None, density=1000.0):
    k = O.Kernels("f32")
    n = 1
    tau = np.zeros((n, 3, 3), np.float32)
    out = np.zeros((n, 3, 3), np.float32)
    flags = np.zeros(3, np.int32)
    mat = MaterialConfig()
    k.CombineGasStress(n, tau, np.asarray(F, np.float32)[None].copy(), np.full(n, porosity or ding.porosity, np.float32),
                       np.full(n, moles), np.full(n, T, np.float32), out, flags, ding.gas_ambient_pressure_pa, ding.porosity,
                       np.full(n, density, np.float32), mat.lame_lambda + 2 * mat.shear_modulus, 1e-4, 0.025, 0.4)
    return out[0], flags


def test_stress_conversion_ambient_and_no_source_conservation():
    """Reference tests/test_gas.py:54-76."""
    ding = DingConfig(gas_enabled=True, porosity=0.2, gas_production_rate=0.0)
    m0 = 101325 * 0.2 / (O.GAS_CONSTANT * 293.15)
    s, _ = _gas_stress(ding, np.eye(3), m0)
    np.testing.assert_allclose(s, 0, atol=0.02)
    s, _ = _gas_stress(ding, np.eye(3), 2 * m0)
    np.testing.assert_allclose(s, -101325 * np.eye(3), rtol=2e-6, atol=0.03)
    vac = replace(ding, gas_ambient_pressure_pa=0)
    s, _ = _gas_stress(vac, np.eye(3) * np.float32(2 ** (1 / 3)), 2 * m0)
    np.testing.assert_allclose(s, -202650 * np.eye(3), rtol=2e-6, atol=0.05)
    pt = OraclePoint(MaterialConfig(), ding)
    m_init = pt.s["moles"].copy()
    for _ in range(10):
        pt.step(np.eye(3), 0.01)
    np.testing.assert_array_equal(pt.s["moles"], m_init)


def test_local_gaussian_source_constant_temperature():
    """Reference tests/test_gas.py:74-88: the Gaussian source integrates to rate * density * time."""
    k = O.Kernels("f32")
    for steps in (1, 20, 100):
        T = np.array([333.15])
        old = np.zeros(1)
        for _ in range(steps):
            new = np.zeros(1)
            pres = np.zeros(1)
            k.LocalGas(1, old, T, T, new, pres, np.eye(3, dtype=np.float32)[None].copy(), np.array([0.2], np.float32),
                       np.array([0.002], np.float32), 1000., 333.15, 15., 0.1 / steps, 1)
            old = new
        np.testing.assert_allclose(old, 0.2, rtol=2e-6)


def test_coupled_compaction_consistency_and_small_pores():
    """Reference tests/test_gas.py:91-117."""
    ding = DingConfig(gas_enabled=True, porosity=0.2, yield_stress=100, viscosity=1e8, gas_production_rate=0)
    pt = OraclePoint(MaterialConfig(), ding)
    for i in range(30):
        s = pt.step(np.eye(3) * (1 - 0.001 * (i + 1)), 0.001)
        assert 0 < s["porosity"][0] <= 0.200001
        np.testing.assert_allclose(s["porosity"], np.linalg.det(s["Fp"]) - 1 + 0.2, atol=2e-6)
        assert s["porosity_residual"][0] < 1e-6
    assert s["porosity"][0] < 0.199
    tiny = DingConfig(gas_enabled=True, porosity=1e-7, gas_production_rate=0)
    st, _ = _gas_stress(tiny, np.eye(3), 101325 * 1e-7 / (O.GAS_CONSTANT * 293.15))
    assert np.isfinite(st).all()
    st, flags = _gas_stress(tiny, np.eye(3), 101325 * 0.2 / (O.GAS_CONSTANT * 293.15))
    assert np.isfinite(st).all() and abs(st[0, 0]) > 1e10 and flags[2] == 1


def _demo_sim(subdiv=3, gravity=(0.0, 0.0, 0.0), v0=(0.2, -0.1, 0.05), precision="f32"):
    mesh = CreateDemoMesh(subdiv)
    cfg = replace(SimulationConfig(), device_name="cpu", gravity=gravity)
    return O.OracleSimulator(mesh, SampleParticles(mesh, 1000.0, 4), cfg, MaterialConfig(), v0, precision=precision)


def test_free_translation_conserves_mass_momentum_and_shape():
    """Reference tests/test_mpm.py:18-37."""
    sim = _demo_sim()
    x0, verts0 = sim.state.x.copy(), sim.UpdateMesh().copy()
    m0 = sim.m0.sum(dtype=np.float64)
    sim.Advance(30)
    t = 30 * sim.dt
    v0 = np.array([0.2, -0.1, 0.05])
    np.testing.assert_allclose(sim.state.x, x0 + t * v0, atol=2e-6)
    np.testing.assert_allclose(sim.state.v, np.tile(v0, (sim.n, 1)), atol=2e-5)
    np.testing.assert_allclose(sim.state.F, np.tile(np.eye(3), (sim.n, 1, 1)), atol=1e-5)
    np.testing.assert_allclose(sim.UpdateMesh(), verts0 + t * v0, atol=2e-6)
    np.testing.assert_allclose(sim.node_m.sum(dtype=np.float64), m0, rtol=1e-6)
    assert sim.unmapped == 0


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_free_fall_matches_symplectic_euler(precision):
    """Reference tests/test_mpm.py:40-53 (and the FP64 run of the same kernels, 1e-9 tighter)."""
    sim = _demo_sim(gravity=(0.0, -9.81, 0.0), v0=(0.0, 0.0, 0.0), precision=precision)
    x0 = sim.state.x.copy()
    n = 40
    sim.Advance(n)
    exp = x0.astype(np.float64).copy()
    exp[:, 1] -= 9.81 * sim.dt ** 2 * n * (n + 1) / 2
    tol = 1.5e-6 if precision == "f32" else 1e-12
    np.testing.assert_allclose(sim.state.x, exp, atol=tol)
    v = np.zeros_like(exp)
    v[:, 1] = -9.81 * sim.dt * n
    np.testing.assert_allclose(sim.state.v, v, atol=1e-5 if precision == "f32" else 1e-11)


def test_apic_reproduces_affine_velocity_and_conserves_momentum():
    """Closed forms of quadratic-B-spline APIC: an affine field v = A x + b (C = A) round-trips exactly,
    and P2G conserves mass and linear momentum for any C and stress."""
    rng = np.random.default_rng(7)
    k = O.Kernels("f64")
    res, dx = 16, 1.0 / 16
    n = 300
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    v = x @ A.T + b
    C = np.tile(A, (n, 1, 1))
    m = rng.uniform(0.5, 2.0, n)
    V0 = np.full(n, dx ** 3 / 8)
    nm = np.zeros(res ** 3)
    npm = np.zeros((res ** 3, 3))
    flags = np.zeros(2, np.int32)
    origin = np.zeros(3)
    k.ParticleToGrid(n, x, v, C, np.zeros((n, 3, 3)), m, V0, nm, npm, origin, dx, res, 0.0, flags)
    np.testing.assert_allclose(nm.sum(), m.sum(), rtol=1e-13)
    np.testing.assert_allclose(npm.sum(axis=0), (m[:, None] * v).sum(axis=0), rtol=1e-12, atol=1e-12)
    # stress only redistributes momentum: its P2G contribution sums to zero
    npm2 = np.zeros_like(npm)
    k.ParticleToGrid(n, x, v, C, rng.normal(size=(n, 3, 3)) * 1e3, m, V0, np.zeros_like(nm), npm2, origin, dx, res, 1e-3, flags)
    np.testing.assert_allclose(npm2.sum(axis=0), npm.sum(axis=0), rtol=1e-9, atol=1e-9)
    nv = np.zeros_like(npm)
    k.UpdateGrid(O.i64(res ** 3), nm, npm, nv, res, 0, np.zeros(3), 0.0)
    xo, vo, Co, Fo = np.zeros_like(x), np.zeros_like(v), np.zeros_like(C), np.zeros_like(C)
    k.GridToParticle(n, x, np.tile(np.eye(3), (n, 1, 1)), nv, xo, vo, Co, Fo, origin, dx, res, 0.0, flags)
    np.testing.assert_allclose(vo, v, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(Co, C, rtol=1e-9, atol=1e-9)
    assert flags.max() == 0


def test_cell_keys_match_closed_form_and_sort_is_stable():
    """The cell key is floor(fl32((x - o) / dx) - 1/2) flattened; the permutation is numpy's stable argsort."""
    rng = np.random.default_rng(3)
    x = rng.uniform(0.1, 0.9, size=(5000, 3)).astype(np.float32)
    x[:100] = np.round(x[:100] * 64) / 64  # exact cell boundaries
    keys, flags = O.cell_keys(x, (0.0, 0.0, 0.0), 1.0 / 64, 64)
    g = (x - np.float32(0.0)) / np.float32(1.0 / 64)
    base = np.floor(g - np.float32(0.5)).astype(np.int64)
    np.testing.assert_array_equal(keys, (base[:, 0] * 64 + base[:, 1]) * 64 + base[:, 2])
    tkeys, _ = O.cell_keys(x, (0.0, 0.0, 0.0), 1.0 / 64, 64, tiled=True)
    blk = ((base[:, 0] >> 2) * 16 + (base[:, 1] >> 2)) * 16 + (base[:, 2] >> 2)
    loc = (base[:, 0] & 3) * 16 + (base[:, 1] & 3) * 4 + (base[:, 2] & 3)
    np.testing.assert_array_equal(tkeys, blk * 64 + loc)
    assert flags.max() == 0
    perm = O.sort_permutation(keys)
    assert np.all(np.diff(keys[perm].astype(np.int64)) >= 0)
    same = keys[perm][1:] == keys[perm][:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])


def test_excessive_time_step_rejected():
    """Reference tests/test_mpm.py:71-73."""
    with pytest.raises(ValueError, match="CFL"):
        replace(SimulationConfig(), time_step=1.0).Validate(MaterialConfig())


def test_inside_residual_shortcut_finds_no_overstress():
    """The fused substep skips the porosity prediction of an elastic particle's residual when the trial porosity t is
    in (1e-30, 0.5] and 1 - (q + a(t) p) > 1e-6 (a the f64 pressure coefficient): the prediction at float(t) then
    finds no overstress and returns pv - 1 + n0.  Pinned here on the oracle's prediction at float(t), over random
    states packed against the 1e-6 margin and across the whole range of t, the largest slope of a included."""
    rng = np.random.default_rng(17)
    checked = 0
    for k in range(4000):
        n0 = np.float32(rng.uniform(0.0, 0.3))
        t = float(10 ** rng.uniform(-12, np.log10(0.5))) if k % 4 else float(rng.uniform(0.3, 0.5))
        pv = 1.0 - float(n0) + t  # the old plastic volume whose trial porosity is t
        a = 9.0 / (4.0 * np.log(1.0 / t) ** 2 * (3.5 * t + 1.0))  # the f64 pressure coefficient (porous_return.py:11-16)
        # q + a p just inside the margin: q in [0, 1), p = (1 - q - margin) / a with the margin a little above 1e-6
        q = np.float32(rng.uniform(0.0, 0.999))
        margin = 1e-6 * (1.0 + 10 ** rng.uniform(-6, 1))
        p = np.float32(max(0.0, (1.0 - float(q) - margin) / a)) if a > 0 else np.float32(rng.uniform(0, 10))
        if not (1.0 - (float(q) + a * float(p)) > 1e-6):
            continue
        df, ms = np.float32(10 ** rng.uniform(-4, 1)), np.float32(10 ** rng.uniform(-4, 1))
        ex = np.float32(rng.uniform(-0.1, 0.1))
        pred = O.porosity_prediction(float(np.float32(t)), q, p, df, ms, ex, pv, n0)
        assert pred == pv - 1.0 + float(n0), (t, q, p, margin)
        checked += 1
    assert checked > 2500
