"""GPU checks of the reverse-mode path (``baking_amd.adjoint``): adjoint kernels behind a checkpointed autograd Function.

Pins, each against the FP64 oracle (the same scene run in double precision):

* the reference's own gradient tests, with their tolerances:
  d mean(x_final - x0)_x / d v0 = 0.003 on the elastic demo cube
  (tests/test_warp_autodiff.py:54-60, rtol 0.01); d mean(tau_xx) / d eta
  through 5 MPM substeps of the Ding cube (tests/test_ding.py:142-170, rtol
  0.03); the same through 4 single-point return maps (tests/test_ding.py:81-113,
  rtol 0.02) -- each also against a tighter FP64 central difference;
* the reference's gas-production gradient, d mean(det F) / d alpha through
  20 substeps of the gas-producing porous cube (tests/test_gas.py:120-160,
  rtol 0.05), with the coupled porosity return and the pore-gas stress;
* the reference's transport gradients: d mean(T) / d h (heat transfer)
  through 4 substeps with heat and water transport and pore gas
  (tests/test_transport.py:96-121, rtol 3e-3), and d mean(tau_xx) / d eta
  through 5 substeps of the heated porous cube (tests/test_ding_thermal.py:
  166-191, rtol 0.03);
* the implicit-function adjoints of the overstress solve (ding.py:90-100) and
  of the coupled porosity solve (porous_return.py:64-100) against central
  differences of the oracle's FP64 overstress and of the FP64 root of its
  porosity-consistency equation (porous_return.py:34-40);
* a random linear functional of the whole final state after 20 substeps
  (checkpoints every 8, so three recompute segments) differentiated with
  respect to 32 random coordinates of the initial state, against FP64
  central differences (eps = 1e-5 of the coordinate's scale) of the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.harness import OraclePoint, oracle_from_spec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import adjoint, scenes  # noqa: E402
from baking_amd.config import DingConfig, MaterialConfig  # noqa: E402
from baking_amd.physics.materials.ding import DingMaterial  # noqa: E402
from baking_amd.physics.state import AllocateState  # noqa: E402


def _oracle_run(spec, steps, setup=None):
    o = oracle_from_spec(spec, "f64")
    if setup is not None:
        setup(o)
    for _ in range(steps):
        o.Step()
    return o


def test_velocity_gradient_closed_form():
    """tests/test_warp_autodiff.py: 3 substeps of the elastic cube at v0 = 0.2 along x; the loss moves rigidly."""
    spec = scenes.demo_autodiff(0.2)
    sim = scenes.build_simulator(spec)
    x0 = sim.current_state.particle_positions.clone()
    n = x0.shape[0]
    vp = torch.tensor([0.2], device="cuda", requires_grad=True)
    v0 = vp * torch.tensor([1.0, 0.0, 0.0], device="cuda").expand(n, 3)
    out = adjoint.rollout(sim, 3, {"v": v0})
    loss = (out["x"][:, 0] - x0[:, 0]).mean()
    loss.backward()
    grad = float(vp.grad)
    x0n = spec.samples.particle_positions.astype(np.float32).astype(np.float64)

    def loss_at(value):
        def setup(o):
            o.state.v[...] = 0.0
            o.state.v[:, 0] = value
        o = _oracle_run(spec, 3, setup)
        return float(np.mean(o.state.x[:, 0] - x0n[:, 0]))

    fd_ref = (loss_at(0.3) - loss_at(0.1)) / 0.2
    fd = (loss_at(0.2 + 1e-6) - loss_at(0.2 - 1e-6)) / 2e-6
    print(f"gradient {grad:.9g}, oracle FD (eps 0.1) {fd_ref:.9g}, oracle FD (eps 1e-6) {fd:.9g}")
    np.testing.assert_allclose(grad, fd_ref, rtol=0.01, atol=1e-6)
    np.testing.assert_allclose(grad, 0.003, rtol=0.01, atol=1e-6)
    np.testing.assert_allclose(grad, fd, rtol=1e-5)
    # the simulator is left at the final state
    assert sim.step_count == 3
    torch.testing.assert_close(sim.current_state.particle_positions, out["x"].detach())


@pytest.mark.parametrize("porosity", [0.0, 0.2])
def test_mpm_viscosity_gradient(porosity):
    """tests/test_ding.py:142-170: d mean(tau_xx) / d eta after 5 substeps of the yielding Ding cube."""
    spec = scenes.demo_viscous_gradient(1000.0, porosity=porosity)
    sim = scenes.build_simulator(spec)
    eta = torch.tensor([1000.0], device="cuda", requires_grad=True)
    out = adjoint.rollout(sim, 5, viscosity_parameter=eta)
    loss = out["tau"][:, 0, 0].mean()
    loss.backward()
    grad = float(eta.grad)

    def loss_at(value):
        def setup(o):
            o.visc_param[0] = value
        o = _oracle_run(spec, 5, setup)
        return float(np.mean(o.state.ding["tau"][:, 0, 0]))

    fd_ref = (loss_at(1010.0) - loss_at(990.0)) / 20.0
    fd = (loss_at(1000.0 + 1e-3) - loss_at(1000.0 - 1e-3)) / 2e-3
    print(f"porosity {porosity}: gradient {grad:.9g}, oracle FD (eps 10) {fd_ref:.9g}, oracle FD (eps 1e-3) {fd:.9g}")
    assert abs(fd_ref) > 1e-7
    np.testing.assert_allclose(grad, fd_ref, rtol=0.03, atol=1e-7)
    np.testing.assert_allclose(grad, fd, rtol=1e-3)


def _point_gradient(viscosity):
    """Four single-point return maps at prescribed total gradients, then the adjoint chain (tests/test_ding.py:81-100)."""
    ding = DingConfig(yield_stress=500.0, viscosity=viscosity)
    mc = MaterialConfig()
    mat = DingMaterial(mc, ding, "cuda")
    state = AllocateState(1, "cuda")
    state.deformation_gradients.copy_(torch.eye(3, device="cuda")[None])
    mat.InitializeState(state)
    states = [state]
    dt = 0.01
    for k in range(4):
        F = np.diag([1.15 + 0.01 * k, 0.93, 1.02])
        out = AllocateState(1, "cuda")
        inc = F @ np.linalg.inv(states[-1].deformation_gradients.cpu().numpy()[0].astype(np.float64))
        out.deformation_gradients.copy_(torch.as_tensor(F, dtype=torch.float32)[None])
        out.affine_velocities.copy_(torch.as_tensor((inc - np.eye(3)) / dt, dtype=torch.float32)[None])
        mat.AdvanceState(states[-1], out, dt)
        states.append(out)
    loss = float(states[-1].material_state.kirchhoff_stresses[0, 0, 0])
    z = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")  # noqa: E731
    g_Fe, g_tau, g_n = z(9), z(9), z(1)
    g_tau[0] = 1.0
    g_eta = z(1)
    for k in reversed(range(4)):
        old, new = states[k].material_state, states[k + 1]
        gFe_old, gn_old = z(9), z(1)
        adjoint.UpdateDingAdjoint(1, old.elastic_gradients, old.porosities, old.gelled, new.deformation_gradients,
                                  new.affine_velocities, old.temperatures, mat.viscosity_parameter, g_Fe, g_tau, g_n, gFe_old, gn_old,
                                  z(9), z(9), g_eta, mc.shear_modulus, mc.lame_lambda, ding.yield_stress, ding.porosity,
                                  int(ding.temperature_enabled), ding.temperature_knots_k, ding.yield_factors,
                                  ding.viscosity_factors, dt)
        g_Fe, g_tau, g_n = gFe_old, z(9), gn_old
    return loss, float(g_eta)


def test_point_viscosity_gradient():
    """tests/test_ding.py:81-113 on one particle: d tau_xx / d eta at eta = 1e10."""
    loss, grad = _point_gradient(1e10)

    def loss_at(value):
        p = OraclePoint(MaterialConfig(), DingConfig(yield_stress=500.0, viscosity=value), precision="f64")
        for k in range(4):
            s = p.step(np.diag([1.15 + 0.01 * k, 0.93, 1.02]))
        return float(s["tau"][0, 0, 0])

    fd_ref = (loss_at(1e10 + 1e8) - loss_at(1e10 - 1e8)) / 2e8
    fd = (loss_at(1e10 + 1e4) - loss_at(1e10 - 1e4)) / 2e4
    print(f"loss {loss:.6g}: gradient {grad:.9g}, oracle FD (eps 1e8) {fd_ref:.9g}, oracle FD (eps 1e4) {fd:.9g}")
    assert abs(fd_ref) > 1e-12
    np.testing.assert_allclose(grad, fd_ref, rtol=0.02, atol=1e-12)
    np.testing.assert_allclose(grad, fd, rtol=1e-3)


def test_overstress_adjoint_matches_oracle_differences():
    rng = np.random.default_rng(5)
    m = 64
    args = np.stack([rng.uniform(1.2, 500.0, m), rng.uniform(0.0, 50.0, m), 10 ** rng.uniform(-2, 1, m),
                     10 ** rng.uniform(-4, 0, m)], axis=1)
    args[:4, :2] = [[0.3, 0.2], [0.5, 0.49], [0.0, 0.9], [0.1, 0.1]]  # inside the yield surface: zero adjoint
    gbar = rng.normal(size=m)
    out = torch.zeros(m * 4, dtype=torch.float64, device="cuda")
    adjoint.OverstressAdjoint(m, torch.as_tensor(args.reshape(-1), device="cuda"), torch.as_tensor(gbar, device="cuda"), out)
    got = out.cpu().numpy().reshape(m, 4)
    solve = O.lib().oracle_overstress64
    want = np.zeros((m, 4))
    for i in range(m):
        for k in range(4):
            h = 1e-7 * max(abs(args[i, k]), 1e-3)
            ap, am = args[i].copy(), args[i].copy()
            ap[k] += h
            am[k] -= h
            want[i, k] = gbar[i] * (solve(*ap) - solve(*am)) / (2 * h)
    assert np.all(got[:4] == 0.0)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-9 * np.abs(want).max())


def _porosity_root(q, p, df, ms, ex, pv, n0):
    """FP64 root of n = prediction(n) (the oracle's porous_return.py:34-40) by 200 bisection steps."""
    lo, hi = 0.0, 1.0 - 1e-9
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid < O.porosity_prediction(mid, q, p, df, ms, ex, pv, n0):
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_solve_porosity_adjoint_matches_oracle_differences():
    rng = np.random.default_rng(3)
    rows, pvs = [], []
    while len(rows) < 24:  # interior roots with the trial state outside the yield surface
        n0 = np.float32(rng.uniform(0.05, 0.3))
        pv = 1.0 - float(n0) + rng.uniform(0.5 * n0, 1.5 * n0)
        q, p = np.float32(10 ** rng.uniform(0, 2.5)), np.float32(10 ** rng.uniform(0, 3))
        df, ms = np.float32(10 ** rng.uniform(-2, 0.5)), np.float32(10 ** rng.uniform(-2, 0.5))
        ex = np.float32(rng.uniform(-0.05, 0.05))
        if q + O.porosity_coefficient(pv - 1.0 + float(n0)) * p <= 1.0:
            continue
        if not 0.01 < _porosity_root(q, p, df, ms, ex, pv, n0) < 0.99:
            continue
        rows.append([q, p, df, ms, ex, n0])
        pvs.append(pv)
    args = np.asarray(rows, np.float32)
    pvs = np.asarray(pvs)
    m = len(rows)
    gbar = rng.normal(size=m)
    out = torch.zeros(m * 7, dtype=torch.float64, device="cuda")
    from baking_amd import _native
    lib = _native.load()
    a_d, pv_d, g_d = (torch.as_tensor(v, device="cuda") for v in (args.reshape(-1), pvs, gbar))
    _native.check(lib.bk_solve_porosity_adjoint(_native.ptr(a_d), _native.ptr(pv_d), _native.ptr(g_d), _native.ptr(out), m,
                                                _native.stream_ptr()), "bk_solve_porosity_adjoint")
    got = out.cpu().numpy().reshape(m, 7)
    want = np.zeros((m, 7))
    for i in range(m):
        for k in range(7):
            if k < 6:  # float arguments: differences between exactly representable neighbours
                x = float(args[i, k])
                up, dn = np.float32(x + 1e-4 * abs(x)), np.float32(x - 1e-4 * abs(x))
                ap, am = list(args[i]), list(args[i])
                ap[k], am[k] = up, dn
                h = float(up) - float(dn)
                fd = (_porosity_root(*ap[:5], pvs[i], ap[5]) - _porosity_root(*am[:5], pvs[i], am[5])) / h
            else:
                h = 1e-7
                fd = (_porosity_root(*args[i, :5], pvs[i] + h, args[i, 5]) - _porosity_root(*args[i, :5], pvs[i] - h, args[i, 5])) / (2 * h)
            want[i, k] = gbar[i] * fd
    err = np.abs(got - want) / (np.abs(want) + 1e-6 * np.abs(want).max())
    print("worst relative difference", err.max())
    np.testing.assert_allclose(got, want, rtol=1e-3, atol=1e-6 * np.abs(want).max())


def test_gas_production_gradient():
    """tests/test_gas.py:131-160: d mean(det F) / d alpha after 20 substeps of the gas-producing porous cube."""
    spec = scenes.demo_gas(rate=0.005)
    sim = scenes.build_simulator(spec)
    alpha = torch.tensor([0.005], device="cuda", requires_grad=True)
    out = adjoint.rollout(sim, 20, production_parameter=alpha)
    loss = torch.linalg.det(out["F"].double()).mean()
    loss.backward()
    grad = float(alpha.grad)

    def loss_at(value):
        def setup(o):
            o.prod_param[0] = value
        o = _oracle_run(spec, 20, setup)
        return float(np.linalg.det(o.state.F).mean())

    fd_ref = (loss_at(0.0051) - loss_at(0.0049)) / 0.0002
    fd = (loss_at(0.005 + 1e-7) - loss_at(0.005 - 1e-7)) / 2e-7
    print(f"gradient {grad:.9g}, oracle FD (eps 1e-4) {fd_ref:.9g}, oracle FD (eps 1e-7) {fd:.9g}")
    assert abs(fd_ref) > 1e-4
    np.testing.assert_allclose(grad, fd_ref, rtol=0.05, atol=1e-4)
    np.testing.assert_allclose(grad, fd, rtol=2e-3)


def test_heat_transfer_gradient():
    """tests/test_transport.py:96-121: d mean(T) / d h after 4 substeps with heat and water transport and pore gas."""
    spec = scenes.demo_transport(20.0)
    sim = scenes.build_simulator(spec)
    h = torch.tensor([20.0], device="cuda", requires_grad=True)
    out = adjoint.rollout(sim, 4, heat_transfer_parameter=h)
    loss = out["T"].mean()
    loss.backward()
    grad = float(h.grad)

    def loss_at(value):
        def setup(o):
            o.heat_param[0] = value
        o = _oracle_run(spec, 4, setup)
        return float(o.state.transport["T"].mean())

    fd_ref = (loss_at(22.0) - loss_at(18.0)) / 4.0
    fd = (loss_at(20.5) - loss_at(19.5)) / 1.0
    print(f"gradient {grad:.9g}, oracle FD (eps 2) {fd_ref:.9g}, oracle FD (eps 0.5) {fd:.9g}")
    assert grad > 0
    np.testing.assert_allclose(grad, fd_ref, rtol=3e-3, atol=1e-10)
    np.testing.assert_allclose(grad, fd, rtol=1e-3)


def test_heated_porous_viscosity_gradient():
    """tests/test_ding_thermal.py:166-191: d mean(tau_xx) / d eta after 5 substeps of the heated porous cube."""
    spec = scenes.demo_heated(viscosity=1000.0, yield_stress=5.0, dt=1e-3)
    sim = scenes.build_simulator(spec)
    eta = torch.tensor([1000.0], device="cuda", requires_grad=True)
    out = adjoint.rollout(sim, 5, viscosity_parameter=eta)
    loss = out["tau"][:, 0, 0].mean()
    loss.backward()
    grad = float(eta.grad)

    def loss_at(value):
        def setup(o):
            o.visc_param[0] = value
        o = _oracle_run(spec, 5, setup)
        return float(np.mean(o.state.ding["tau"][:, 0, 0]))

    fd_ref = (loss_at(1010.0) - loss_at(990.0)) / 20.0
    fd = (loss_at(1000.0 + 1e-3) - loss_at(1000.0 - 1e-3)) / 2e-3
    print(f"gradient {grad:.9g}, oracle FD (eps 10) {fd_ref:.9g}, oracle FD (eps 1e-3) {fd:.9g}")
    assert abs(fd_ref) > 1e-7
    np.testing.assert_allclose(grad, fd_ref, rtol=0.03, atol=1e-7)
    np.testing.assert_allclose(grad, fd, rtol=1e-3)


def _state_scales(fields):
    # typical magnitudes (or variations) of each field in these scenes: the random functional weights
    # each about equally, and central differences step by 1e-5 of them
    return dict(x=0.05, v=0.5, C=2.0, F=0.1, Fe=0.1, porosity=0.05, tau=50.0, moles=10.0, T=1.0, W=0.05, losses=1e-8,
                exchanges=0.1)


@pytest.mark.parametrize("case", ["neo_hookean", "ding", "ding_porous", "ding_gas", "transport", "heated"])
def test_rollout_gradient_matches_fp64_differences(case):
    """32 random initial-state coordinates of a random linear functional of the state after 20 substeps."""
    if case == "neo_hookean":
        spec = scenes.strain_rate_field(scenes.scaled(scenes.demo_autodiff(0.0), time_step=5e-4), np.diag([1.0, -0.6, 0.3]))
    elif case == "ding_gas":
        spec = scenes.strain_rate_field(scenes.demo_gas(rate=0.005), np.diag([0.5, -0.25, 0.1]))
    elif case == "transport":
        spec = scenes.strain_rate_field(scenes.demo_transport(20.0, water_stress=10000.0), np.diag([0.5, -0.25, 0.1]))
    elif case == "heated":
        spec = scenes.demo_heated(viscosity=1000.0, yield_stress=5.0, dt=5e-4)
    else:
        spec = scenes.demo_viscous_gradient(1000.0, porosity=0.2 if case == "ding_porous" else 0.0)
        spec = scenes.scaled(spec, time_step=5e-4)
    steps = 20
    sim = scenes.build_simulator(spec)
    fields = adjoint._fields_of(sim)
    from baking_amd.physics.mpm import state_field_tensors
    start = {f: t.detach().clone() for f, t in state_field_tensors(sim.current_state).items() if f in fields}
    inputs = {f: start[f].clone().requires_grad_(True) for f in fields}
    out = adjoint.rollout(sim, steps, inputs, checkpoint_every=8)
    rng = np.random.default_rng(7)
    scales = _state_scales(fields)
    weights = {f: rng.normal(size=tuple(out[f].shape)) / scales[f] for f in fields}
    loss = sum((out[f].double() * torch.as_tensor(weights[f], device="cuda")).sum() for f in fields)
    loss.backward()
    grads = {f: inputs[f].grad.double().cpu().numpy() for f in fields}

    o_fields = dict(x=("x", None), v=("v", None), C=("C", None), F=("F", None), Fe=("Fe", "ding"), porosity=("porosity", "ding"),
                    tau=("tau", "ding"), moles=("moles", "ding"), T=("T", "transport"), W=("W", "transport"),
                    losses=("losses", "transport"), exchanges=("exchanges", "transport"))
    start_np = {f: start[f].cpu().numpy().astype(np.float64) for f in fields}

    def view(o, f):
        name, part = o_fields[f]
        return getattr(o.state, part)[name] if part else getattr(o.state, name)

    def loss_at(field, index, delta):
        def setup(o):
            for f in fields:
                view(o, f)[...] = start_np[f]
            view(o, field)[index] += delta
        o = _oracle_run(spec, steps, setup)
        return sum(float((view(o, f) * weights[f]).sum()) for f in fields)

    picks = []
    n = start["x"].shape[0]
    # with n0 = 0 a zero porosity stays 0 (ding.py:207-209): no neighbourhood to difference in
    pickable = [f for f in fields if not (f == "porosity" and spec.ding.porosity == 0.0)]
    for _ in range(32):
        f = pickable[rng.integers(len(pickable))]
        idx = (int(rng.integers(n)),) + tuple(int(rng.integers(s)) for s in start[f].shape[1:])
        picks.append((f, idx))
    rows = []
    for f, idx in picks:
        h = 1e-5 * scales[f]
        fd = (loss_at(f, idx, h) - loss_at(f, idx, -h)) / (2 * h)
        rows.append((f, idx, grads[f][idx], fd))
    # The adjoint differentiates the float32 trajectory, the differences the FP64 one: agreement to ~1e-4
    # of a field's gradient scale is that precision.  A coordinate whose gradient is far below its
    # field's scale is therefore compared with a floor of 0.1 of the field's largest picked gradient;
    # a missing or wrong term shows up as O(1) relative errors on many coordinates.
    scale = {f: max(abs(r[3]) for r in rows if r[0] == f) for f in {r[0] for r in rows}}
    rel = [abs(r[2] - r[3]) / (abs(r[3]) + 0.1 * scale[r[0]]) for r in rows]
    plain = [abs(r[2] - r[3]) / max(abs(r[3]), 1e-300) for r in rows]
    for r in rows:
        print(f"{r[0]:9s} {str(r[1]):14s} adjoint {r[2]: .8e}  FD {r[3]: .8e}")
    print(f"{case}: worst relative difference {max(rel):.2e} (floor 0.1 of the field's largest gradient), "
          f"median plain relative difference {np.median(plain):.2e}")
    assert max(rel) < 2e-3
    assert np.median(plain) < 1e-4
