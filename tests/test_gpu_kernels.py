"""The unfused per-kernel C entries (one per reference Warp kernel, in the Warp kernel's argument order)
against the oracle's kernel of the same name, FP32 and FP64, on identical seeded inputs.

Inputs come from the small sphere scene after three FP64 oracle substeps (so the particles carry real
transport, gas and stress fields), with seeded perturbations: deformation gradients and affine velocities
spread around their values, stresses of a few hundred Pa, porosities, saturations, temperatures and gas
contents spread over their ranges, and part of the mesh lowered onto the tray so contact areas appear.
Each output must satisfy err(GPU, FP32 oracle) <= max(1e-5, 4 err(FP32, FP64)), errors relative to the
output's largest magnitude (the same kernel from the same inputs: only rounding and summation order
differ), and every flag must equal the FP32 oracle's.

These are the entries the reference's own tests call directly (tests/test_transport.py:195-196,
tests/test_gas.py:84-86); the fused substep runs the same arithmetic on its tiled grid.
"""
import numpy as np
import pytest

from tests.harness import O, oracle_from_spec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import scenes  # noqa: E402
from baking_amd.physics import grid as G_  # noqa: E402
from baking_amd.physics import thermal_water as TW  # noqa: E402
from baking_amd.physics import transfer as TR  # noqa: E402
from baking_amd.physics.materials import elastic as EL  # noqa: E402
from baking_amd.physics.materials import gas as GAS  # noqa: E402


class V3(tuple):
    """A host vec3 argument (wp.vec3): a list for the GPU entry, an array of the oracle's precision."""


def _run(gpu_kernel, oracle_name, n, args, outs, dim64=False):
    """Run one kernel on the GPU and on both oracles from the same inputs; return {out index: (gpu, o32, o64)}."""
    dev = []
    for a in args:
        if isinstance(a, np.ndarray):
            dev.append(torch.tensor(np.ascontiguousarray(a)).cuda())
        elif isinstance(a, V3):
            dev.append(list(a))
        else:
            dev.append(a)
    gpu_kernel(n, *dev)
    torch.cuda.synchronize()
    res = {i: [dev[i].cpu().numpy()] for i in outs}
    for precision in ("f32", "f64"):
        k = O.Kernels(precision)
        R = k.R
        conv = []
        for a in args:
            if isinstance(a, np.ndarray):
                conv.append(np.ascontiguousarray(a.astype(R) if a.dtype == np.float32 else a.copy()))
            elif isinstance(a, V3):
                conv.append(np.asarray(a, R))
            else:
                conv.append(a)
        getattr(k, oracle_name)(O.i64(n) if dim64 else n, *conv)
        for i in outs:
            res[i].append(conv[i])
    return res


def _check(name, res, names, rtol=1e-5, factor=4.0):
    for i, label in zip(sorted(res), names):
        g, a32, a64 = (np.asarray(x) for x in res[i])
        if g.dtype.kind == "i":
            np.testing.assert_array_equal(g, a32, err_msg=f"{name}: {label}")
            continue
        g, a32, a64 = (x.astype(np.float64) for x in (g, a32, a64))
        scale = max(np.abs(a64).max(initial=0.0), 1e-300)
        e, e_ref = np.abs(g - a32).max(initial=0.0) / scale, np.abs(a32 - a64).max(initial=0.0) / scale
        print(f"{name} {label}: |GPU-FP32| {e:.1e}, |FP32-FP64| {e_ref:.1e} (of max {scale:.3e})", flush=True)
        assert e <= max(rtol, factor * e_ref), (name, label, e, e_ref)


@pytest.fixture(scope="module")
def scene():
    spec = scenes.sphere_small()
    o = oracle_from_spec(spec, "f64")
    for _ in range(3):
        o.Step()
    rng = np.random.default_rng(20)
    s = o.state
    n = o.n
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    d = dict(spec=spec, o=o, n=n, rng=rng)
    d["x"] = f32(s.x)
    d["v"] = f32(s.v + rng.normal(0, 0.05, (n, 3)))
    d["C"] = f32(s.C + rng.normal(0, 2.0, (n, 3, 3)))
    d["F"] = f32(s.F + rng.normal(0, 0.03, (n, 3, 3)))
    d["tau"] = f32(rng.normal(0, 300.0, (n, 3, 3)))
    d["masses"] = f32(o.Masses())
    d["V0"] = f32(o.V0)
    d["porosity"] = f32(rng.uniform(0.02, 0.2, n))
    d["moles"] = s.ding["moles"] * rng.uniform(0.8, 1.5, n)
    d["T"] = s.transport["T"] + rng.uniform(0.0, 60.0, n)
    d["T_f32"] = f32(d["T"])
    d["W"] = s.transport["W"] * rng.uniform(0.7, 1.1, n)
    d["sat"] = rng.uniform(0.2, 0.6, n)
    d["density"] = f32(s.ding["density"])
    return d


def test_ComputeElasticStress(scene):
    n, mat = scene["n"], scene["spec"].material
    res = _run(EL.ComputeElasticStress, "ComputeElasticStress", n,
               [scene["F"], np.zeros((n, 3, 3), np.float32), float(mat.shear_modulus), float(mat.lame_lambda),
                np.zeros(2, np.int32)], outs=(1, 4))
    _check("ComputeElasticStress", res, ("tau", "flags"))


def test_CombineGasStress(scene):
    n, spec = scene["n"], scene["spec"]
    c, mat, sim = spec.ding, spec.material, spec.simulation
    res = _run(GAS.CombineGasStress, "CombineGasStress", n,
               [scene["tau"], scene["F"], scene["porosity"], scene["moles"], scene["T_f32"], np.zeros((n, 3, 3), np.float32),
                np.zeros(3, np.int32), float(c.gas_ambient_pressure_pa), float(c.porosity), scene["density"],
                float(mat.lame_lambda + 2 * mat.shear_modulus), float(sim.time_step), float(sim.grid_spacing),
                float(sim.cfl_limit)], outs=(5, 6))
    _check("CombineGasStress", res, ("total stress", "flags"))


def test_RefreshGas(scene):
    n = scene["n"]
    res = _run(GAS.RefreshGas, "RefreshGas", n,
               [scene["moles"], np.zeros(n), scene["F"], scene["porosity"], np.zeros(n), scene["T_f32"]], outs=(1, 4))
    _check("RefreshGas", res, ("moles", "pressure"))


def test_WaterStress(scene):
    n, t = scene["n"], scene["spec"].transport
    res = _run(TW.WaterStress, "WaterStress", n,
               [scene["F"], scene["sat"], scene["tau"], np.zeros((n, 3, 3), np.float32), float(t.initial_saturation),
                float(t.water_stress)], outs=(3,))
    _check("WaterStress", res, ("stress",))


def test_Saturations(scene):
    n, t = scene["n"], scene["spec"].transport
    res = _run(TW.Saturations, "Saturations", n, [scene["W"], scene["V0"], scene["F"], np.zeros(n), float(t.water_density)],
               outs=(3,))
    _check("Saturations", res, ("saturation",))


def test_LocalGas(scene):
    n, c, spec = scene["n"], scene["spec"].ding, scene["spec"]
    T_old = scene["T"]
    T_new = T_old + scene["rng"].uniform(0.0, 2.0, n)
    for producing in (1, 0):
        res = _run(TW.LocalGas, "LocalGas", n,
                   [scene["moles"], T_old, T_new, np.zeros(n), np.zeros(n), scene["F"], scene["porosity"],
                    np.array([c.gas_production_rate * 50.0], np.float32), float(spec.material.density),
                    float(c.gas_reaction_temperature_k), float(c.gas_reaction_width_k), float(spec.simulation.time_step),
                    producing], outs=(3, 4))
        _check(f"LocalGas producing={producing}", res, ("moles", "pressure"))


def _lowered_mesh(scene):
    """The sphere's reconstructed vertices lowered onto the tray, with the bottom cap (the lowest 2 cm) flattened
    into the plane 0.2 contact distances above the floor: faces there touch the tray, the rest are exposed."""
    o, spec = scene["o"], scene["spec"]
    sim = spec.simulation
    verts = np.zeros((len(o.reference_vertices), 3))
    o.reconstruct(o.state, verts)
    floor = sim.grid_origin[1] + (sim.boundary_cells - 1) * sim.grid_spacing
    verts[:, 1] += floor + 0.2 * spec.transport.contact_distance - verts[:, 1].min()
    verts[:, 1] = np.maximum(verts[:, 1], verts[:, 1].min() + 0.02)
    verts[:, 1] -= 0.02
    return np.ascontiguousarray(verts, np.float32), floor


def test_SurfaceAreas(scene):
    o, spec = scene["o"], scene["spec"]
    verts, floor = _lowered_mesh(scene)
    nf = len(o.faces)
    res = _run(TW.SurfaceAreas, "SurfaceAreas", nf,
               [verts, np.ascontiguousarray(o.faces, np.int32), np.ascontiguousarray(o.owners, np.int32), np.zeros(o.n),
                np.zeros(o.n), float(floor), float(spec.transport.contact_distance)], outs=(3, 4))
    assert res[4][1].sum() > 0, "the lowered mesh has faces on the tray"
    _check("SurfaceAreas", res, ("exposed", "contact"))


def _scalar_to_grid_args(scene):
    o, spec, n = scene["o"], scene["spec"], scene["n"]
    t, sim = spec.transport, spec.simulation
    verts, floor = _lowered_mesh(scene)
    exposed, contact = np.zeros(n), np.zeros(n)
    O.Kernels("f64").SurfaceAreas(len(o.faces), verts.astype(np.float64), o.faces, o.owners, exposed, contact, floor,
                                  t.contact_distance)
    G = o.node_count
    return [scene["x"], scene["F"], scene["V0"], np.ascontiguousarray(o.capacities), scene["T"], scene["W"], exposed, contact,
            np.array([t.heat_transfer], np.float32), float(t.oven_temperature_k), float(t.tray_temperature_k),
            float(t.tray_heat_transfer), float(t.surface_water_loss), float(t.water_density), float(t.latent_heat),
            int(t.water_enabled)] + [np.zeros(G) for _ in range(7)] + [np.zeros(n), np.zeros(n),
                                                                         V3(sim.grid_origin), float(sim.grid_spacing),
                                                                         int(sim.grid_resolution)]


def test_ScalarToGrid(scene):
    res = _run(TW.ScalarToGrid, "ScalarToGrid", scene["n"], _scalar_to_grid_args(scene), outs=tuple(range(16, 25)))
    _check("ScalarToGrid", res, ("V", "C", "E", "W", "heat", "loss", "exchange", "particle heat", "particle loss"))


@pytest.fixture(scope="module")
def grid_fields(scene):
    """The FP64 oracle's scalar P2G of the scene: the node fields GridFields, DiffuseGrid and ScalarToParticles read."""
    args = _scalar_to_grid_args(scene)
    conv = [np.asarray(a, np.float64) if isinstance(a, V3) else (a.astype(np.float64) if isinstance(a, np.ndarray) and
                                                                  a.dtype == np.float32 else a) for a in args]
    O.Kernels("f64").ScalarToGrid(scene["n"], *conv)
    names = ("V", "C", "E", "W", "H", "L", "X")
    out = {k: conv[16 + i] for i, k in enumerate(names)}
    out["p_heat"], out["p_loss"] = conv[23], conv[24]
    G = scene["o"].node_count
    out["Tn"], out["Sn"] = np.zeros(G), np.zeros(G)
    O.lib().oracle_GridFields(O.ctypes.c_int64(G), O._p(out["V"]), O._p(out["C"]), O._p(out["E"]), O._p(out["W"]),
                              O._p(out["Tn"]), O._p(out["Sn"]), O.ctypes.c_double(float(np.float32(
                                  scene["spec"].transport.water_density))))
    return out


def test_GridFields(scene, grid_fields):
    G, rw = scene["o"].node_count, float(scene["spec"].transport.water_density)
    g = grid_fields
    args = [g["V"], g["C"], g["E"], g["W"], np.zeros(G), np.zeros(G), rw]
    dev = [torch.tensor(a).cuda() if isinstance(a, np.ndarray) else a for a in args]
    TW.GridFields(G, *dev)
    torch.cuda.synchronize()
    for i, name, want in ((4, "T", g["Tn"]), (5, "s", g["Sn"])):
        got = dev[i].cpu().numpy()
        occupied = g["C"] > 0
        assert occupied.sum() > 0
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"GridFields {name}: |GPU-FP64 oracle| {err:.1e} over {int(occupied.sum())} occupied nodes", flush=True)
        assert err <= 1e-15, name  # float64 division both sides
        assert np.all(got[~occupied] == 0.0)


def test_DiffuseGrid(scene, grid_fields):
    spec, G = scene["spec"], scene["o"].node_count
    t, sim = spec.transport, spec.simulation
    g = grid_fields
    # hot spots so the fluxes are well above rounding: +40 K on a seeded tenth of the occupied nodes
    T = g["Tn"].copy()
    occ = np.flatnonzero(g["C"] > 0)
    hot = scene["rng"].choice(occ, len(occ) // 10, replace=False)
    T[hot] += 40.0
    args = [g["V"], g["C"], T, g["Sn"], g["W"], g["H"], g["L"], g["X"], np.zeros(G), np.zeros(G), np.zeros(2, np.int32),
            float(t.conductivity * 50), float(t.water_diffusivity * 1e4), float(t.water_density), int(t.water_enabled),
            float(sim.grid_spacing), int(sim.grid_resolution), float(sim.time_step)]
    res = _run(TW.DiffuseGrid, "DiffuseGrid", G, args, outs=(8, 9, 10), dim64=True)
    _check("DiffuseGrid", res, ("dT", "dW", "flags"))


def test_ScalarToParticles(scene, grid_fields):
    spec, n, o = scene["spec"], scene["n"], scene["o"]
    t, sim = spec.transport, spec.simulation
    g = grid_fields
    G = o.node_count
    rng = scene["rng"]
    occ = g["C"] > 0
    dT = np.where(occ, rng.normal(0.0, 0.5, G), 0.0)
    dW = np.where(g["W"] > 0, rng.normal(0.0, 1e-3, G) * g["W"], 0.0)
    args = [scene["x"], scene["T"], scene["W"], np.abs(rng.normal(0, 1e-9, n)), rng.normal(0, 1e-3, n), np.ascontiguousarray(o.dry),
            scene["V0"], scene["F"], g["W"], dT, dW, g["p_heat"], g["p_loss"], np.zeros(n), np.zeros(n, np.float32), np.zeros(n),
            np.zeros(n), np.zeros(n), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n), float(t.water_density),
            V3(sim.grid_origin), float(sim.grid_spacing), int(sim.grid_resolution), float(sim.time_step)]
    res = _run(TW.ScalarToParticles, "ScalarToParticles", n, args, outs=tuple(range(13, 21)))
    _check("ScalarToParticles", res, ("T", "T (material, f32)", "W", "losses", "exchanges", "masses", "reference density",
                                      "saturation"))


def _p2g_args(scene):
    sim, n = scene["spec"].simulation, scene["n"]
    G = scene["o"].node_count
    return [scene["x"], scene["v"], scene["C"], scene["tau"], scene["masses"], scene["V0"], np.zeros(G, np.float32),
            np.zeros((G, 3), np.float32), V3(sim.grid_origin), float(sim.grid_spacing), int(sim.grid_resolution),
            float(sim.time_step), np.zeros(2, np.int32)]


def test_ParticleToGrid(scene):
    res = _run(TR.ParticleToGrid, "ParticleToGrid", scene["n"], _p2g_args(scene), outs=(6, 7, 12))
    _check("ParticleToGrid", res, ("node masses", "node momenta", "flags"))
    # mass conservation (reference tests/test_mpm.py:36) on the GPU's grid
    np.testing.assert_allclose(res[6][0].sum(dtype=np.float64), scene["masses"].sum(dtype=np.float64), rtol=1e-6)


@pytest.fixture(scope="module")
def p2g_grid(scene):
    args = _p2g_args(scene)
    conv = [np.asarray(a, np.float32) if isinstance(a, V3) else a for a in args]
    O.Kernels("f32").ParticleToGrid(scene["n"], *conv)
    return conv[6], conv[7]


def test_UpdateGrid(scene, p2g_grid):
    """The GPU contracts p/m + dt g into one fused multiply-add, the oracle (built with contraction off) rounds
    twice, so nodes where p/m nearly cancels dt g differ in the last bits of the small sum: held to the FP64
    rule like every other output.  The wall projection's zeroed components must agree exactly where the
    FP64 free velocity is clearly outward."""
    sim = scene["spec"].simulation
    nm, npm = p2g_grid
    G = len(nm)
    for res_, bc in ((int(sim.grid_resolution), int(sim.boundary_cells)), (int(sim.grid_resolution), 20)):
        # bc = 20 puts every wall band through the dough, so the projection acts on occupied nodes of all six walls
        res = _run(G_.UpdateGrid, "UpdateGrid", G, [nm, npm, np.zeros((G, 3), np.float32), res_, bc,
                                                   V3(sim.gravity), float(sim.time_step)], outs=(2,), dim64=True)
        _check(f"UpdateGrid bc={bc}", res, ("node velocities",))
        g, a64 = res[2][0], res[2][2]
        zeroed = (a64 == 0.0) & (nm[:, None] > 0)
        m = nm.astype(np.float64)[:, None]
        free = np.where(m > 0, npm.astype(np.float64) / np.where(m > 0, m, 1.0), 0.0) + float(sim.time_step) * np.asarray(sim.gravity)
        clear = zeroed & (np.abs(free) > 1e-6 * np.abs(free).max())
        assert bc == int(sim.boundary_cells) or clear.sum() > 0  # the deep bands cut through the dough
        assert np.all(g[clear] == 0.0), f"bc={bc}: a clearly outward component was not removed"


def test_GridToParticle(scene, p2g_grid):
    sim, n = scene["spec"].simulation, scene["n"]
    nm, npm = p2g_grid
    G = len(nm)
    nv = np.zeros((G, 3), np.float32)
    O.Kernels("f32").UpdateGrid(O.i64(G), nm, npm, nv, int(sim.grid_resolution), int(sim.boundary_cells),
                                np.asarray(sim.gravity, np.float32), float(sim.time_step))
    args = [scene["x"], scene["F"], nv, np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32),
            np.zeros((n, 3, 3), np.float32), np.zeros((n, 3, 3), np.float32), V3(sim.grid_origin), float(sim.grid_spacing),
            int(sim.grid_resolution), float(sim.time_step), np.zeros(2, np.int32)]
    res = _run(TR.GridToParticle, "GridToParticle", n, args, outs=(3, 4, 5, 6, 11))
    _check("GridToParticle", res, ("x", "v", "C", "F", "flags"))
