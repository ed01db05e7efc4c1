"""CPU oracle driver for the MLS-MPM baking substep.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker (or the timed CPU
baseline).  The product path (``baking_amd``) never imports it.

It restates, in numpy + the serial C++ kernels of ``baking_oracle.cpp``:

* ``MpmSimulator.Step`` orchestration (reference ``physics/mpm.py:163-225``),
* ``ThermalWaterTransport.Advance/AddWaterStress/FinishStep``
  (``physics/thermal_water.py:296-367``),
* ``DingMaterial.InitializeState/ComputeStress/AdvanceState``
  (``physics/materials/ding.py:243-327``),
* the host-side ``BoundaryFaces`` (``physics/thermal_water.py:51-74``) and the
  vertex-particle binding (``geometry/mesh_tracking.py:41-75``).

``precision`` is ``"f32"`` (the reference's float32 state) or ``"f64"`` (every
reference ``float`` promoted to double: the FP64 run of the same kernels).

Parity pins: every function here is checked against the reference tests'
closed forms and invariants in ``tests/test_oracle_*.py`` (free fall, free
translation, Hencky elastic limit, stationarity, ideal-gas scalings, the
near-closure porosity known answer, insulated/conservation balances, cosine
heat mode orders, boundary-face areas, affine mesh reconstruction), and the
SVD against LAPACK.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_TIMING_DIR = _HERE / "_timing"
GAS_CONSTANT = 8.31446261815324  # physics/materials/gas.py:8

_lib = None


def build(force=False):
    """Compile ``liboracle.so`` from ``baking_oracle.cpp`` with g++ (serial, IEEE, no FMA contraction)."""
    src = _HERE / "baking_oracle.cpp"
    if _LIB_PATH.exists() and not force and _LIB_PATH.stat().st_mtime >= src.stat().st_mtime:
        return _LIB_PATH
    import subprocess
    tmp = _LIB_PATH.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           str(src), "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def build_timing(openmp, march="native"):
    """Compile a timing build of the CPU baseline into ``oracle/_timing/``: the same source at -O3 -march=<march>,
    contraction still off -- serial (``liboracle_o3.so``, the checker's loops) or with OpenMP over each kernel's
    loop and atomic scatters (``liboracle_omp.so``).  Built on the host that runs it (the machine code is specific
    to that CPU); returns its path."""
    import subprocess
    src = _HERE / "baking_oracle.cpp"
    _TIMING_DIR.mkdir(exist_ok=True)
    path = _TIMING_DIR / ("liboracle_omp.so" if openmp else "liboracle_o3.so")
    tmp = path.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", f"-march={march}", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"]
    cmd += ["-fopenmp", "-DORACLE_OPENMP"] if openmp else []
    subprocess.run(cmd + [str(src), "-o", str(tmp)], check=True)
    os.replace(tmp, path)
    return path


_libs = {}


def _configure(handle):
    handle.oracle_porosity_coefficient.restype = ctypes.c_float
    handle.oracle_porosity_coefficient.argtypes = [ctypes.c_float]
    handle.oracle_solve_overstress.restype = ctypes.c_float
    handle.oracle_solve_overstress.argtypes = [ctypes.c_float] * 4
    handle.oracle_overstress64.restype = ctypes.c_double
    handle.oracle_overstress64.argtypes = [ctypes.c_double] * 4
    handle.oracle_solve_porosity.restype = ctypes.c_float
    handle.oracle_solve_porosity.argtypes = [ctypes.c_float] * 5 + [ctypes.c_double, ctypes.c_float]
    handle.oracle_porosity_prediction.restype = ctypes.c_double
    handle.oracle_porosity_prediction.argtypes = [ctypes.c_double] + [ctypes.c_float] * 5 + [ctypes.c_double, ctypes.c_float]
    handle.oracle_gas_pressure.restype = ctypes.c_double
    handle.oracle_gas_pressure.argtypes = [ctypes.c_double, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    handle.oracle_temperature_factor.restype = ctypes.c_float
    handle.oracle_temperature_factor.argtypes = [ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
    handle.oracle_set_threads.argtypes = [ctypes.c_int]
    handle.oracle_threads.restype = ctypes.c_int
    return handle


def lib(path=None):
    """The checker library (serial, ``liboracle.so``), or the library at ``path`` (the timing build)."""
    global _lib
    if path is None:
        if _lib is None:
            if not _LIB_PATH.exists():
                build()
            _lib = _configure(ctypes.CDLL(str(_LIB_PATH)))
        return _lib
    key = str(Path(path).resolve())
    if key not in _libs:
        _libs[key] = _configure(ctypes.CDLL(key))
    return _libs[key]


# ------------------------------------------------------------------ argument marshalling
def _p(a):
    if a is None:
        return ctypes.c_void_p(0)
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return ctypes.c_void_p(a.ctypes.data)


class Kernels:
    """Typed wrappers around the serial C++ kernels for one precision."""

    def __init__(self, precision="f32", library=None):
        assert precision in ("f32", "f64")
        self.precision = precision
        self.R = np.float32 if precision == "f32" else np.float64
        self._c = ctypes.c_float if precision == "f32" else ctypes.c_double
        self._lib = lib(library)

    def _call(self, name, *args):
        symbol = f"oracle_{name}_{self.precision}"
        fn = getattr(self._lib, symbol)
        types = _signatures()[symbol]
        if len(types) != len(args):
            raise TypeError(f"{symbol} takes {len(types)} arguments, got {len(args)}")
        conv = []
        for t, a in zip(types, args):
            if t is ctypes.c_void_p:
                if not (isinstance(a, np.ndarray) or a is None):
                    raise TypeError(f"{symbol}: expected an array, got {type(a).__name__}")
                conv.append(_p(a))
            elif t is ctypes.c_int64:
                conv.append(ctypes.c_int64(a.v if isinstance(a, _I64) else int(a)))
            elif t is ctypes.c_int:
                conv.append(ctypes.c_int(int(a)))
            else:
                conv.append(t(float(a)))
        fn.argtypes = types
        fn.restype = None
        fn(*conv)

    def arr(self, a, dtype=None):
        return np.ascontiguousarray(a, dtype=self.R if dtype is None else dtype)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return lambda *args: self._call(name, *args)


_SIGS = None


def _signatures():
    """ctypes argtypes of every exported kernel, parsed from the ORACLE_BOTH declarations of the C++ source."""
    global _SIGS
    if _SIGS is not None:
        return _SIGS
    import re
    text = (_HERE / "baking_oracle.cpp").read_text()
    sigs = {}
    for m in re.finditer(r"ORACLE_BOTH\((\w+),\s*\(([^)]*)\),\s*\(([^)]*)\),", text):
        name = m.group(1)
        for suffix, params in (("f32", m.group(2)), ("f64", m.group(3))):
            types = []
            for p in params.split(","):
                p = p.strip()
                if "*" in p:
                    types.append(ctypes.c_void_p)
                elif p.startswith("int64_t"):
                    types.append(ctypes.c_int64)
                elif p.startswith("int"):
                    types.append(ctypes.c_int)
                elif p.startswith("float"):
                    types.append(ctypes.c_float)
                elif p.startswith("double"):
                    types.append(ctypes.c_double)
                else:
                    raise ValueError(f"unparsed parameter '{p}' of {name}")
            sigs[f"oracle_{name}_{suffix}"] = types
    _SIGS = sigs
    return sigs


class _I64:
    def __init__(self, v):
        self.v = int(v)


def i64(v):
    return _I64(v)


# ------------------------------------------------------------------ scalar helpers (KATs)
def porosity_coefficient(n):
    return float(lib().oracle_porosity_coefficient(float(n)))


def solve_overstress(q2, p2, df, mf):
    return float(lib().oracle_solve_overstress(q2, p2, df, mf))


def solve_porosity(q, p, df, ms, ex, pv, n0):
    return float(lib().oracle_solve_porosity(q, p, df, ms, ex, pv, n0))


def porosity_prediction(n, q, p, df, ms, ex, pv, n0):
    return float(lib().oracle_porosity_prediction(n, q, p, df, ms, ex, pv, n0))


def gas_pressure(moles, T, J, n):
    return float(lib().oracle_gas_pressure(moles, T, J, n))


def temperature_factor(t, knots, factors):
    k = np.ascontiguousarray(knots, dtype=np.float32)
    f = np.ascontiguousarray(factors, dtype=np.float32)
    return float(lib().oracle_temperature_factor(float(t), _p(k), _p(f)))


def svd3(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(3, 3)
    U = np.zeros((3, 3)); S = np.zeros(3); V = np.zeros((3, 3))
    lib().oracle_svd3(_p(F), _p(U), _p(S), _p(V))
    return U, S, V


def cell_keys(positions, origin, dx, res, tiled=False, precision="f32"):
    k = Kernels(precision)
    x = k.arr(positions)
    keys = np.zeros(len(x), dtype=np.uint32)
    flags = np.zeros(2, dtype=np.int32)
    k.CellKeys(len(x), x, k.arr(origin), dx, res, int(tiled), keys, flags)
    return keys, flags


def sort_permutation(keys):
    """The reference-grade stable sort: numpy's stable argsort of the keys."""
    return np.argsort(keys, kind="stable").astype(np.int32)


# ------------------------------------------------------------------ host setup restatements
def boundary_faces(vertex_positions, tet_indices, source_tets, particle_positions):
    """Outward boundary triangles and their owner particle (physics/thermal_water.py:51-74)."""
    seen = {}
    for t, tet in enumerate(tet_indices):
        for opp in range(4):
            face = np.delete(tet, opp).copy()
            key = tuple(sorted(face))
            if key in seen:
                seen[key] = None
                continue
            pts = vertex_positions[face]
            if np.dot(np.cross(pts[1] - pts[0], pts[2] - pts[0]), vertex_positions[tet[opp]] - pts[0]) > 0:
                face[[1, 2]] = face[[2, 1]]
            seen[key] = (face, t)
    faces, owners = [], []
    for entry in seen.values():
        if entry is None:
            continue
        face, t = entry
        members = np.flatnonzero(source_tets == t)
        centre = vertex_positions[face].mean(axis=0)
        owners.append(members[np.argmin(np.linalg.norm(particle_positions[members] - centre, axis=1))])
        faces.append(face)
    return np.asarray(faces, dtype=np.int32).reshape(-1, 3), np.asarray(owners, dtype=np.int32)


def sample_particles(vertex_positions, tet_indices, reference_volumes, density, particles_per_tet=0, target_volume=None,
                     random_seed=42):
    """Tet-by-tet particle seeding (geometry/particle_sampling.py:13-53): symmetric 4-point rule, centroid,
    or seeded Dirichlet weights drawn per tet; returns (positions, volumes, masses, source tets)."""
    if particles_per_tet == 0:
        counts = np.maximum(4.0, np.ceil(reference_volumes / target_volume)).astype(np.int64)
    else:
        counts = np.full(len(tet_indices), particles_per_tet, dtype=np.int64)
    source = np.repeat(np.arange(len(counts)), counts)
    positions = np.empty((len(source), 3))
    volumes = np.repeat(reference_volumes / counts, counts)
    rng = np.random.default_rng(random_seed)
    sym = np.full((4, 4), (5.0 - np.sqrt(5.0)) / 20.0)
    np.fill_diagonal(sym, (5.0 + 3.0 * np.sqrt(5.0)) / 20.0)
    start = 0
    for t, k in enumerate(counts):
        w = sym if k == 4 else (np.full((1, 4), 0.25) if k == 1 else rng.dirichlet(np.ones(4), size=k))
        positions[start:start + k] = w @ vertex_positions[tet_indices[t]]
        start += k
    return positions, volumes, density * volumes, source


def mesh_binding(vertex_positions, tet_indices, source_tets, particle_positions, reference_volumes):
    """Vertex -> particle CSR with normalised volume weights (geometry/mesh_tracking.py:42-75)."""
    n = len(particle_positions)
    nv = len(vertex_positions)
    bv = tet_indices[source_tets].reshape(-1)
    bp = np.repeat(np.arange(n, dtype=np.int32), 4)
    order = np.argsort(bv, kind="stable")
    bv, bp = bv[order], bp[order]
    counts = np.bincount(bv, minlength=nv)
    offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    w = np.asarray(reference_volumes, dtype=np.float64)[bp]
    vw = np.bincount(bv, weights=w, minlength=nv)
    w = w / vw[bv]
    offs = vertex_positions[bv] - particle_positions[bp]
    return dict(offsets=offsets, particles=bp.astype(np.int32), weights=w, reference_offsets=offs,
                unmapped=int(np.count_nonzero(counts == 0)))


# ------------------------------------------------------------------ state
@dataclass
class OracleState:
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    ding: dict | None = None
    transport: dict | None = None

    def copy(self):
        return OracleState(self.x.copy(), self.v.copy(), self.F.copy(), self.C.copy(),
                           None if self.ding is None else {k: v.copy() for k, v in self.ding.items()},
                           None if self.transport is None else {k: v.copy() for k, v in self.transport.items()})


class OracleSimulator:
    """Serial CPU restatement of ``MpmSimulator`` (forward mode) for parity checks.

    Arguments follow the reference constructor (``physics/mpm.py:24-25``); the
    configs are duck-typed (attributes only).
    """

    def __init__(self, tet_mesh, particle_samples, simulation_config, material_config, initial_velocity=(0.0, 0.0, 0.0),
                 ding_config=None, transport_config=None, precision="f32", library=None, sticky_floor=False):
        self.k = Kernels(precision, library)
        # no reference counterpart: after UpdateGrid's walls, velocity 0 on every node of the floor's layers
        self.sticky_floor = bool(sticky_floor)
        R = self.k.R
        self.R = R
        self.sim = simulation_config
        self.mat = material_config
        self.ding_cfg = ding_config if material_config.model == "ding2019" else None
        if self.ding_cfg is None and material_config.model == "ding2019":
            raise ValueError("ding2019 needs a ding_config")
        self.tr = transport_config if (transport_config is not None and transport_config.enabled) else None
        self.mesh = tet_mesh
        self.samples = particle_samples
        self.n = len(particle_samples.particle_positions)
        self.res = int(simulation_config.grid_resolution)
        self.node_count = self.res ** 3
        self.origin = np.asarray(simulation_config.grid_origin, dtype=R)
        self.gravity = np.asarray(simulation_config.gravity, dtype=R)
        self.dx = float(simulation_config.grid_spacing)
        self.dt = float(simulation_config.time_step)
        self.V0 = np.ascontiguousarray(particle_samples.reference_volumes, dtype=R)
        self.m0 = np.ascontiguousarray(particle_samples.particle_masses, dtype=R)
        b = mesh_binding(tet_mesh.vertex_positions, tet_mesh.tet_indices, particle_samples.source_tets,
                         particle_samples.particle_positions, particle_samples.reference_volumes)
        self.binding = dict(offsets=b["offsets"], particles=b["particles"], weights=b["weights"].astype(R),
                            reference_offsets=np.ascontiguousarray(b["reference_offsets"], dtype=R))
        self.unmapped = b["unmapped"]
        self.reference_vertices = np.ascontiguousarray(tet_mesh.vertex_positions, dtype=R)
        self.initial_velocity = np.asarray(initial_velocity, dtype=np.float32)
        if self.ding_cfg is not None:
            c = self.ding_cfg
            self.visc_param = np.array([c.viscosity], dtype=R)
            self.prod_param = np.array([c.gas_production_rate], dtype=R)
            self.knots = np.asarray(c.temperature_knots_k, dtype=R)
            self.yf = np.asarray(c.yield_factors, dtype=R)
            self.vf = np.asarray(c.viscosity_factors, dtype=R)
            self.initial_temperature_k = self.tr.initial_temperature_k if self.tr is not None else 293.15
        if self.tr is not None:
            t = self.tr
            vols = np.asarray(particle_samples.reference_volumes, dtype=np.float64)
            self.dry = (material_config.density - t.water_density * t.initial_saturation) * vols
            self.dry = np.ascontiguousarray(self.dry, dtype=np.float64)
            self.capacities = np.ascontiguousarray(material_config.density * t.specific_heat * vols, dtype=np.float64)
            self.heat_param = np.array([t.heat_transfer], dtype=R)
            faces, owners = boundary_faces(tet_mesh.vertex_positions, tet_mesh.tet_indices, particle_samples.source_tets,
                                           particle_samples.particle_positions)
            self.faces, self.owners = faces, owners
        self.Reset()

    # -------------------------------------------------------------- init (mpm.py:129-161)
    def _ding_init(self, n):
        c = self.ding_cfg
        R = self.R
        T0 = self.initial_temperature_k
        eye = np.tile(np.eye(3, dtype=R), (n, 1, 1))
        gel = int(c.temperature_enabled and T0 >= c.temperature_knots_k[3])
        yf = float(np.interp(T0, c.temperature_knots_k, c.yield_factors)) if c.temperature_enabled else 1.0
        vf = float(np.interp(T0, c.temperature_knots_k, c.viscosity_factors)) if c.temperature_enabled else 1.0
        moles = c.gas_initial_pressure_pa * c.porosity / (GAS_CONSTANT * T0) if c.gas_enabled else 0.0
        return dict(Fe=eye.copy(), Fp=eye.copy(), tau=np.zeros((n, 3, 3), R), residual=np.zeros(n, R),
                    porosity=np.full(n, c.porosity, R), T=np.full(n, T0, R), gelled=np.full(n, gel, np.int32),
                    yield_stress=np.full(n, c.yield_stress * yf, R), viscosity=np.full(n, float(self.visc_param[0]) * vf, R),
                    moles=np.full(n, moles, np.float64), pressure=np.full(n, c.gas_initial_pressure_pa if c.gas_enabled else 0.0),
                    porosity_residual=np.zeros(n, R), density=np.full(n, self.mat.density, R))

    def Reset(self):
        R = self.R
        n = self.n
        x = np.ascontiguousarray(self.samples.particle_positions.astype(np.float32), dtype=R)
        v = np.ascontiguousarray(np.tile(self.initial_velocity, (n, 1)), dtype=R)
        F = np.tile(np.eye(3, dtype=R), (n, 1, 1))
        C = np.zeros((n, 3, 3), R)
        self.state = OracleState(x, v, F, C)
        if self.ding_cfg is not None:
            self.state.ding = self._ding_init(n)
        if self.tr is not None:
            t = self.tr
            vols = np.asarray(self.samples.reference_volumes, dtype=np.float64)
            self.state.transport = dict(T=np.full(n, t.initial_temperature_k), W=t.water_density * t.initial_saturation * vols,
                                        s=np.full(n, t.initial_saturation), losses=np.zeros(n), exchanges=np.zeros(n),
                                        masses=self.m0.copy())
        self.flags = np.zeros(2, np.int32)
        self.ding_flags = np.zeros(3, np.int32)
        self.transport_flags = np.zeros(2, np.int32)
        self.step_count = 0
        # settling controller (physics/mpm.py:150-157)
        self._baking_start_step = None if self.tr is not None and self.tr.settling_duration_s > 0 else 0
        self.settling_quiet_steps = 0
        self.settling_reference_step = 0
        self.settling_plastic_rate = 0.0
        self.settling_reference_plastic = self.state.ding["Fp"].copy() if self.ding_cfg is not None else None
        self.settling_reason = "pending" if self._baking_start_step is None else "disabled"
        self.settling_diagnostics = dict(max_speed_m_s=0.0, max_strain_rate_per_s=0.0, max_plastic_rate_per_s=0.0)
        self.node_m = np.zeros(self.node_count, R)
        self.node_p = np.zeros((self.node_count, 3), R)
        self.node_v = np.zeros((self.node_count, 3), R)
        self.vertices = self.reference_vertices.copy()

    # -------------------------------------------------------------- settling (mpm.py:57-109)
    @property
    def settling_steps(self):
        if self._baking_start_step is not None and self._baking_start_step > 0:
            return self._baking_start_step
        if self.tr is None:
            return 0
        return int(math.ceil(self.tr.settling_duration_s / self.dt - 1e-10))

    @property
    def IsSettling(self):
        return self.tr is not None and self.tr.settling_duration_s > 0 and self._baking_start_step is None

    @property
    def baking_steps(self):
        return 0 if self.IsSettling else max(0, self.step_count - self.settling_steps)

    @property
    def baking_time(self):
        return self.baking_steps * self.dt

    def update_settling(self):
        """physics/mpm.py:77-109 after an accepted resting substep.

        The tolerance comparisons take the float32 metrics against the configuration's floats in
        double precision (NumPy 1.x promotion; under NumPy 2 the reference compares in float32,
        which differs only for a metric equal to a tolerance rounded up to float32)."""
        t = self.tr
        dt = self.dt
        if not t.settling_auto:
            if self.step_count >= self.settling_steps:
                self._baking_start_step = self.step_count
                self.settling_reason = "fixed_duration"
            return
        values = np.zeros(3, self.R)  # the kernel's precision: float32 as the reference, float64 for the FP64 run
        elapsed = (self.step_count - self.settling_reference_step) * dt
        self.k.SettlingMetrics(self.n, self.state.v, self.state.C, self.settling_reference_plastic, self.state.ding["Fp"],
                               values, elapsed)
        if elapsed + 1e-12 >= min(.01, t.settling_hold_duration_s):
            self.settling_plastic_rate = float(values[2])
            self.settling_reference_plastic = self.state.ding["Fp"].copy()
            self.settling_reference_step = self.step_count
        values[2] = self.settling_plastic_rate
        self.settling_diagnostics = dict(max_speed_m_s=float(values[0]), max_strain_rate_per_s=float(values[1]),
                                         max_plastic_rate_per_s=float(values[2]))
        eligible = (self.step_count - 1) * dt + 1e-12 >= t.settling_min_duration_s
        small = float(values[0]) <= t.settling_speed_tolerance and float(max(values[1:])) <= t.settling_rate_tolerance
        self.settling_quiet_steps = self.settling_quiet_steps + 1 if eligible and small else 0
        if self.settling_quiet_steps * dt + 1e-12 >= t.settling_hold_duration_s:
            self._baking_start_step = self.step_count
            self.settling_reason = "quasi_static"
        elif self.step_count >= self.settling_steps:
            self.settling_reason = "timeout"
            raise RuntimeError("Settling did not reach quasi-static tolerances before settling_duration_s; "
                               "baking was not started. Increase the maximum duration or review tolerances/material.")

    # -------------------------------------------------------------- CheckState (mpm.py:243-308)
    def check_state(self):
        """The reference's CheckState on the oracle state with numpy: raises its errors in its order and returns
        its statistics.  GasWaveSpeed follows physics/materials/ding.py:285-292."""
        sim, mat = self.sim, self.mat
        if self.flags[0]:
            raise RuntimeError("nonpositive/nonfinite J; reduce dt or inspect loading and sampling")
        if self.flags[1]:
            raise RuntimeError("particle stencil left the grid; enlarge the domain or reduce dt")
        s = self.state
        x, v, F, C = (np.asarray(a, np.float32) for a in (s.x, s.v, s.F, s.C))
        if not all(np.isfinite(a).all() for a in (x, v, F, C)):
            raise RuntimeError("nonfinite particle state; simulation stopped")
        J = np.linalg.det(F)
        if J.min() <= 0.0:
            raise RuntimeError("inverted particle deformation; reduce dt or loading")
        g = (x - np.array(sim.grid_origin)) / sim.grid_spacing
        if np.any(g < 0.5) or np.any(g >= sim.grid_resolution - 1.5):
            raise RuntimeError("particle outside valid grid stencil")
        masses = np.asarray(self.Masses(), np.float32).astype(np.float64)
        vmax = float(np.linalg.norm(v, axis=1).max())
        wave = mat.wave_speed
        d = s.ding
        if self.ding_cfg is not None and self.ding_cfg.gas_enabled:
            n = np.asarray(d["porosity"], np.float32).astype(np.float64)
            bulk = d["pressure"] * (1 + (1 - self.ding_cfg.porosity + n) / n)
            modulus = mat.lame_lambda + 2 * mat.shear_modulus
            Jd = np.linalg.det(F.astype(np.float64))
            with np.errstate(divide="ignore", invalid="ignore"):
                gas = float(np.sqrt((modulus + bulk) * Jd / np.asarray(d["density"], np.float32)).max())
            wave = max(wave, gas)
        if self.tr is not None and self.tr.water_enabled:
            water = self.tr.water_stress * s.transport["s"] * J / np.asarray(d["density"], np.float32)
            wave = float(np.sqrt(wave * wave + water.max()))
        cfl = (vmax + wave) * sim.time_step / sim.grid_spacing
        if cfl > sim.cfl_limit:
            raise RuntimeError(f"estimated CFL={cfl:.3f} > {sim.cfl_limit}; reduce dt")
        if self.ding_cfg is not None:
            self.check_ding()
            if not np.isfinite(d["Fe"]).all() or not np.isfinite(d["Fp"]).all():
                raise RuntimeError("nonfinite Ding internal state")
            if float(np.max(d["residual"])) > 1e-4:
                raise RuntimeError("Ding local return mapping did not converge")
        verts = np.zeros((len(self.reference_vertices), 3), self.R)
        self.reconstruct(s, verts)
        self.vertices = verts
        verts = verts.astype(np.float32)
        if not np.isfinite(verts).all():
            raise RuntimeError("nonfinite reconstructed mesh vertices")
        corners = verts[self.mesh.tet_indices]
        signed = np.linalg.det(corners[:, 1:] - corners[:, :1]) / 6.0
        return {
            "step_count": self.step_count,
            "simulation_time": self.step_count * sim.time_step,
            "particle_count": self.n,
            "particle_volume_m3": float(np.dot(np.asarray(self.samples.reference_volumes, np.float64), J)),
            "total_mass": float(masses.sum()),
            "grid_mass": float(np.asarray(self.node_m, np.float32).sum(dtype=np.float64)),
            "minimum_j": float(J.min()),
            "maximum_j": float(J.max()),
            "maximum_speed": vmax,
            "kinetic_energy": float(0.5 * np.sum(masses * np.sum(v ** 2, axis=1))),
            "estimated_cfl": cfl,
            "unmapped_mesh_vertices": self.unmapped,
            "inverted_mesh_tets": int(np.count_nonzero(signed <= 0.0)),
        }

    def Masses(self, state=None):
        s = self.state if state is None else state
        return s.transport["masses"] if s.transport is not None else self.m0

    # -------------------------------------------------------------- pieces
    def reconstruct(self, state, out):
        b = self.binding
        self.k.ReconstructMeshVertices(len(out), state.x, state.F, self.reference_vertices, b["offsets"], b["particles"],
                                       b["weights"], b["reference_offsets"], out)

    def transport_advance(self, state, settling=False, dt=None):
        """thermal_water.py:296-355; returns the prepared state and the workspace."""
        k, t, n, G = self.k, self.tr, self.n, self.node_count
        dt = self.dt if dt is None else dt
        sim = self.sim
        R = self.R
        heat_param = np.zeros(1, R) if settling else self.heat_param
        verts = np.zeros((len(self.reference_vertices), 3), R)
        self.reconstruct(state, verts)
        exposed = np.zeros(n)
        contact = np.zeros(n)
        floor = sim.grid_origin[1] + (sim.boundary_cells - 1) * sim.grid_spacing
        k.SurfaceAreas(len(self.faces), verts, self.faces, self.owners, exposed, contact, floor, t.contact_distance)
        w = {name: np.zeros(G) for name in ("V", "C", "E", "W", "H", "L", "X", "T", "S", "dT", "dW")}
        p_heat, p_loss = np.zeros(n), np.zeros(n)
        old = state.transport
        out = dict(T=np.zeros(n), W=np.zeros(n), s=np.zeros(n), losses=np.zeros(n), exchanges=np.zeros(n), masses=np.zeros(n, R))
        T_mat = np.zeros(n, R)
        rho_ref = np.zeros(n, R)
        k.ScalarToGrid(n, state.x, state.F, self.V0, self.capacities, old["T"], old["W"], exposed, contact, heat_param,
                       t.oven_temperature_k, t.tray_temperature_k, 0.0 if settling else t.tray_heat_transfer,
                       0.0 if settling else t.surface_water_loss, t.water_density, t.latent_heat, int(t.water_enabled),
                       w["V"], w["C"], w["E"], w["W"], w["H"], w["L"], w["X"], p_heat, p_loss, self.origin, self.dx, self.res)
        self.k._lib.oracle_GridFields(ctypes.c_int64(G), _p(w["V"]), _p(w["C"]), _p(w["E"]), _p(w["W"]), _p(w["T"]), _p(w["S"]),
                                ctypes.c_double(float(R(t.water_density))))
        k.DiffuseGrid(i64(G), w["V"], w["C"], w["T"], w["S"], w["W"], w["H"], w["L"], w["X"], w["dT"], w["dW"], self.transport_flags,
                      0.0 if settling else t.conductivity, 0.0 if settling else t.water_diffusivity, t.water_density,
                      int(t.water_enabled), self.dx, self.res, dt)
        if self.transport_flags[0] or self.transport_flags[1]:
            raise RuntimeError('Heat/water explicit stability bound exceeded; reduce time_step; no temperature or water clipping is applied')
        k.ScalarToParticles(n, state.x, old["T"], old["W"], old["losses"], old["exchanges"], self.dry, self.V0, state.F, w["W"],
                            w["dT"], w["dW"], p_heat, p_loss, out["T"], T_mat, out["W"], out["losses"], out["exchanges"],
                            out["masses"], rho_ref, out["s"], t.water_density, self.origin, self.dx, self.res, dt)
        ding = dict(state.ding)
        ding["T"] = T_mat
        ding["density"] = rho_ref
        c = self.ding_cfg
        if c.gas_enabled:
            ding["moles"] = np.zeros(n)
            ding["pressure"] = np.zeros(n)
            k.LocalGas(n, state.ding["moles"], old["T"], out["T"], ding["moles"], ding["pressure"], state.F,
                       state.ding["porosity"], self.prod_param, self.mat.density, c.gas_reaction_temperature_k,
                       c.gas_reaction_width_k, dt, int(not settling))
        prepared = OracleState(state.x, state.v, state.F, state.C, ding, out)
        work = dict(w, exposed=exposed, contact=contact, vertices=verts, p_heat=p_heat, p_loss=p_loss)
        return prepared, work

    def compute_stress(self, state, tau_out):
        k = self.k
        if self.ding_cfg is None:
            k.ComputeElasticStress(self.n, state.F, tau_out, self.mat.shear_modulus, self.mat.lame_lambda, self.flags)
            return
        c = self.ding_cfg
        if not c.gas_enabled:
            tau_out[...] = state.ding["tau"]
            return
        sim = self.sim
        k.CombineGasStress(self.n, state.ding["tau"], state.F, state.ding["porosity"], state.ding["moles"], state.ding["T"],
                           tau_out, self.ding_flags, c.gas_ambient_pressure_pa, c.porosity, state.ding["density"],
                           self.mat.lame_lambda + 2 * self.mat.shear_modulus, sim.time_step, sim.grid_spacing, sim.cfl_limit)
        self.check_ding()

    def check_ding(self):
        f = self.ding_flags
        if f[2]:
            raise RuntimeError("Gas acoustic CFL exceeds the explicit limit; reduce time_step. Pressure is not capped")
        if f[1]:
            raise RuntimeError("Coupled porosity return has no resolved interior solution; reduce dt/loading. No pore-volume clipping or closed-pore law is applied")
        if f[0]:
            raise RuntimeError("Ding porosity left [0, 1); no constitutive clamping is applied. Reduce dt or loading; the paper gives no closure beyond this domain")

    def advance_ding(self, inp, out):
        """ding.py:303-327 (UpdateDing, density copy, RefreshGas)."""
        k, c, n = self.k, self.ding_cfg, self.n
        if out.ding is None:
            out.ding = self._ding_init(n)
        od, nd = inp.ding, out.ding
        k.UpdateDing(n, od["Fe"], od["porosity"], od["gelled"], out.F, out.C, nd["Fe"], nd["Fp"], nd["tau"], nd["residual"],
                     nd["porosity"], nd["T"], nd["gelled"], nd["yield_stress"], nd["viscosity"], self.visc_param, self.ding_flags,
                     nd["porosity_residual"], int(c.gas_enabled), self.mat.shear_modulus, self.mat.lame_lambda, c.yield_stress,
                     c.porosity, int(c.temperature_enabled), self.knots, self.yf, self.vf, self.dt, od["T"])
        nd["density"][...] = od["density"]
        if c.gas_enabled:
            k.RefreshGas(n, od["moles"], nd["moles"], out.F, nd["porosity"], nd["pressure"], nd["T"])

    # -------------------------------------------------------------- Step (mpm.py:163-225)
    def Step(self):
        k, n = self.k, self.n
        R = self.R
        inp = self.state
        prep = inp
        was_settling = self.IsSettling
        if self.tr is not None:
            prep, self.work = self.transport_advance(inp, settling=was_settling)
        out = OracleState(np.zeros_like(inp.x), np.zeros_like(inp.v), np.zeros_like(inp.F), np.zeros_like(inp.C))
        if self.ding_cfg is not None:
            out.ding = {kk: vv.copy() for kk, vv in inp.ding.items()}  # ping-pong buffer contents are overwritten below
        self.node_m[...] = 0
        self.node_p[...] = 0
        self.node_v[...] = 0
        tau = np.zeros((n, 3, 3), R)
        self.compute_stress(prep, tau)
        if self.tr is not None:
            t = self.tr
            tau2 = np.zeros((n, 3, 3), R)
            k.WaterStress(n, prep.F, prep.transport["s"], tau, tau2, t.initial_saturation, t.water_stress if t.water_enabled else 0.0)
            tau = tau2
        self.stress = tau
        k.ParticleToGrid(n, prep.x, prep.v, prep.C, tau, self.Masses(prep), self.V0, self.node_m, self.node_p, self.origin,
                         self.dx, self.res, self.dt, self.flags)
        k.UpdateGrid(i64(self.node_count), self.node_m, self.node_p, self.node_v, self.res, self.sim.boundary_cells, self.gravity,
                     self.dt)
        if self.sticky_floor:
            self.node_v.reshape(self.res, self.res, self.res, 3)[:, :self.sim.boundary_cells] = 0.0
        k.GridToParticle(n, prep.x, prep.F, self.node_v, out.x, out.v, out.C, out.F, self.origin, self.dx, self.res, self.dt,
                         self.flags)
        if self.ding_cfg is not None:
            self.advance_ding(prep, out)
            self.check_ding()
        if self.tr is not None:
            out.transport = dict(prep.transport)
            out.transport["s"] = np.zeros(n)
            k.Saturations(n, out.transport["W"], self.V0, out.F, out.transport["s"], self.tr.water_density)
        self.prepared = prep
        self.state = out
        self.step_count += 1
        if was_settling:
            self.update_settling()

    def Advance(self, steps):
        for _ in range(steps):
            self.Step()
        if self.flags[0]:
            raise RuntimeError("nonpositive/nonfinite J; reduce dt or inspect loading and sampling")
        if self.flags[1]:
            raise RuntimeError("particle stencil left the grid; enlarge the domain or reduce dt")

    def UpdateMesh(self):
        self.reconstruct(self.state, self.vertices)
        return self.vertices

    # -------------------------------------------------------------- aggregates (experiments/recording.py:46-112)
    def Measure(self):
        """The reference recorder's measurement of the current state (experiments/recording.py:46-112): mesh
        extents and volume after UpdateMesh, particle volume sum V0 det F, Ding porosity/temperature/gelation
        statistics, gas content and pressures, capacity-weighted and surface/core temperatures, water and
        energy balances."""
        s = self.state
        verts = np.asarray(self.UpdateMesh(), dtype=np.float64)
        F = np.asarray(s.F, dtype=np.float64)
        if not np.isfinite(verts).all() or not np.isfinite(F).all():
            raise ValueError("Cannot record nonfinite simulation state")
        corners = verts[self.mesh.tet_indices]
        signed = np.linalg.det(corners[:, 1:] - corners[:, :1]) / 6.0
        J = np.linalg.det(F)
        extent = np.ptp(verts, axis=0)
        V0 = np.asarray(self.samples.reference_volumes, dtype=np.float64)
        out = {}
        if s.ding is not None:
            d = s.ding
            por = np.asarray(d["porosity"], np.float32)
            out.update(temperature_k=float(np.asarray(d["T"], np.float32).mean()), porosity_min=float(por.min()),
                       porosity_mean=float(por.mean(dtype=np.float64)), porosity_max=float(por.max()),
                       yield_stress_pa=float(np.asarray(d["yield_stress"], np.float32)[0]),
                       viscosity_pa3_s=float(np.asarray(d["viscosity"], np.float32)[0]),
                       gelled_fraction=float(np.asarray(d["gelled"]).mean()))
            if self.ding_cfg.gas_enabled:
                pr = d["pressure"]
                out.update(gas_total_mol=float(np.dot(V0, d["moles"])), gas_pressure_min_pa=float(pr.min()),
                           gas_pressure_max_pa=float(pr.max()), gas_pressure_mean_pa=float(pr.mean()),
                           pore_volume_m3=float(np.dot(V0, J * por)))
        if s.transport is not None:
            t = s.transport
            T, W, sat = t["T"], t["W"], t["s"]
            cap = self.capacities
            surface = np.zeros(self.n, dtype=bool)
            surface[self.owners] = True
            pos = np.asarray(self.samples.particle_positions)
            radius = np.linalg.norm(pos - np.mean(self.mesh.vertex_positions, axis=0), axis=1)
            core = radius < 0.4 * radius.max()
            if not core.any():
                core = radius == radius.min()
            initial_water = self.tr.water_density * self.tr.initial_saturation * V0.sum()
            energy = float(np.dot(cap, T))
            exchange = float(t["exchanges"].sum())
            losses = float(t["losses"].sum())
            out.update(temperature_min_k=float(T.min()), temperature_max_k=float(T.max()),
                       temperature_mean_k=float(np.average(T, weights=cap)), temperature_surface_k=float(T[surface].mean()),
                       temperature_core_k=float(T[core].mean()), water_mass_kg=float(W.sum()), water_lost_kg=losses,
                       water_balance_error_kg=float(W.sum() + losses - initial_water), saturation_min=float(sat.min()),
                       saturation_max=float(sat.max()), saturation_mean=float(sat.mean()), thermal_energy_j=energy,
                       heat_exchange_j=exchange,
                       energy_balance_error_j=float(energy - self.tr.initial_temperature_k * cap.sum() - exchange))
        out.update(baking_time_s=self.baking_time, phase="settling" if self.IsSettling else "baking",
                   step=self.step_count, time_s=self.step_count * self.dt, length_x_m=float(extent[0]),
                   width_z_m=float(extent[2]), height_y_m=float(extent[1]), particle_volume_m3=float(np.dot(V0, J)),
                   mesh_volume_m3=float(np.abs(signed).sum()), inverted_tets=int(np.count_nonzero(signed <= 0)),
                   invalid_particles=int(np.count_nonzero(J <= 0)))
        return out
