"""MLS-MPM baking simulator driver (reference ``physics/mpm.py:17-340``).

Same constructor, methods and attributes as the reference's ``MpmSimulator``;
the substep itself is one call of the fused gfx950 pipeline ``bk_substep``
(transport pre-pass, stresses, P2G, grid update, G2P + Ding return map + gas
refresh + saturations) on preallocated ping-pong state buffers.

Error semantics follow the reference:

* ``Step()`` reads the device flag block once per substep and raises the
  transport or Ding error before the new state is accepted
  (``physics/mpm.py:205-207``); J and stencil flags are raised by
  ``CheckState`` (``physics/mpm.py:243-248``).
* ``Advance(n)`` runs the n substeps without host synchronisation, then reads
  the flags once.  If any transport/Ding flag is set it restores the state it
  started from and replays substep by substep, so the error is raised at the
  same substep, with the same accepted state, as ``Step()`` would.

Settling (``physics/mpm.py:57-109``): while the dough rests, the substep's
parameters carry ``settling = 2`` and the library's settling controller --
a per-dough record on the device -- reduces the rest metrics after every
substep and starts baking, so resting substeps need no host round trip.  The
host reads the record back after ``Step`` and at the end of each ``Advance``
chunk; ``Advance`` ends a chunk at ``settling_steps`` so that a timeout raises
at the reference's substep with the reference's accepted state.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np
import torch

from .. import _native
from ..geometry.mesh_tracking import ParticleMeshBinding
from .grid import GridState
from .materials.ding import DingMaterial, DingState
from .materials.elastic import ElasticMaterial
from .state import AllocateState, CreateParticleState
from .thermal_water import ThermalWaterTransport, TransportState

_FLAG_J, _FLAG_STENCIL, _FLAG_POROSITY, _FLAG_COUPLED, _FLAG_GAS, _FLAG_HEAT, _FLAG_WATER = range(7)


def _resolve_device(name):
    if name == "cpu":
        raise RuntimeError("baking_amd has no CPU substep: use device_name 'auto' or 'cuda[:N]' on a ROCm GPU "
                           "(the CPU restatement in oracle/ is test infrastructure only)")
    _native.require_gpu()
    if name == "auto":
        return "cuda:0"
    return name if ":" in name else "cuda:0"


class MpmSimulator:
    """Forward-mode MLS-MPM simulator on one ROCm GPU; history mode is not provided by the fused path."""

    def __init__(self, tet_mesh, particle_samples, simulation_config, material_config, initial_velocity=(0.0, 0.0, 0.0),
                 retain_history=False, ding_config=None, transport_config=None):
        simulation_config.Validate(material_config)
        if retain_history:
            raise NotImplementedError("retain_history: use baking_amd.adjoint for gradients; the fused substep is forward-only")
        self.device_name = _resolve_device(simulation_config.device_name)
        torch.cuda.set_device(torch.device(self.device_name))
        _native.load()
        self.tet_mesh = tet_mesh
        self.particle_samples = particle_samples
        self.simulation_config = simulation_config
        self.material = ElasticMaterial(material_config)
        if material_config.model == "ding2019":
            from ..config import DingConfig
            self.material = DingMaterial(material_config, ding_config or DingConfig(), self.device_name, retain_history)
        if isinstance(self.material, DingMaterial):
            self.material.simulation_config = simulation_config
        if (isinstance(self.material, DingMaterial) and self.material.ding_config.gas_enabled
                and not (transport_config and transport_config.enabled)):
            raise ValueError("Ding gas simulation requires transport.enabled=true; prescribed temperature schedules were removed")
        self.initial_velocity = initial_velocity
        self.retain_history = retain_history
        self.ValidateInitialDomain()
        self.mesh_binding = ParticleMeshBinding(tet_mesh, particle_samples, self.device_name)
        self.transport = None
        if transport_config is not None and transport_config.enabled:
            transport_config.Validate(material_config)
            if (transport_config.settling_duration_s > 0 and transport_config.settling_auto
                    and not isinstance(self.material, DingMaterial)):
                raise ValueError("automatic settling measures the plastic-gradient rate and needs the ding2019 material")
            self.material.initial_temperature_k = transport_config.initial_temperature_k
            self.transport = ThermalWaterTransport(self, transport_config)
        self._workspace = None
        self.Reset()

    # ------------------------------------------------------------------ properties kept from the reference
    @property
    def _settling_configured(self):
        return self.transport is not None and self.transport.config.settling_duration_s > 0

    @property
    def settling_steps(self):
        if self._baking_start_step is not None and self._baking_start_step > 0:
            return self._baking_start_step
        if self.transport is None:
            return 0
        return int(math.ceil(self.transport.config.settling_duration_s / self.simulation_config.time_step - 1e-10))

    @property
    def IsSettling(self):
        return self._settling_configured and self._baking_start_step is None

    @property
    def baking_steps(self):
        return 0 if self.IsSettling else max(0, self.step_count - self.settling_steps)

    @property
    def baking_time(self):
        return self.baking_steps * self.simulation_config.time_step

    def Masses(self, state=None):
        state = self.current_state if state is None else state
        return state.transport_state.particle_masses if state.transport_state is not None else self.particle_data.particle_masses

    def ValidateInitialDomain(self):
        sim = self.simulation_config
        origin = np.asarray(sim.grid_origin)
        lower = origin + (sim.boundary_cells - 1) * sim.grid_spacing
        upper = origin + (sim.grid_resolution - sim.boundary_cells) * sim.grid_spacing
        points = np.concatenate([self.tet_mesh.vertex_positions, self.particle_samples.particle_positions])
        if np.any(points < lower) or np.any(points > upper):
            raise ValueError(f"model must initially fit inside {lower} .. {upper}; "
                             "adjust [mesh] mesh_scale/mesh_translation or [simulation] grid settings in the config file")

    # ------------------------------------------------------------------ buffers
    def _new_state(self):
        """A full set of particle buffers (x, v, F, C + Ding + transport) for the ping-pong pair."""
        state = AllocateState(self.particle_data.particle_count, self.device_name)
        if isinstance(self.material, DingMaterial):
            self.material.InitializeState(state)
        if self.transport is not None:
            self.transport.InitializeState(state)
        return state

    def Reset(self):
        if isinstance(self.material, DingMaterial):
            self.material.Reset()
        self.current_state, self.particle_data = CreateParticleState(self.particle_samples, self.device_name,
                                                                     self.initial_velocity, self.retain_history)
        if isinstance(self.material, DingMaterial):
            self.material.InitializeState(self.current_state)
        if self.transport is not None:
            self.transport.InitializeState(self.current_state)
        self.next_state = self._new_state()
        self.grid_state = GridState(self.simulation_config, self.device_name, self.retain_history)
        n = self.particle_data.particle_count
        self.kirchhoff_stresses = torch.zeros(n, 3, 3, dtype=torch.float32, device=self.device_name)
        self.vertex_positions = torch.as_tensor(self.tet_mesh.vertex_positions.astype(np.float32), device=self.device_name)
        self.error_flags = torch.zeros(2, dtype=torch.int32, device=self.device_name)
        self.step_count = 0
        self._baking_start_step = None if self._settling_configured else 0
        self.settling_quiet_steps = 0
        self.settling_reference_step = 0
        self.settling_plastic_rate = 0.0
        self.settling_reason = "pending" if self._baking_start_step is None else "disabled"
        self.settling_diagnostics = dict(max_speed_m_s=0.0, max_strain_rate_per_s=0.0, max_plastic_rate_per_s=0.0)
        self.mesh_step_count = 0
        self._build_native()
        if self._settling_configured:
            self._start_settling()
        if getattr(self, "result_recorder", None) is not None:
            self.result_recorder.Reset()

    def _build_native(self):
        lib = _native.load()
        self._destroy_graphs()
        if self._workspace is not None:
            lib.bk_workspace_destroy(self._workspace)
            self._workspace = None
        sim = self.simulation_config
        n = self.particle_data.particle_count
        dev = self.device_name
        nv = len(self.tet_mesh.vertex_positions) if self.transport is not None else 0
        nslots = int(len(self.transport.slots["slot_owner"])) if self.transport is not None else 0
        ws = lib.bk_workspace_create(n, sim.grid_resolution, nv, nslots)
        if not ws:
            _native.check(-2, "bk_workspace_create")
        self._workspace = ctypes.c_void_p(ws)
        flags_ptr = lib.bk_workspace_flags(self._workspace)
        self._ws_flags = flags_ptr
        self._flags_host = torch.zeros(8, dtype=torch.int32).pin_memory()
        tiled = self.grid_state.tiled_count
        if self.transport is not None:
            self.thermal_grid = {k: torch.zeros(tiled, dtype=torch.float64, device=dev)
                                 for k in ("V", "C", "E", "W", "H", "L", "X", "dT", "dW")}
        else:
            self.thermal_grid = {}
        # static scene
        keep = {}
        keep["V0"] = self.particle_data.reference_volumes
        keep["m0"] = self.particle_data.particle_masses
        if isinstance(self.material, DingMaterial):
            keep["visc_param"] = self.material.viscosity_parameter
            keep["prod_param"] = self.material.production_parameter
        if self.transport is not None:
            t = self.transport
            keep.update(capacities=t.capacities, dry=t.dry_masses, heat_param=t.heat_transfer_parameter,
                        ref_vertices=self.mesh_binding.reference_vertices, bind_offsets=self.mesh_binding.binding_offsets,
                        bind_particles=self.mesh_binding.binding_particles, bind_weights=self.mesh_binding.binding_weights,
                        bind_ref_offsets=self.mesh_binding.reference_offsets, faces=t.faces,
                        slot_face_offsets=t.slots["slot_face_offsets"], slot_faces=t.slots["slot_faces"],
                        slot_owner=t.slots["slot_owner"], particle_slot=t.slots["particle_slot"])
        self._scene_tensors = keep
        scene = _native.BkScene()
        for name in _native.SCENE_FIELDS:
            setattr(scene, name, keep[name].data_ptr() if name in keep else None)
        self._scene = scene
        self._refresh_grid_struct()
        self._settle_record = None
        self._settle_reference = None
        self._check_buffers = None
        self._params = self._make_params()

    def _refresh_grid_struct(self):
        grid = _native.BkGrid()
        grid.m = self.grid_state.tiled_masses.data_ptr()
        grid.p = self.grid_state.tiled_momenta.data_ptr()
        grid.v = self.grid_state.tiled_velocities.data_ptr()
        for k, v in self.thermal_grid.items():
            setattr(grid, k, v.data_ptr())
        self._grid = grid

    def __del__(self):
        try:
            self._destroy_graphs()
            if self._workspace is not None:
                _native.load().bk_workspace_destroy(self._workspace)
        except Exception:
            pass

    def _make_params(self, settling=None):
        """bk_params of this simulator.  settling: 0 bake, 1 rest, 2 follow the settling controller;
        None picks 2 while the dough rests and 0 once baking has started."""
        if settling is None:
            settling = 2 if self.IsSettling else 0
        settling = int(settling)
        rest = settling == 1
        sim, mat = self.simulation_config, self.material.material_config
        P = _native.BkParams()
        P.origin = (ctypes.c_float * 3)(*sim.grid_origin)
        P.dx = sim.grid_spacing
        P.res = sim.grid_resolution
        P.boundary_cells = sim.boundary_cells
        P.gravity = (ctypes.c_float * 3)(*sim.gravity)
        P.dt = sim.time_step
        P.cfl_limit = sim.cfl_limit
        P.mu, P.lam, P.density = mat.shear_modulus, mat.lame_lambda, mat.density
        P.model = 1 if isinstance(self.material, DingMaterial) else 0
        if P.model == 1:
            c = self.material.ding_config
            P.yield_stress, P.n0 = c.yield_stress, c.porosity
            P.temperature_enabled, P.gas_enabled = int(c.temperature_enabled), int(c.gas_enabled)
            P.knots = (ctypes.c_float * 4)(*c.temperature_knots_k)
            P.yield_factors = (ctypes.c_float * 4)(*c.yield_factors)
            P.viscosity_factors = (ctypes.c_float * 4)(*c.viscosity_factors)
            P.p_ambient, P.gas_peak, P.gas_width = c.gas_ambient_pressure_pa, c.gas_reaction_temperature_k, c.gas_reaction_width_k
        if self.transport is not None:
            t = self.transport.config
            P.transport_enabled, P.water_enabled, P.settling = 1, int(t.water_enabled), settling
            P.T_oven, P.T_tray = t.oven_temperature_k, t.tray_temperature_k
            P.tray_h = 0.0 if rest else t.tray_heat_transfer
            P.loss_coef = 0.0 if rest else t.surface_water_loss
            P.rho_w, P.latent = t.water_density, t.latent_heat
            P.kappa = 0.0 if rest else t.conductivity
            P.diffusivity = 0.0 if rest else t.water_diffusivity
            P.s0, P.water_stress, P.contact_distance = t.initial_saturation, t.water_stress, t.contact_distance
            P.n_vertices = len(self.tet_mesh.vertex_positions)
            P.n_faces = len(self.transport.faces)
            P.n_slots = len(self.transport.slots["slot_owner"])
        P.n = self.particle_data.particle_count
        P.return_solver = {"bisection": 0, "bracketed": 1}[self.return_solver]
        P.surrogate = int(self._surrogate_on)
        P.sticky_floor = int(getattr(self, "_sticky_floor", False))
        return P

    # ------------------------------------------------------------------ settling controller (physics/mpm.py:77-109)
    def _start_settling(self):
        """Device record of a dough starting its rest at step 0; the reference plastic gradients are Fp now."""
        t = self.transport.config
        self._settle_reference = None
        if isinstance(self.material, DingMaterial):
            n = self.particle_data.particle_count
            self._settle_reference = torch.zeros(2, n, 3, 3, dtype=torch.float32, device=self.device_name)
            _native.check(_native.load().bk_workspace_settling_reference(self._workspace, _native.ptr(self._settle_reference)),
                          "bk_workspace_settling_reference")
        rec = _native.BkSettling()
        rec.auto_mode = int(t.settling_auto)
        rec.settling_steps = self.settling_steps
        rec.time_step = self.simulation_config.time_step
        rec.min_duration_s, rec.hold_duration_s = t.settling_min_duration_s, t.settling_hold_duration_s
        rec.speed_tolerance, rec.rate_tolerance = t.settling_speed_tolerance, t.settling_rate_tolerance
        rec.phase, rec.reason, rec.step_count, rec.baking_start_step = 1, 0, self.step_count, -1
        self._write_settling(rec)
        if isinstance(self.material, DingMaterial):
            self.settling_reference_plastic.copy_(self.current_state.material_state.plastic_gradients)

    def _write_settling(self, rec):
        _native.check(_native.load().bk_workspace_settling_set(self._workspace, 0, 1, ctypes.byref(rec), _native.stream_ptr()),
                      "bk_workspace_settling_set")
        self._settle_record = _native.BkSettling.from_buffer_copy(rec)

    @property
    def settling_reference_plastic(self):
        """The live (N, 3, 3) half of the controller's reference plastic gradients (None without a rest)."""
        if self._settle_record is None or self._settle_reference is None:
            return None
        return self._settle_reference[self._settle_record.ref_slot]

    def _save_settling(self):
        """Record and live reference of the controller, to undo substeps that are replayed."""
        if self._settle_record is None:
            return None
        return _native.BkSettling.from_buffer_copy(self._settle_record), self.settling_reference_plastic.clone()

    def _restore_settling(self, saved):
        if saved is None:
            return
        rec, ref = saved
        self._write_settling(rec)
        self.settling_reference_plastic.copy_(ref)
        self._apply_settling(rec)

    def _apply_settling(self, rec):
        """Mirror the device record in the reference's simulator attributes."""
        self.settling_quiet_steps = int(rec.quiet_steps)
        self.settling_reference_step = int(rec.reference_step)
        self.settling_plastic_rate = float(rec.plastic_rate)
        self.settling_reason = _native.SETTLING_REASONS[rec.reason]
        if rec.auto_mode:
            d = rec.diagnostics
            self.settling_diagnostics = dict(max_speed_m_s=float(d[0]), max_strain_rate_per_s=float(d[1]),
                                             max_plastic_rate_per_s=float(d[2]))
        if rec.phase == 0 and self._baking_start_step is None:
            self._baking_start_step = int(rec.baking_start_step)
            self._params = self._make_params()
        elif rec.phase == 1 and self._baking_start_step is not None and self._settling_configured:
            self._baking_start_step = None
            self._params = self._make_params()

    def _pull_settling(self):
        """Read the record after the substeps already enqueued; raise the reference's timeout error."""
        rec = _native.BkSettling()
        _native.check(_native.load().bk_workspace_settling_get(self._workspace, 0, 1, ctypes.byref(rec), _native.stream_ptr()),
                      "bk_workspace_settling_get")
        if rec.step_count != self.step_count if rec.phase else rec.step_count > self.step_count:  # counts resting substeps
            raise RuntimeError(f"settling controller is at substep {rec.step_count}, the simulator at {self.step_count}")
        self._settle_record = rec
        self._apply_settling(rec)
        if rec.reason == 3 and rec.phase == 1:
            raise RuntimeError('Settling did not reach quasi-static tolerances before settling_duration_s; '
                               'baking was not started. Increase the maximum duration or review tolerances/material.')

    def _sync_settling(self):
        """Follow changes of ``transport.config`` made since the last substep, as the reference, which reads it
        live: the rest/bake mode of the parameters and the controller's configuration."""
        if self.transport is None:
            return
        mode = 2 if self.IsSettling else 0
        if self._params.settling != mode:
            self._params = self._make_params()
        rec = self._settle_record
        if not self.IsSettling or rec is None:
            return
        t = self.transport.config
        want = (int(t.settling_auto), self.settling_steps, t.settling_min_duration_s, t.settling_hold_duration_s,
                t.settling_speed_tolerance, t.settling_rate_tolerance)
        have = (rec.auto_mode, rec.settling_steps, rec.min_duration_s, rec.hold_duration_s, rec.speed_tolerance, rec.rate_tolerance)
        if want != have:
            if want[0] and self._settle_reference is None:
                raise ValueError("automatic settling measures the plastic-gradient rate and needs the ding2019 material")
            new = _native.BkSettling.from_buffer_copy(rec)
            (new.auto_mode, new.settling_steps, new.min_duration_s, new.hold_duration_s, new.speed_tolerance,
             new.rate_tolerance) = want
            self._write_settling(new)

    def _settling_chunk(self, remaining):
        """Substeps Advance may run before reading the controller: a timeout can only come at settling_steps."""
        if not self.IsSettling or self.step_count + remaining < self.settling_steps:
            return remaining
        return max(1, self.settling_steps - self.step_count)

    @staticmethod
    def _state_struct(state):
        s = _native.BkState()
        for field, t in state_field_tensors(state).items():
            setattr(s, field, t.data_ptr())
        return s

    # ------------------------------------------------------------------ substep
    use_graphs = True
    graph_steps = 4  # substeps per graph launch in runs of substeps (even; 1: one launch per substep)
    # Coupled-porosity solve of the fused substep: "bracketed" (default) solves the
    # reference's end-state porosity equation to a few ulps in ~10-20 residual
    # evaluations; "bisection" runs the reference's 64 fixed bisection steps.
    return_solver = "bracketed"

    # Start coupled-porosity solves from the return-map network (baking_amd.surrogate.attach sets it)
    _surrogate_on = 0

    def SetSurrogate(self, on):
        """0/False: no network; 1/True: the network guides the coupled solves (exact results); 2: it stands in for
        them when its one-step correction passes (see baking_amd.surrogate.attach)."""
        mode = int(on)
        if mode not in (0, 1, 2):
            raise ValueError("surrogate mode must be 0, 1 or 2")
        self._surrogate_on = mode
        self._params = self._make_params()

    def SetStickyFloor(self, on=True):
        """Floor option without a reference counterpart: every node of the floor's boundary_cells layers gets
        velocity 0 after the reference's frictionless unilateral wall projection (physics/grid.py:32-37), so dough
        on the tray does not slide.  Off by default (the reference's walls)."""
        self._sticky_floor = bool(on)
        self._params = self._make_params()

    def SetReturnSolver(self, name):
        if name not in ("bisection", "bracketed"):
            raise ValueError("return solver must be 'bisection' or 'bracketed'")
        self.return_solver = name
        self._params = self._make_params()

    def SetTuning(self, tile_chunk=None, particle_block=None, grid_cap=None, deferred_wgs=None):
        """Launch shapes of the fused substep (``bk_tuning``); captured graphs are dropped."""
        lib = _native.load()
        t = self.Tuning()
        for name, value in (("tile_chunk", tile_chunk), ("particle_block", particle_block), ("grid_cap", grid_cap),
                            ("deferred_wgs", deferred_wgs)):
            if value is not None:
                t[name] = int(value)
        struct = _native.BkTuning(t["tile_chunk"], t["particle_block"], t["grid_cap"], t["deferred_wgs"])
        _native.check(lib.bk_workspace_set_tuning(self._workspace, ctypes.byref(struct)), "bk_workspace_set_tuning")
        self._destroy_graphs()

    def Tuning(self):
        struct = _native.BkTuning()
        _native.check(_native.load().bk_workspace_get_tuning(self._workspace, ctypes.byref(struct)), "bk_workspace_get_tuning")
        return dict(tile_chunk=struct.tile_chunk, particle_block=struct.particle_block, grid_cap=struct.grid_cap,
                    deferred_wgs=struct.deferred_wgs)

    def _launch_substeps(self, count, stream=None):
        """Enqueue ``count`` substeps and accept each.  As HIP graphs, a run of two or more is pipelined (each
        substep's transport pre-pass and P2G set-up beside the previous substep's deferred return maps; the last
        one's are finished here), so the state is complete in stream order when this returns either way."""
        if count < 2 or not self.use_graphs:
            for _ in range(count):
                self._launch_substep(stream)
                self._accept()
            return
        k = self.graph_steps
        while k > 1 and count >= k:  # k substeps per graph launch (k even: each launch ends on its first buffers)
            self._launch_graph_steps(k, stream)
            count -= k
        for _ in range(count):
            self._launch_substep(stream, pipelined=True)
            self._accept()
        lib = _native.load()
        s_in, s_out = self._state_struct(self.next_state), self._state_struct(self.current_state)  # the last substep's
        _native.check(lib.bk_substep_finish(ctypes.byref(self._params), ctypes.byref(self._scene), ctypes.byref(s_in),
                                            ctypes.byref(s_out), self._workspace, _native.stream_ptr(stream)), "bk_substep_finish")

    def _save_for_tuning(self):
        flags = self._read_flags()
        tensors = _state_tensors(self.current_state) + _state_tensors(self.next_state)
        return (self.current_state, self.next_state, self.step_count, flags, [t.clone() for t in tensors],
                self._save_settling())

    def _restore_after_tuning(self, saved):
        current, nxt, steps, flags, clones, settling = saved
        torch.cuda.synchronize()
        self.current_state, self.next_state, self.step_count = current, nxt, steps
        for dst, src in zip(_state_tensors(current) + _state_tensors(nxt), clones):
            dst.copy_(src)
        self.grid_state.Clear()
        for t in self.thermal_grid.values():
            t.zero_()
        self._write_flags(flags)
        self._restore_settling(settling)

    def _destroy_graphs(self):
        lib = _native.load()
        for handle in getattr(self, "_graphs", {}).values():
            lib.bk_graph_destroy(handle)
        self._graphs = {}

    def _launch_substep(self, stream=None, use_graph=None, pipelined=False):
        """Enqueue one substep; replays a captured HIP graph for this ping-pong direction when possible.
        ``pipelined`` (graphs only): the graph that runs beside the previous substep's deferred return maps and
        leaves its own pending (see ``_launch_substeps``)."""
        lib = _native.load()
        s_in = self._state_struct(self.current_state)
        s_out = self._state_struct(self.next_state)
        use_graph = self.use_graphs if use_graph is None else use_graph
        if use_graph:
            key = bytes(s_in) + bytes(s_out) + bytes(self._params) + bytes([int(pipelined)])
            graph = self._graphs.get(key)
            if graph is None:
                if len(self._graphs) >= 8:
                    self._destroy_graphs()
                create = lib.bk_substep_graph_create_pipelined if pipelined else lib.bk_substep_graph_create
                handle = create(ctypes.byref(self._params), ctypes.byref(self._scene), ctypes.byref(s_in), ctypes.byref(s_out),
                                ctypes.byref(self._grid), self._workspace)
                if not handle:
                    _native.check(-2, "bk_substep_graph_create")
                graph = self._graphs[key] = ctypes.c_void_p(handle)
            _native.check(lib.bk_graph_launch(graph, _native.stream_ptr(stream)), "bk_graph_launch")
            return
        rc = lib.bk_substep(ctypes.byref(self._params), ctypes.byref(self._scene), ctypes.byref(s_in), ctypes.byref(s_out),
                            ctypes.byref(self._grid), self._workspace, _native.stream_ptr(stream))
        _native.check(rc, "bk_substep")

    def _launch_graph_steps(self, k, stream=None):
        """Enqueue and accept ``k`` pipelined substeps captured in one HIP graph (alternating ping-pong buffers)."""
        lib = _native.load()
        s_in = self._state_struct(self.current_state)
        s_out = self._state_struct(self.next_state)
        key = bytes(s_in) + bytes(s_out) + bytes(self._params) + bytes([2, k])
        graph = self._graphs.get(key)
        if graph is None:
            if len(self._graphs) >= 8:
                self._destroy_graphs()
            handle = lib.bk_substep_graph_create_steps(ctypes.byref(self._params), ctypes.byref(self._scene), ctypes.byref(s_in),
                                                       ctypes.byref(s_out), ctypes.byref(self._grid), self._workspace, k, 1)
            if not handle:
                _native.check(-2, "bk_substep_graph_create_steps")
            graph = self._graphs[key] = ctypes.c_void_p(handle)
        _native.check(lib.bk_graph_launch(graph, _native.stream_ptr(stream)), "bk_graph_launch")
        for _ in range(k):
            self._accept()

    def _accept(self):
        self.current_state, self.next_state = self.next_state, self.current_state
        self.step_count += 1

    def _read_flags(self):
        """Blocking read of the 8-int device flag block (waits for the substeps already enqueued)."""
        lib = _native.load()
        _native.check(lib.bk_workspace_read_flags(self._workspace, ctypes.c_void_p(self._flags_host.data_ptr()),
                                                  _native.stream_ptr()), "bk_workspace_read_flags")
        return self._flags_host.tolist()

    def _write_flags(self, values):
        lib = _native.load()
        arr = (ctypes.c_int32 * 8)(*values)
        _native.check(lib.bk_workspace_write_flags(self._workspace, ctypes.cast(arr, ctypes.c_void_p), _native.stream_ptr()),
                      "bk_workspace_write_flags")

    def _raise_for(self, flags):
        if flags[_FLAG_HEAT] or flags[_FLAG_WATER]:
            raise RuntimeError('Heat/water explicit stability bound exceeded; reduce time_step; '
                               'no temperature or water clipping is applied')
        if isinstance(self.material, DingMaterial):
            self.material.RaiseFor([flags[_FLAG_POROSITY], flags[_FLAG_COUPLED], flags[_FLAG_GAS]])

    def _sync_api_flags(self, flags):
        self.error_flags.copy_(torch.tensor(flags[:2], dtype=torch.int32))
        if isinstance(self.material, DingMaterial):
            self.material.error_flags.copy_(torch.tensor(flags[2:5], dtype=torch.int32))
        if self.transport is not None:
            self.transport.error_flags.copy_(torch.tensor(flags[5:7], dtype=torch.int32))

    def Step(self):
        self._sync_settling()
        was_settling = self.IsSettling
        saved = self._save_settling() if was_settling else None
        self._launch_substep()
        flags = self._read_flags()
        self._sync_api_flags(flags)
        try:
            self._raise_for(flags)
        except RuntimeError:
            self._restore_settling(saved)  # the rejected substep never ran for the controller either
            raise
        self._accept()
        if was_settling:
            self._pull_settling()
        if getattr(self, "result_recorder", None) is not None:
            self.result_recorder.Record()

    def Advance(self, step_count, check=True):
        """Run ``step_count`` substeps with one flag read at the end, then ``CheckState``."""
        if step_count < 1:
            raise ValueError("step_count must be positive")
        if getattr(self, "result_recorder", None) is not None:
            for _ in range(step_count):
                self.Step()
            return self.CheckState() if check else None
        remaining = step_count
        while remaining > 0:
            chunk = self._settling_chunk(remaining)
            self._advance_chunk(chunk)
            remaining -= chunk
        self._sync_api_flags(self._read_flags())
        return self.CheckState() if check else None

    def _advance_chunk(self, count):
        self._sync_settling()
        was_settling = self.IsSettling
        flags_before = self._read_flags()
        start_state, start_next, start_steps = self.current_state, self.next_state, self.step_count
        saved = [t.clone() for t in _state_tensors(start_state)]
        saved_settling = self._save_settling() if was_settling else None
        self._launch_substeps(count)
        flags = self._read_flags()
        if any(flags[i] and not flags_before[i] for i in range(_FLAG_POROSITY, _FLAG_WATER + 1)):
            # Replay from the starting state, checking after every substep, so the
            # error surfaces at the same substep with the same accepted state.
            self.current_state, self.next_state, self.step_count = start_state, start_next, start_steps
            for dst, src in zip(_state_tensors(start_state), saved):
                dst.copy_(src)
            self._write_flags(flags_before)
            self._restore_settling(saved_settling)
            for _ in range(count):
                self.Step()
            return
        del saved
        if was_settling:
            self._pull_settling()

    def UpdateMesh(self, force_update=False):
        if force_update or self.mesh_step_count != self.step_count:
            self.mesh_binding.Reconstruct(self.current_state, self.vertex_positions)
            self.mesh_step_count = self.step_count

    def CheckState(self):
        """Health checks and run statistics (physics/mpm.py:243-308), reduced on the GPU in one pass
        (``bk_check_state``): the host reads 15 numbers, not the particle state.  Errors are raised in the
        reference's order with its texts; the mesh is reconstructed into ``vertex_positions`` only when every
        check passes, as the reference's ``UpdateMesh`` call comes last."""
        flags = self._read_flags()
        self._sync_api_flags(flags)
        if flags[_FLAG_J]:
            raise RuntimeError("nonpositive/nonfinite J; reduce dt or inspect loading and sampling")
        if flags[_FLAG_STENCIL]:
            raise RuntimeError("particle stencil left the grid; enlarge the domain or reduce dt")
        c = self._device_check()
        if c["nonfinite_state"]:
            raise RuntimeError("nonfinite particle state; simulation stopped")
        if c["minimum_j"] <= 0.0:
            raise RuntimeError("inverted particle deformation; reduce dt or loading")
        if c["outside_stencil"]:
            raise RuntimeError("particle outside valid grid stencil")
        sim = self.simulation_config
        vmax = c["maximum_speed"]
        wave = self.material.material_config.wave_speed
        if isinstance(self.material, DingMaterial) and self.material.ding_config.gas_enabled:
            wave = max(wave, c["gas_wave"])
        if self.transport is not None and self.transport.config.water_enabled and isinstance(self.material, DingMaterial):
            wave = float(np.sqrt(wave * wave + c["water_wave_sq"]))
        cfl = (vmax + wave) * sim.time_step / sim.grid_spacing
        if cfl > sim.cfl_limit:
            raise RuntimeError(f"estimated CFL={cfl:.3f} > {sim.cfl_limit}; reduce dt")
        if isinstance(self.material, DingMaterial):
            self.material.RaiseFor(flags[2:5])
            if c["nonfinite_ding"]:
                raise RuntimeError("nonfinite Ding internal state")
            if c["maximum_residual"] > 1e-4:
                raise RuntimeError("Ding local return mapping did not converge")
        if c["nonfinite_vertices"]:
            raise RuntimeError("nonfinite reconstructed mesh vertices")
        self.vertex_positions.copy_(self._check_vertices)
        self.mesh_step_count = self.step_count
        return {
            "step_count": self.step_count,
            "simulation_time": self.step_count * sim.time_step,
            "particle_count": self.particle_data.particle_count,
            "particle_volume_m3": c["particle_volume"],
            "total_mass": c["total_mass"],
            "grid_mass": c["grid_mass"],
            "minimum_j": c["minimum_j"],
            "maximum_j": c["maximum_j"],
            "maximum_speed": vmax,
            "kinetic_energy": c["kinetic_energy"],
            "estimated_cfl": cfl,
            "unmapped_mesh_vertices": self.mesh_binding.unmapped_vertex_count,
            "inverted_mesh_tets": int(c["inverted_tets"]),
        }

    def _device_check(self):
        """One bk_check_state pass over the current state, the grid masses and a freshly reconstructed mesh."""
        lib = _native.load()
        dev = self.device_name
        n = self.particle_data.particle_count
        if getattr(self, "_check_buffers", None) is None:
            tets = torch.as_tensor(np.ascontiguousarray(self.tet_mesh.tet_indices, dtype=np.int32), device=dev)
            nodes = self.grid_state.tiled_masses.numel()
            size = int(lib.bk_check_scratch_size(n, nodes, len(self.vertex_positions), len(tets)))
            self._check_buffers = dict(tets=tets, scratch=torch.empty(size, dtype=torch.float64, device=dev),
                                       out=torch.empty(len(_native.CHECK_FIELDS), dtype=torch.float64, device=dev))
            self._check_vertices = torch.empty_like(self.vertex_positions)
        b = self._check_buffers
        self.mesh_binding.Reconstruct(self.current_state, self._check_vertices)
        sim = self.simulation_config
        geometry = (ctypes.c_double * 4)(*sim.grid_origin, sim.grid_spacing)
        state = self._state_struct(self.current_state)
        masses = self.grid_state.tiled_masses
        _native.check(lib.bk_check_state(ctypes.byref(self._params), ctypes.byref(self._scene), ctypes.byref(state),
                                         ctypes.cast(geometry, ctypes.c_void_p), _native.ptr(masses), masses.numel(),
                                         _native.ptr(self._check_vertices), len(self._check_vertices), _native.ptr(b["tets"]),
                                         len(b["tets"]), _native.ptr(b["scratch"]), _native.ptr(b["out"]), _native.stream_ptr()),
                      "bk_check_state")
        return dict(zip(_native.CHECK_FIELDS, b["out"].cpu().tolist()))

    def CaptureSnapshot(self):
        """Host copy of the full state with the reference's NPZ keys (physics/mpm.py:310-334)."""
        self.UpdateMesh()
        s = self.current_state
        data = dict(
            particle_positions=s.particle_positions.cpu().numpy().copy(),
            particle_velocities=s.particle_velocities.cpu().numpy().copy(),
            deformation_gradients=s.deformation_gradients.cpu().numpy().copy(),
            affine_velocities=s.affine_velocities.cpu().numpy().copy(),
            particle_masses=self.Masses().cpu().numpy().copy(),
            reference_volumes=self.particle_data.reference_volumes.cpu().numpy().copy(),
            vertices=self.vertex_positions.cpu().numpy().copy(), tetrahedra=self.tet_mesh.tet_indices.copy(),
            reference_vertices=self.tet_mesh.vertex_positions.copy(),
            source_tets=self.particle_samples.source_tets.copy(),
            simulation_time=self.step_count * self.simulation_config.time_step,
            baking_time_s=self.baking_time, settling_steps=self.settling_steps)
        if isinstance(self.material, DingMaterial):
            for name in ("elastic_gradients", "plastic_gradients", "kirchhoff_stresses", "return_residuals", "porosities",
                         "temperatures", "gelled", "yield_stresses", "viscosities", "gas_moles", "gas_pressures",
                         "porosity_residuals", "reference_densities"):
                data["ding_" + name] = getattr(s.material_state, name).cpu().numpy().copy()
        if s.transport_state is not None:
            from dataclasses import fields
            for f in fields(s.transport_state):
                data["transport_" + f.name] = getattr(s.transport_state, f.name).cpu().numpy().copy()
        return data

    def SaveSnapshot(self, output_path):
        self.UpdateMesh()
        output_path = Path(output_path)
        output_path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(output_path, **self.CaptureSnapshot())


_MATERIAL_FIELDS = (("Fe", "elastic_gradients"), ("Fp", "plastic_gradients"), ("tau", "kirchhoff_stresses"),
                    ("residual", "return_residuals"), ("porosity", "porosities"), ("T_mat", "temperatures"),
                    ("gelled", "gelled"), ("yield_stress", "yield_stresses"), ("viscosity", "viscosities"),
                    ("porosity_residual", "porosity_residuals"), ("density", "reference_densities"),
                    ("moles", "gas_moles"), ("pressure", "gas_pressures"))
_TRANSPORT_FIELDS = (("T", "temperatures"), ("W", "water_masses"), ("s", "saturations"), ("losses", "water_losses"),
                     ("exchanges", "heat_exchanges"), ("masses", "particle_masses"))


def state_field_slots(state):
    """Where each ``bk_state`` field of a state lives: {field: (object, attribute name)}."""
    out = dict(x=(state, "particle_positions"), v=(state, "particle_velocities"), F=(state, "deformation_gradients"),
               C=(state, "affine_velocities"))
    if state.material_state is not None:
        out.update((field, (state.material_state, name)) for field, name in _MATERIAL_FIELDS)
    if state.transport_state is not None:
        out.update((field, (state.transport_state, name)) for field, name in _TRANSPORT_FIELDS)
    return out


def state_field_tensors(state):
    """The particle tensors of a state keyed by their ``bk_state`` field names (reference layouts, contiguous)."""
    out = {field: getattr(obj, name) for field, (obj, name) in state_field_slots(state).items()}
    for field, t in out.items():
        if not t.is_contiguous():
            raise ValueError(f"state field {field} must be contiguous")
    return out


def _state_tensors(state):
    out = [state.particle_positions, state.particle_velocities, state.deformation_gradients, state.affine_velocities]
    if state.material_state is not None:
        from dataclasses import fields
        out += [getattr(state.material_state, f.name) for f in fields(state.material_state)]
    if state.transport_state is not None:
        from dataclasses import fields
        out += [getattr(state.transport_state, f.name) for f in fields(state.transport_state)]
    return out


