"""Reverse-mode gradients through the fused substep: a checkpointed ``torch.autograd.Function``.

The reference differentiates a rollout by recording ``MpmSimulator.Step`` on a
``wp.Tape`` in history mode (``physics/mpm.py:18-22, 163-225``; used by
``tests/test_warp_autodiff.py`` and ``tests/test_ding.py:81-170``), which keeps
every substep's state and grid alive.  Here the forward runs the fused HIP
substep (graphs included) and keeps only a checkpoint of the state every
``checkpoint_every`` substeps.  The backward walks the segments in reverse:
it recomputes one segment from its checkpoint, keeping each substep's input
state, then, substep by substep in reverse, reruns that substep's forward
(which leaves its grid and its sorted P2G tiles in the workspace) and runs
the library's adjoint kernels on it (``bk_adjoint_substep``: the Ding return
map, G2P with the F update, the grid update, P2G with the stress).  Adjoints
are FP64.

Covered: the substep with the neo-Hookean model or the Ding model (pore gas,
the coupled porosity return and the temperature-dependent yield stress and
viscosity included) and the heat and water transport (mesh areas, Robin
terms, thermal P2G, FV diffusion, scalar G2P, gas source, water stress).
Gradients with respect to the initial positions, velocities, affine
velocities and deformation gradients, the Ding elastic gradients,
porosities, Kirchhoff stresses and gas moles, the temperatures, water
masses and loss/exchange accumulators, and the Ding viscosity, gas
production and heat-transfer parameters and the oven temperature.  Resting substeps before baking
(``physics/mpm.py:57-109``) are differentiated with the phase the forward's
settling controller gave them; the stopping decision itself has no
derivative, as in the reference (``physics/quasistatic.py:1-6``).  Not
covered: batched doughs.

    out = rollout(sim, steps, {"v": v0})          # tensors that may require grad
    loss = (out["x"][:, 0] - x0[:, 0]).mean()
    loss.backward()
"""
from __future__ import annotations

import ctypes
from dataclasses import replace

import torch

from . import _native
from .launch import Kernel
from .physics.materials.ding import DingMaterial
from .physics.mpm import state_field_tensors

# Adjoint of UpdateDing without pore gas (ding.py:140-221), arguments as UpdateDing's inputs, then the
# adjoints of its outputs Fe, tau, porosity (read), then of old_Fe, old_porosity, F_total, C_affine and
# the viscosity parameter (accumulated).
UpdateDingAdjoint = Kernel("bk_UpdateDing_adjoint",
                           ["f32", "f32", "i32", "f32", "f32", "f32", "f32", "f64", "f64", "f64", "f64", "f64", "f64", "f64",
                            "f64", "f", "f", "f", "f", "i", "v4", "v4", "v4", "f"])
# Implicit-function adjoint of the overstress solve (ding.py:90-100): (q2, p2, df, mf) rows, overstress adjoints -> 4 adjoints.
OverstressAdjoint = Kernel("bk_overstress_adjoint", ["f64", "f64", "f64"])

MECHANICAL_FIELDS = ("x", "v", "C", "F")
DING_FIELDS = ("Fe", "porosity", "tau")
GAS_FIELDS = ("moles",)
TRANSPORT_FIELDS = ("T", "W", "losses", "exchanges")
_WIDTH = dict(x=3, v=3, C=9, F=9, Fe=9, tau=9, porosity=1, moles=1, T=1, W=1, losses=1, exchanges=1)


def _gas(sim):
    return isinstance(sim.material, DingMaterial) and sim.material.ding_config.gas_enabled


def _transport(sim):
    return getattr(sim, "transport", None) is not None


def _fields_of(sim):
    fields = MECHANICAL_FIELDS
    if isinstance(sim.material, DingMaterial):
        fields += DING_FIELDS + (GAS_FIELDS if _gas(sim) else ())
    return fields + (TRANSPORT_FIELDS if _transport(sim) else ())


def _state_list(state):
    return list(state_field_tensors(state).values())


def _clone_state(state):
    """A copy of every state array (one multi-tensor copy)."""
    src = _state_list(state)
    dst = [torch.empty_like(t) for t in src]
    torch._foreach_copy_(dst, src)
    return dst


def _load_state(state, saved):
    torch._foreach_copy_(_state_list(state), saved)


def _binding_reverse(sim):
    """The mesh bindings indexed by particle, for the gather in the mesh adjoint (cached per simulator)."""
    rev = getattr(sim, "_adjoint_bindings", None)
    if rev is not None:
        return rev
    sc = sim._scene_tensors
    offs = sc["bind_offsets"].long()
    parts = sc["bind_particles"].long()
    n, nv = sim.particle_data.particle_count, offs.numel() - 1
    counts = offs[1:] - offs[:-1]
    vertex_of = torch.repeat_interleave(torch.arange(nv, device=offs.device), counts)
    order = torch.argsort(parts, stable=True)
    rev_offsets = torch.zeros(n + 1, dtype=torch.long, device=offs.device)
    rev_offsets[1:] = torch.cumsum(torch.bincount(parts, minlength=n), 0)
    # per-vertex weight sums as differences of one prefix sum (bindings are grouped by vertex)
    csum = torch.zeros(parts.numel() + 1, dtype=torch.float64, device=offs.device)
    csum[1:] = torch.cumsum(sc["bind_weights"].double(), 0)
    wsum = csum[offs[1:]] - csum[offs[:-1]]
    rev = dict(rev_offsets=rev_offsets.int().contiguous(), rev_bindings=order.int().contiguous(),
               rev_vertices=vertex_of[order].int().contiguous(), anchor_coef=(1.0 - wsum).contiguous())
    sim._adjoint_bindings = rev
    return rev


def _run(sim, count):
    for _ in range(count):
        sim._launch_substep()
        sim._accept()


def _run_forward(sim, count):
    """``count`` substeps as ``Advance`` runs them: while the dough rests, the settling controller decides
    when baking starts, and a timeout raises at the reference's substep."""
    while count > 0:
        chunk = sim._settling_chunk(count)
        resting = sim.IsSettling
        _run(sim, chunk)
        if resting:
            sim._pull_settling()
        count -= chunk


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sim, steps, every, fields, visc, prod, heat, oven, *inputs):
        lib = _native.load()
        state = state_field_tensors(sim.current_state)
        for f, t in zip(fields, inputs):
            state[f].copy_(t.detach().reshape(state[f].shape).to(state[f].dtype))
        ding = isinstance(sim.material, DingMaterial)
        if ding:
            sim.material.viscosity_parameter.copy_(visc.detach().reshape(1).to(torch.float32))
        if _gas(sim):
            sim.material.production_parameter.copy_(prod.detach().reshape(1).to(torch.float32))
        if _transport(sim):
            sim.transport.heat_transfer_parameter.copy_(heat.detach().reshape(1).to(torch.float32))
            if oven.numel():
                # the oven temperature is a plain kernel argument (physics/thermal_water.py:100, :325), read from
                # the live transport config like the reference's
                sim.transport.config = replace(sim.transport.config, oven_temperature_k=float(oven.detach().reshape(-1)[0]))
                sim._params = sim._make_params()
        flags_before = sim._read_flags()
        first, resting = sim.step_count, sim.IsSettling
        checkpoints = []
        for start in range(0, steps, every):
            checkpoints.append(_clone_state(sim.current_state))
            _run_forward(sim, min(every, steps - start))
        flags = sim._read_flags()
        if any(f and not b for f, b in zip(flags, flags_before)):
            sim._sync_api_flags(flags)
            sim._raise_for(flags)
            raise RuntimeError(f"rollout: the forward substeps raised error flags {flags}")
        ctx.sim, ctx.steps, ctx.every, ctx.fields, ctx.checkpoints = sim, steps, every, fields, checkpoints
        # substeps [0, rest) of the rollout rested (the stopping decision itself has no derivative)
        if resting:
            ctx.rest = steps if sim._baking_start_step is None else max(0, min(steps, sim._baking_start_step - first))
        else:
            ctx.rest = 0
        ctx.shapes = [t.shape for t in inputs]
        ctx.dtypes = [t.dtype for t in inputs]
        ctx.params = [(visc.shape, visc.dtype), (prod.shape, prod.dtype), (heat.shape, heat.dtype), (oven.shape, oven.dtype)]
        ctx.lib = lib
        final = state_field_tensors(sim.current_state)
        return tuple(final[f].clone() for f in fields)

    @staticmethod
    def backward(ctx, *grads):
        sim, lib, fields = ctx.sim, ctx.lib, ctx.fields
        n = sim.particle_data.particle_count
        dev = sim.current_state.particle_positions.device
        adj = {f: torch.zeros(n * _WIDTH[f], dtype=torch.float64, device=dev) for f in _fields_of(sim)}
        for f, g in zip(fields, grads):
            if g is not None:
                adj[f] += g.detach().reshape(-1).to(torch.float64)
        bk_adj = _native.BkAdjoint()
        for f, t in adj.items():
            setattr(bk_adj, f, t.data_ptr())
        if _transport(sim):
            for f, t in _binding_reverse(sim).items():
                setattr(bk_adj, f, t.data_ptr())
        param = torch.zeros(4, dtype=torch.float64, device=dev)
        scratch = torch.zeros(int(lib.bk_adjoint_scratch_size(ctypes.byref(sim._params))), dtype=torch.float64, device=dev)
        # the simulator is left where the forward ended
        keep = (sim.current_state, sim.next_state, sim.step_count, _clone_state(sim.current_state), sim._read_flags(),
                sim._params)
        # each substep runs again with its own phase, without the settling controller
        phase_params = (sim._make_params(settling=0), sim._make_params(settling=1))
        stream = _native.stream_ptr()
        # the ping-pong pair keeps its arrays throughout: one bk_state each
        structs = {id(s): sim._state_struct(s) for s in (sim.current_state, sim.next_state)}
        bufs = [[torch.empty_like(t) for t in _state_list(sim.current_state)] for _ in range(ctx.every - 1)]
        for seg in reversed(range(len(ctx.checkpoints))):
            count = min(ctx.every, ctx.steps - seg * ctx.every)
            _load_state(sim.current_state, ctx.checkpoints[seg])
            states = [ctx.checkpoints[seg]]
            for k in range(count - 1):
                sim._params = phase_params[int(seg * ctx.every + k < ctx.rest)]
                _run(sim, 1)
                torch._foreach_copy_(bufs[k], _state_list(sim.current_state))
                states.append(bufs[k])
            for k in reversed(range(count)):
                if k != count - 1:
                    _load_state(sim.current_state, states[k])
                sim._params = phase_params[int(seg * ctx.every + k < ctx.rest)]
                # the substep's forward again: its output, grid and P2G tiles for the adjoint
                sim._launch_substep()
                s_in, s_out = structs[id(sim.current_state)], structs[id(sim.next_state)]
                _native.check(lib.bk_adjoint_substep(ctypes.byref(sim._params), ctypes.byref(sim._scene), ctypes.byref(s_in),
                                                     ctypes.byref(s_out), ctypes.byref(sim._grid), sim._workspace,
                                                     ctypes.byref(bk_adj), _native.ptr(scratch), _native.ptr(param), stream),
                              "bk_adjoint_substep")
        del bufs
        sim.current_state, sim.next_state, sim.step_count, final, flags, sim._params = keep
        _load_state(sim.current_state, final)
        sim._write_flags(flags)
        out = [adj[f].reshape(shape).to(dtype) for f, shape, dtype in zip(fields, ctx.shapes, ctx.dtypes)]
        pgrads = [param[k:k + 1].reshape(shape).to(dtype) if shape.numel() else None
                  for k, (shape, dtype) in enumerate(ctx.params)]
        return (None, None, None, None, *pgrads, *out)  # sim, steps, every, fields; parameters; fields


def rollout(sim, steps, inputs=None, viscosity_parameter=None, production_parameter=None, heat_transfer_parameter=None,
            oven_temperature=None, checkpoint_every=16):
    """Run ``steps`` substeps of ``sim`` from ``inputs`` and return the final state as differentiable tensors.

    ``inputs`` maps state fields (``x, v, C, F``; for the Ding model also ``Fe, porosity, tau``, ``moles``
    with pore gas; with transport ``T, W, losses, exchanges``) to tensors of the state's shapes (N x 3,
    N x 3, N x 3 x 3, ...); fields left out start from the simulator's current state.
    ``viscosity_parameter`` (Ding model), ``production_parameter`` (pore gas) and
    ``heat_transfer_parameter`` (transport), one-element tensors, replace the material's and the
    transport's parameters; ``oven_temperature`` (transport), a one-element tensor, replaces the transport
    config's ``oven_temperature_k`` (the tray keeps its own temperature).  Returns {field: tensor} for every field; the simulator is left at the final
    state, as after ``Advance(steps)``.
    """
    if steps < 1:
        raise ValueError("steps must be positive")
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every must be positive")
    inputs = dict(inputs or {})
    fields = _fields_of(sim)
    unknown = set(inputs) - set(fields)
    if unknown:
        raise ValueError(f"rollout: no adjoint for fields {sorted(unknown)}; differentiable fields are {fields}")
    current = state_field_tensors(sim.current_state)
    tensors = []
    for f in fields:
        t = inputs.get(f)
        if t is None:
            t = current[f].detach().clone()
        elif t.numel() != current[f].numel():
            raise ValueError(f"rollout: input {f} has {t.numel()} elements, the state has {current[f].numel()}")
        tensors.append(t)
    dev = current["x"].device
    if isinstance(sim.material, DingMaterial):
        visc = viscosity_parameter if viscosity_parameter is not None else sim.material.viscosity_parameter.detach().clone()
    else:
        if viscosity_parameter is not None:
            raise ValueError("rollout: viscosity_parameter needs the Ding model")
        visc = torch.zeros(0, device=dev)
    if _gas(sim):
        prod = production_parameter if production_parameter is not None else sim.material.production_parameter.detach().clone()
    else:
        if production_parameter is not None:
            raise ValueError("rollout: production_parameter needs the Ding model with pore gas")
        prod = torch.zeros(0, device=dev)
    if _transport(sim):
        heat = (heat_transfer_parameter if heat_transfer_parameter is not None
                else sim.transport.heat_transfer_parameter.detach().clone())
    else:
        if heat_transfer_parameter is not None:
            raise ValueError("rollout: heat_transfer_parameter needs transport")
        heat = torch.zeros(0, device=dev)
    if oven_temperature is not None and not _transport(sim):
        raise ValueError("rollout: oven_temperature needs transport")
    oven = oven_temperature if oven_temperature is not None else torch.zeros(0, device=dev)
    outs = _Rollout.apply(sim, int(steps), int(checkpoint_every), fields, visc, prod, heat, oven, *tensors)
    return dict(zip(fields, outs))
