"""Many independent doughs advanced by one fused substep (no reference counterpart).

The reference steps one dough per ``MpmSimulator`` (physics/mpm.py:163-225).
``BatchSimulator`` takes several fully built simulators and packs their
particles, tiled grids, meshes and boundary faces into one set of device
buffers and one dough table, so each stage of a substep is a single launch
over every dough (``bk_batch_substep``) and a whole substep replays as one HIP
graph.  Small doughs no longer leave most of the GPU idle between launches.

Each dough evolves as it would alone: the kernels are the single-dough
kernels reading their dough's parameters from the table, each dough keeps its
own grid region and its own error flags.  After packing, every simulator's
state, next state and grid are views into the batch buffers, so the
per-dough reference API (``CheckState``, ``CaptureSnapshot``, ``UpdateMesh``,
``Masses``, the state tensors) keeps working on each dough; step the doughs
through the batch only.

Doughs that are still resting before baking (``physics/mpm.py:57-109``) keep
resting in the batch: each has its own settling-controller record in the
batch workspace, and ``Advance`` ends a chunk of substeps where a dough's
rest could time out, as ``MpmSimulator.Advance`` does.  Changes of a dough's
``transport.config`` after packing are not followed.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _native
from .mpm import _state_tensors, state_field_slots

_THERMAL_GRID = ("V", "C", "E", "W", "H", "L", "X", "dT", "dW")
_FLAG_NAMES = ("J", "stencil", "porosity", "coupled", "gas_cfl", "heat", "water")


class BatchSimulator:
    use_graphs = True
    graph_steps = 4  # substeps per graph launch in runs of substeps (even; 1: one launch per substep)

    def __init__(self, simulators, tuning=None):
        lib = _native.load()
        sims = list(simulators)
        if not sims:
            raise ValueError("BatchSimulator needs at least one simulator")
        self.doughs = sims
        self.dough_count = nd = len(sims)
        self.device_name = dev = sims[0].device_name
        for s in sims:
            if s.device_name != dev:
                raise ValueError("all doughs of a batch must live on one device")
            s._destroy_graphs()
        counts = [s.particle_data.particle_count for s in sims]
        blocks = [((s.simulation_config.grid_resolution + 3) // 4) ** 3 for s in sims]
        self.particle_offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.block_offsets = np.concatenate([[0], np.cumsum(blocks)]).astype(np.int64)
        self.particle_count = n = int(self.particle_offsets[-1])
        total_blocks = int(self.block_offsets[-1])
        if n >= 2 ** 31 or total_blocks * 64 >= 2 ** 32:
            raise ValueError("batch too large for 32-bit particle and node indices")
        self._states = [self._pack_state("current_state"), self._pack_state("next_state")]
        self._cur = 0
        self._pack_grid(total_blocks)
        self._pack_scene(counts)
        self.step_count = 0
        self._graphs = {}

        table = (_native.BkDough * nd)()
        for d, s in enumerate(sims):
            table[d].params = s._make_params()  # settling 2 for a dough still resting
            table[d].particle_offset = int(self.particle_offsets[d])
            table[d].block_offset = int(self.block_offsets[d])
        self._table = table
        ws = lib.bk_workspace_create_batch(n, total_blocks, nd, self.vertex_count, self.slot_count)
        if not ws:
            _native.check(-2, "bk_workspace_create_batch")
        self._workspace = ctypes.c_void_p(ws)
        _native.check(lib.bk_workspace_set_doughs(self._workspace, table, nd), "bk_workspace_set_doughs")
        # a dough that raises a flag stops on the device at its last accepted state; the others run on
        _native.check(lib.bk_workspace_set_freezing(self._workspace, 1), "bk_workspace_set_freezing")
        if tuning is not None:
            self.SetTuning(**tuning)
        self._flags_host = torch.zeros(nd * 8, dtype=torch.int32).pin_memory()
        self._settle_reference = None
        if any(s.IsSettling for s in sims):
            self._pack_settling()

    def _pack_settling(self):
        """Move the resting doughs' controller records and reference plastic gradients into the batch."""
        lib = _native.load()
        nd, n = self.dough_count, self.particle_count
        self._settle_reference = torch.zeros(2, n, 3, 3, dtype=torch.float32, device=self.device_name)
        _native.check(lib.bk_workspace_settling_reference(self._workspace, _native.ptr(self._settle_reference)),
                      "bk_workspace_settling_reference")
        records = (_native.BkSettling * nd)()
        for d, s in enumerate(self.doughs):
            sl = self.dough_slice(d)
            if s.IsSettling and s._settle_record is not None:
                ctypes.memmove(ctypes.byref(records[d]), ctypes.byref(s._settle_record), ctypes.sizeof(_native.BkSettling))
                if s._settle_reference is not None:
                    self._settle_reference[:, sl] = s._settle_reference
            else:
                records[d].phase, records[d].time_step = 0, s.simulation_config.time_step
            if s._settle_reference is not None:
                s._settle_reference = self._settle_reference[:, sl]
        _native.check(lib.bk_workspace_settling_set(self._workspace, 0, nd, records, _native.stream_ptr()),
                      "bk_workspace_settling_set")

    def _read_settling(self):
        records = (_native.BkSettling * self.dough_count)()
        _native.check(_native.load().bk_workspace_settling_get(self._workspace, 0, self.dough_count, records,
                                                               _native.stream_ptr()), "bk_workspace_settling_get")
        return records

    def _pull_settling(self, resting):
        """Mirror the records of the doughs that rested into their simulators; raise for timed-out rests."""
        records = self._read_settling()
        timed_out = []
        for d in resting:
            s, rec = self.doughs[d], records[d]
            if rec.step_count != s.step_count if rec.phase else rec.step_count > s.step_count:  # counts resting substeps
                raise RuntimeError(f"settling controller of dough {d} is at substep {rec.step_count}, the dough at {s.step_count}")
            s._settle_record = _native.BkSettling.from_buffer_copy(rec)
            s._apply_settling(rec)
            if rec.reason == 3 and rec.phase == 1:
                timed_out.append(d)
        if self._table_rests and not any(s.IsSettling for s in self.doughs):
            self._refresh_table()  # every dough bakes: drop the controller from the substep
        if timed_out:
            raise RuntimeError(f"Settling did not reach quasi-static tolerances before settling_duration_s in doughs {timed_out}; "
                               "baking was not started. Increase the maximum duration or review tolerances/material.")

    @property
    def _table_rests(self):
        return any(self._table[d].params.settling == 2 for d in range(self.dough_count))

    def _refresh_table(self):
        for d, s in enumerate(self.doughs):
            self._table[d].params = s._make_params()
        _native.check(_native.load().bk_workspace_set_doughs(self._workspace, self._table, self.dough_count),
                      "bk_workspace_set_doughs")
        self._destroy_graphs()

    # ------------------------------------------------------------------ packing
    def _pack_state(self, which):
        """Concatenate one ping-pong side of every dough; rebind each dough's tensors to views of it."""
        dev = self.device_name
        slots = [state_field_slots(getattr(s, which)) for s in self.doughs]
        packed = {}
        for field in _native.STATE_FIELDS:
            owners = [sl[field] for sl in slots if field in sl]
            if not owners:
                continue
            proto = getattr(*owners[0])
            parts = []
            for d, sl in enumerate(slots):
                if field in sl:
                    parts.append(getattr(*sl[field]))
                else:
                    m = int(self.particle_offsets[d + 1] - self.particle_offsets[d])
                    parts.append(torch.zeros((m,) + tuple(proto.shape[1:]), dtype=proto.dtype, device=dev))
            packed[field] = torch.cat(parts).contiguous()
            for d, sl in enumerate(slots):
                if field in sl:
                    obj, name = sl[field]
                    setattr(obj, name, packed[field][self.particle_offsets[d]:self.particle_offsets[d + 1]])
        return packed

    def _pack_grid(self, total_blocks):
        dev = self.device_name
        nodes = total_blocks * 64
        self.thermal = any(s.transport is not None for s in self.doughs)
        g = dict(m=torch.zeros(nodes, dtype=torch.float32, device=dev),
                 p=torch.zeros(nodes, 3, dtype=torch.float32, device=dev),
                 v=torch.zeros(nodes, 3, dtype=torch.float32, device=dev))
        if self.thermal:
            g.update((k, torch.zeros(nodes, dtype=torch.float64, device=dev)) for k in _THERMAL_GRID)
        self.grid = g
        grid = _native.BkGrid()
        for k, t in g.items():
            setattr(grid, k, t.data_ptr())
        self._grid = grid
        for d, s in enumerate(self.doughs):
            lo, hi = int(self.block_offsets[d]) * 64, int(self.block_offsets[d + 1]) * 64
            s.grid_state.tiled_masses = g["m"][lo:hi]
            s.grid_state.tiled_momenta = g["p"][lo:hi]
            s.grid_state.tiled_velocities = g["v"][lo:hi]
            s.thermal_grid = {k: g[k][lo:hi] for k in s.thermal_grid}
            s._refresh_grid_struct()

    def _pack_scene(self, counts):
        dev = self.device_name
        sims = self.doughs
        keep = {}
        i32 = dict(dtype=torch.int32, device=dev)

        def cat(name, dtype, shape=()):
            parts = []
            for d, s in enumerate(sims):
                t = s._scene_tensors.get(name)
                parts.append(t if t is not None else torch.zeros((counts[d],) + shape, dtype=dtype, device=dev))
            return torch.cat(parts).contiguous()

        keep["V0"] = cat("V0", torch.float32)
        keep["m0"] = cat("m0", torch.float32)
        keep["capacities"] = cat("capacities", torch.float64)
        keep["dry"] = cat("dry", torch.float64)
        for name in ("visc_param", "prod_param", "heat_param"):
            keep[name] = torch.stack([s._scene_tensors[name].reshape(-1)[0].float() if name in s._scene_tensors
                                      else torch.zeros((), dtype=torch.float32, device=dev) for s in sims]).contiguous()
        keep["dough_of"] = torch.repeat_interleave(torch.arange(len(sims), **i32),
                                                   torch.as_tensor(counts, device=dev)).to(torch.int32).contiguous()
        # meshes and boundary faces: concatenated with global indices
        ref_v, b_off, b_part, b_w, b_ref, faces, s_off, s_faces, s_owner, p_slot = ([] for _ in range(10))
        nv = nb = nf = ns = nsf = 0
        for d, s in enumerate(sims):
            p0 = int(self.particle_offsets[d])
            sc = s._scene_tensors
            if s.transport is None:
                p_slot.append(torch.full((counts[d],), -1, **i32))
                continue
            ref_v.append(sc["ref_vertices"])
            b_off.append(sc["bind_offsets"][:-1] + nb)
            b_part.append(sc["bind_particles"] + p0)
            b_w.append(sc["bind_weights"])
            b_ref.append(sc["bind_ref_offsets"])
            faces.append(sc["faces"] + nv)
            s_off.append(sc["slot_face_offsets"][:-1] + nsf)
            s_faces.append(sc["slot_faces"] + nf)
            s_owner.append(sc["slot_owner"] + p0)
            ps = sc["particle_slot"]
            p_slot.append(torch.where(ps >= 0, ps + ns, ps))
            nv += sc["ref_vertices"].shape[0]
            nb += int(sc["bind_offsets"][-1])
            nf += sc["faces"].shape[0]
            ns += sc["slot_owner"].shape[0]
            nsf += int(sc["slot_face_offsets"][-1])
        self.vertex_count, self.slot_count = nv, ns
        keep["particle_slot"] = torch.cat(p_slot).to(torch.int32).contiguous()
        if ref_v:
            keep["ref_vertices"] = torch.cat(ref_v).contiguous()
            keep["bind_offsets"] = torch.cat(b_off + [torch.tensor([nb], **i32)]).to(torch.int32).contiguous()
            keep["bind_particles"] = torch.cat(b_part).to(torch.int32).contiguous()
            keep["bind_weights"] = torch.cat(b_w).contiguous()
            keep["bind_ref_offsets"] = torch.cat(b_ref).contiguous()
            keep["faces"] = torch.cat(faces).to(torch.int32).contiguous()
            keep["slot_face_offsets"] = torch.cat(s_off + [torch.tensor([nsf], **i32)]).to(torch.int32).contiguous()
            keep["slot_faces"] = torch.cat(s_faces).to(torch.int32).contiguous()
            keep["slot_owner"] = torch.cat(s_owner).to(torch.int32).contiguous()
        self._scene_tensors = keep
        scene = _native.BkScene()
        for name in _native.SCENE_FIELDS:
            setattr(scene, name, keep[name].data_ptr() if name in keep else None)
        self._scene = scene

    def SetSurrogate(self, using):
        """Per-dough network mode of the coupled-porosity solve: 0 none, 1 guided, 2 stand-in (see
        baking_amd.surrogate.attach)."""
        for d, on in enumerate(using):
            mode = int(on)
            if mode not in (0, 1, 2):
                raise ValueError("surrogate mode must be 0, 1 or 2")
            self.doughs[d]._surrogate_on = mode
        self._refresh_table()

    # ------------------------------------------------------------------ tuning
    def SetTuning(self, tile_chunk=None, particle_block=None, grid_cap=None, deferred_wgs=None):
        """Launch shapes of the batch substep (see ``bk_tuning``); captured graphs are dropped."""
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

    # ------------------------------------------------------------------ stepping
    def _state_struct(self, k):
        s = _native.BkState()
        for field, t in self._states[k].items():
            setattr(s, field, t.data_ptr())
        return s

    def _destroy_graphs(self):
        lib = _native.load()
        for handle in self._graphs.values():
            lib.bk_graph_destroy(handle)
        self._graphs = {}

    def _graph(self, cur, pipelined, steps=1):
        """The captured graph of ``steps`` substeps from state ``cur`` (alternating buffers; plain or pipelined),
        captured on first use."""
        key = (cur, bool(pipelined), steps)
        graph = self._graphs.get(key)
        if graph is None:
            lib = _native.load()
            s_in, s_out = self._state_struct(cur), self._state_struct(1 - cur)
            handle = lib.bk_batch_graph_create_steps(self._workspace, ctypes.byref(self._scene), ctypes.byref(s_in),
                                                     ctypes.byref(s_out), ctypes.byref(self._grid), steps, int(pipelined))
            if not handle:
                _native.check(-2, "bk_batch_graph_create_steps")
            graph = self._graphs[key] = ctypes.c_void_p(handle)
        return graph

    def _ensure_graphs(self):
        """Capture both ping-pong directions, plain and pipelined, one substep and graph_steps, now (capture
        launches nothing)."""
        for cur in (0, 1):
            for pipelined in (False, True):
                self._graph(cur, pipelined)
            if self.graph_steps > 1:
                self._graph(cur, True, self.graph_steps)

    def _launch(self, stream=None, use_graph=None, pipelined=False):
        """One substep.  ``pipelined`` (graphs only): beside the previous substep's deferred return maps, leaving
        its own pending (see ``_launch_substeps``)."""
        lib = _native.load()
        use_graph = self.use_graphs if use_graph is None else use_graph
        if use_graph:
            _native.check(lib.bk_graph_launch(self._graph(self._cur, pipelined), _native.stream_ptr(stream)), "bk_graph_launch")
        else:
            s_in, s_out = self._state_struct(self._cur), self._state_struct(1 - self._cur)
            _native.check(lib.bk_batch_substep(self._workspace, ctypes.byref(self._scene), ctypes.byref(s_in),
                                               ctypes.byref(s_out), ctypes.byref(self._grid), _native.stream_ptr(stream)),
                          "bk_batch_substep")
        self._cur ^= 1

    def _sync_doughs(self, steps):
        for s in self.doughs:
            if steps % 2:
                s.current_state, s.next_state = s.next_state, s.current_state
            s.step_count += steps

    def ReadFlags(self):
        """Blocking read of every dough's flag block: int array (doughs, 8), reference indices."""
        lib = _native.load()
        _native.check(lib.bk_workspace_read_flags_all(self._workspace, ctypes.c_void_p(self._flags_host.data_ptr()),
                                                      self.dough_count, _native.stream_ptr()), "bk_workspace_read_flags_all")
        return self._flags_host.numpy().reshape(self.dough_count, 8).copy()

    def ClearFlags(self):
        lib = _native.load()
        zeros = (ctypes.c_int32 * (8 * self.dough_count))()
        _native.check(lib.bk_workspace_write_flags_all(self._workspace, ctypes.cast(zeros, ctypes.c_void_p), self.dough_count,
                                                       _native.stream_ptr()), "bk_workspace_write_flags_all")

    def Advance(self, step_count, check=True, stream=None):
        """Run ``step_count`` substeps of every dough, then read the flags once.

        Flags are per dough and sticky.  A dough that raises a flag stops on the device from the next
        substep on (no keys, tiles, grid or material work for it) and keeps the state it had before the
        substep that raised the flag -- the reference rejects that substep (physics/mpm.py:205-207) -- with
        its ``step_count`` at its last accepted substep; the other doughs run on.  With ``check``, a flag
        raised during the call makes this raise ``RuntimeError`` naming the doughs and their flags.
        Returns the (doughs, 8) flag array.
        """
        if step_count < 1:
            raise ValueError("step_count must be positive")
        remaining = step_count
        while True:
            # a resting dough can time out only at its settling_steps: end the chunk there
            chunk = min(s._settling_chunk(remaining) for s in self.doughs)
            flags = self._advance_chunk(chunk, check, stream)
            remaining -= chunk
            if remaining == 0:
                return flags

    def FreezeModes(self):
        """(doughs, 2) int array: each dough's stop mode after the last substep (0 running, 1 stopped in that
        substep, 2 stopped earlier) and the substeps run since it stopped."""
        lib = _native.load()
        host = (ctypes.c_int32 * (2 * self.dough_count))()
        _native.check(lib.bk_workspace_freeze_modes(self._workspace, ctypes.cast(host, ctypes.c_void_p), self.dough_count, 0,
                                                    _native.stream_ptr()), "bk_workspace_freeze_modes")
        return np.frombuffer(host, dtype=np.int32).reshape(self.dough_count, 2).copy()

    def _stop_flagged(self, before, flags, step_count):
        """After a chunk of ``step_count`` substeps: a dough stopped before the chunk stays at its step count; one
        that raised its first flag in the chunk takes the state and step count of its last accepted substep."""
        for d in np.flatnonzero(before.any(axis=1)):
            self.doughs[d].step_count -= step_count
        new = flags.any(axis=1) & ~before.any(axis=1)
        if not new.any():
            return
        modes = self.FreezeModes()
        restore = False
        for d in np.flatnonzero(new):
            s = self.doughs[d]
            mode, age = int(modes[d, 0]), int(modes[d, 1])
            if mode == 0:
                # the flag came in the chunk's last substep: its input is the other buffer
                for dst, src in zip(_state_tensors(s.current_state), _state_tensors(s.next_state)):
                    dst.copy_(src)
                modes[d] = (1, 0)
                restore = True
            s.step_count -= age + 1  # the substeps after the rejected one and the rejected one itself
        if restore:
            lib = _native.load()
            host = (ctypes.c_int32 * (2 * self.dough_count))(*modes.reshape(-1).tolist())
            _native.check(lib.bk_workspace_freeze_modes(self._workspace, ctypes.cast(host, ctypes.c_void_p), self.dough_count, 1,
                                                        _native.stream_ptr()), "bk_workspace_freeze_modes")

    def _advance_chunk(self, step_count, check, stream):
        resting = [d for d, s in enumerate(self.doughs) if s.IsSettling]
        before = self.ReadFlags()
        self._launch_substeps(step_count, stream)
        self.step_count += step_count
        self._sync_doughs(step_count)
        flags = self.ReadFlags()
        self._stop_flagged(before, flags, step_count)
        for d, s in enumerate(self.doughs):
            s._sync_api_flags(flags[d].tolist())
        if check:
            new = (flags != 0) & (before == 0)
            if new.any():
                bad = {int(d): [_FLAG_NAMES[i] for i in np.flatnonzero(new[d, :7])] for d in np.flatnonzero(new.any(axis=1))}
                raise RuntimeError(f"batched substep flags raised in doughs {bad}; each stopped at its last accepted substep")
        if resting:
            self._pull_settling(resting)
        return flags

    def Step(self):
        return self.Advance(1)

    def CheckState(self, dough):
        """The reference's ``CheckState`` (physics/mpm.py:243-308) for one dough, with the batch's flags for it."""
        s = self.doughs[dough]
        s._write_flags(self.ReadFlags()[dough].tolist())
        return s.CheckState()

    # ------------------------------------------------------------------ autotuner hooks
    def _launch_substeps(self, count, stream=None):
        """``count`` substeps.  As HIP graphs, a run of two or more is pipelined (each substep's transport pre-pass
        and P2G set-up beside the previous substep's deferred return maps; the last one's are finished here), so
        the state is complete in stream order when this returns either way."""
        if count < 2 or not self.use_graphs:
            for _ in range(count):
                self._launch(stream)
            return
        lib = _native.load()
        k = self.graph_steps
        while k > 1 and count >= k:  # k substeps per graph launch
            _native.check(lib.bk_graph_launch(self._graph(self._cur, True, k), _native.stream_ptr(stream)), "bk_graph_launch")
            self._cur ^= k & 1
            count -= k
        for _ in range(count):
            self._launch(stream, pipelined=True)
        s_in, s_out = self._state_struct(1 - self._cur), self._state_struct(self._cur)  # the last substep's
        _native.check(lib.bk_batch_finish(self._workspace, ctypes.byref(self._scene), ctypes.byref(s_in), ctypes.byref(s_out),
                                          _native.stream_ptr(stream)), "bk_batch_finish")

    def _save_for_tuning(self):
        flags = self.ReadFlags()
        settling = None
        if self._settle_reference is not None:
            settling = (self._read_settling(), self._settle_reference.clone())
        return self._cur, flags, [t.clone() for side in self._states for t in side.values()], settling

    def _restore_after_tuning(self, saved):
        cur, flags, clones, settling = saved
        if settling is not None:
            records, reference = settling
            self._settle_reference.copy_(reference)
            _native.check(_native.load().bk_workspace_settling_set(self._workspace, 0, self.dough_count, records,
                                                                   _native.stream_ptr()), "bk_workspace_settling_set")
        torch.cuda.synchronize()
        for dst, src in zip([t for side in self._states for t in side.values()], clones):
            dst.copy_(src)
        for t in self.grid.values():
            t.zero_()
        self._cur = cur
        arr = (ctypes.c_int32 * flags.size)(*flags.reshape(-1).tolist())
        _native.check(_native.load().bk_workspace_write_flags_all(self._workspace, ctypes.cast(arr, ctypes.c_void_p),
                                                                  self.dough_count, _native.stream_ptr()),
                      "bk_workspace_write_flags_all")

    # ------------------------------------------------------------------ views
    def dough_slice(self, dough):
        return slice(int(self.particle_offsets[dough]), int(self.particle_offsets[dough + 1]))

    def __del__(self):
        try:
            self._destroy_graphs()
            if getattr(self, "_workspace", None) is not None:
                _native.load().bk_workspace_destroy(self._workspace)
        except Exception:
            pass
