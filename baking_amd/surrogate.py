"""Return-map surrogate network (no reference counterpart).

The coupled (gas) Ding return map (reference physics/materials/ding.py:168-221
with porous_return.py:43-61) solves a nested fp64 equation for the porosity
after each yielding substep.  A small network -- 7 inputs, 4 hidden layers of
64 SiLU units, 4 outputs -- predicts that returned state from the trial state:

* inputs: the three principal trial log-strains (descending), the temperature
  the yield and viscosity factors are taken at, the saturation, the porosity
  before the substep, |Fp - I|_F;
* outputs: the three returned log-strains (same order) and the porosity
  increment of the substep.

In the substep the network runs on the matrix cores in BF16 with FP32
accumulation over the queue of particles that need the coupled solve, and
each solve then starts from its prediction (porosity, and the overstress that
maps the trial deviator onto the predicted one) with plain Newton steps.  A
solve is accepted only when a Newton step falls below 1e-10, which pins the
float32 result to the one the exact solve returns; otherwise that particle
runs the exact bracketed solve.  The network therefore changes how fast the
return map converges, not what it converges to.

Training data are rows the exact solve emits during a simulation
(``collect_samples``); training runs on the GPU with the library's own FP32
forward/backward kernels and Adam (``ReturnMapSurrogate.fit``).
"""
from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native

N_IN, ROW, WIDTH, N_OUT = 7, 8, 64, 4
SAMPLE_ROW = 12  # 7 inputs, 4 exact outputs, 1 unused
LAYERS = (("W1", (WIDTH, ROW)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)), ("W3", (WIDTH, WIDTH)),
          ("b3", (WIDTH,)), ("W4", (WIDTH, WIDTH)), ("b4", (WIDTH,)), ("W5", (N_OUT, WIDTH)), ("b5", (N_OUT,)))
THETA_COUNT = sum(int(np.prod(shape)) for _, shape in LAYERS)
DEFAULT_MODEL = Path(__file__).resolve().parent / "data" / "return_map_surrogate.npz"
BATCHED_MODEL = Path(__file__).resolve().parent / "data" / "return_map_surrogate_batched48.npz"


def default_model(workload="loaf_256"):
    """The packaged network of a workload: the loaf's (dt 5e-5), or the batched doughs' (dt 1.25e-5) when it is
    packaged."""
    if workload == "batched_48" and BATCHED_MODEL.exists():
        return BATCHED_MODEL
    return DEFAULT_MODEL

# the material parameters a trained network is valid for (the map depends on them)
SIGNATURE_KEYS = ("shear_modulus", "lame_lambda", "yield_stress", "viscosity", "porosity", "time_step",
                  "temperature_enabled", "temperature_knots_k", "yield_factors", "viscosity_factors")


def theta_views(theta):
    """Named views (W1, b1, ..., W5, b5) into a flat parameter vector."""
    out, off = {}, 0
    for name, shape in LAYERS:
        size = int(np.prod(shape))
        out[name] = theta[off:off + size].reshape(shape)
        off += size
    return out


def init_theta(seed=0):
    """He-style initialisation (normal, std sqrt(1/fan_in)); biases zero; W1's unused input column zero."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(THETA_COUNT, np.float32)
    v = theta_views(theta)
    for name, shape in LAYERS:
        if name.startswith("W"):
            fan_in = N_IN if name == "W1" else shape[1]
            v[name][...] = rng.normal(0.0, math.sqrt(1.0 / fan_in), size=shape).astype(np.float32)
    v["W1"][:, N_IN:] = 0.0
    return theta


def bf16_bits(x):
    """Round-to-nearest-even float32 -> bfloat16 bit patterns (uint16)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def _tiles(W, mus, kts):
    """16x16 A operands of v_mfma_f32_16x16x16_bf16 for a weight matrix [out][in]: tile (mu, kt), lane l, element i holds
    W[16 mu + l % 16, 16 kt + 4 (l // 16) + i] (zero outside W)."""
    pad = np.zeros((16 * mus, 16 * kts), np.float32)
    pad[:W.shape[0], :W.shape[1]] = W
    lane = np.arange(64)
    out = np.empty((mus, kts, 64, 4), np.float32)
    for mu in range(mus):
        for kt in range(kts):
            for i in range(4):
                out[mu, kt, :, i] = pad[16 * mu + lane % 16, 16 * kt + 4 * (lane // 16) + i]
    return bf16_bits(out.reshape(-1))


def _tiles32(W, mus, kcs):
    """16x32 A operands of v_mfma_f32_16x16x32_bf16 for a weight matrix [out][in]: tile (mu, kc), lane l,
    element j holds W[16 mu + l % 16, unit] with unit = 32 kc + 4 (l // 16) + j for j < 4 and
    32 kc + 16 + 4 (l // 16) + j - 4 for j >= 4 (zero outside W): the units the previous layer's output
    tiles 2 kc and 2 kc + 1 leave in lane l."""
    pad = np.zeros((16 * mus, 32 * kcs), np.float32)
    pad[:W.shape[0], :W.shape[1]] = W
    lane = np.arange(64)
    out = np.empty((mus, kcs, 64, 8), np.float32)
    for mu in range(mus):
        for kc in range(kcs):
            for j in range(8):
                unit = 32 * kc + 4 * (lane // 16) + (j if j < 4 else 16 + j - 4)
                out[mu, kc, :, j] = pad[16 * mu + lane % 16, unit]
    return bf16_bits(out.reshape(-1))


def build_blob(theta, in_shift, in_scale, out_shift, out_scale):
    """The device inference blob (layout in baking_amd/csrc/bk_surrogate.h) as a uint8 array."""
    v = theta_views(np.asarray(theta, np.float32))
    head = np.zeros(24, np.float32)
    head[0:7], head[8:15] = in_shift[:N_IN], in_scale[:N_IN]
    head[16:20], head[20:24] = out_shift, out_scale
    bias = np.concatenate([v["b1"], v["b2"], v["b3"], v["b4"], np.pad(v["b5"], (0, 12))]).astype(np.float32)
    w1 = np.pad(v["W1"], ((0, 0), (0, 8)))  # 64 x 16
    parts = [head.view(np.uint8), bias.view(np.uint8), _tiles(w1, 4, 1).view(np.uint8)]
    for name in ("W2", "W3", "W4"):
        parts.append(_tiles32(v[name], 4, 2).view(np.uint8))
    parts.append(_tiles32(v["W5"], 1, 2).view(np.uint8))
    blob = np.concatenate(parts)
    expected = _native.load().bk_mlp_blob_bytes() if _native.library_path().exists() else blob.size
    if blob.size != expected:
        raise RuntimeError(f"surrogate blob is {blob.size} bytes, the library expects {expected}")
    return blob


def material_signature(sim):
    """The material parameters of a simulator the return map depends on (for attach's check)."""
    m = sim.material.material_config
    d = sim.material.ding_config
    return dict(shear_modulus=float(np.float32(m.shear_modulus)), lame_lambda=float(np.float32(m.lame_lambda)),
                yield_stress=float(d.yield_stress), viscosity=float(d.viscosity), porosity=float(d.porosity),
                time_step=float(sim.simulation_config.time_step), temperature_enabled=bool(d.temperature_enabled),
                temperature_knots_k=[float(k) for k in d.temperature_knots_k],
                yield_factors=[float(k) for k in d.yield_factors], viscosity_factors=[float(k) for k in d.viscosity_factors])


@dataclass
class ReturnMapSurrogate:
    theta: np.ndarray
    in_shift: np.ndarray
    in_scale: np.ndarray
    out_shift: np.ndarray
    out_scale: np.ndarray
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self._blob = None

    # ---------------------------------------------------------------- storage
    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, theta=self.theta, in_shift=self.in_shift, in_scale=self.in_scale, out_shift=self.out_shift,
                 out_scale=self.out_scale, meta=np.frombuffer(json.dumps(self.meta).encode(), np.uint8))

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            return cls(z["theta"].astype(np.float32), z["in_shift"].astype(np.float32), z["in_scale"].astype(np.float32),
                       z["out_shift"].astype(np.float32), z["out_scale"].astype(np.float32), meta)

    # ---------------------------------------------------------------- device
    def blob_tensor(self, device="cuda"):
        import torch
        if self._blob is None or self._blob.device.type != torch.device(device).type:
            blob = build_blob(self.theta, self.in_shift, self.in_scale, self.out_shift, self.out_scale)
            self._blob = torch.as_tensor(blob, device=device)
        return self._blob

    def predict(self, rows):
        """BF16 matrix-core outputs for feature rows (torch (n, 8) float32 on the GPU) -> (n, 4)."""
        import torch
        rows = rows.contiguous()
        out = torch.empty(rows.shape[0], N_OUT, dtype=torch.float32, device=rows.device)
        _native.check(_native.load().bk_mlp_infer(_native.ptr(self.blob_tensor(rows.device)), _native.ptr(rows), _native.ptr(out),
                                                  rows.shape[0], _native.stream_ptr()), "bk_mlp_infer")
        return out

    def predict_fp32(self, rows):
        """FP32 outputs of the same network (the training forward) -> (n, 4)."""
        import torch
        dev = rows.device
        x = (rows[:, :ROW].float() - torch.as_tensor(np.pad(self.in_shift, (0, ROW - len(self.in_shift))), device=dev)) \
            * torch.as_tensor(np.pad(self.in_scale, (0, ROW - len(self.in_scale))), device=dev)
        x[:, N_IN:] = 0.0
        x = x.contiguous()
        y = torch.empty(rows.shape[0], N_OUT, dtype=torch.float32, device=dev)
        z = torch.empty(4 * rows.shape[0] * WIDTH, dtype=torch.float32, device=dev)
        th = torch.as_tensor(self.theta, device=dev)
        _native.check(_native.load().bk_mlp_forward(_native.ptr(th), _native.ptr(x), _native.ptr(y), rows.shape[0], _native.ptr(z),
                                                    _native.stream_ptr()), "bk_mlp_forward")
        return y * torch.as_tensor(self.out_scale, device=dev) + torch.as_tensor(self.out_shift, device=dev)

    # ---------------------------------------------------------------- training
    @classmethod
    def fit(cls, samples, epochs=8, batch=16384, lr=2e-3, final_lr=2e-5, seed=0, holdout=0.05, log=None, meta=None):
        """Train on emitted rows (torch (m, 12) float32 on the GPU) with the library's FP32 kernels and Adam.

        Inputs and outputs are normalised by their mean and standard deviation;
        the loss is the mean squared error of the normalised outputs.  The last
        ``holdout`` fraction of a seeded shuffle is kept out of training.
        Returns (model, report) with the held-out error table.
        """
        import torch
        lib = _native.load()
        dev = samples.device
        m = samples.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        perm = torch.randperm(m, generator=gen).to(dev)
        n_hold = max(1, int(m * holdout))
        train, hold = samples[perm[n_hold:]], samples[perm[:n_hold]]
        X, Y = train[:, :N_IN].double(), train[:, N_IN:N_IN + N_OUT].double()
        # a (nearly) constant column keeps a finite scale: its spread is floored at 1e-6 of its magnitude
        x_mean, y_mean = X.mean(0), Y.mean(0)
        in_shift = x_mean.float().cpu().numpy()
        in_scale = (1.0 / torch.maximum(X.std(0), 1e-6 * x_mean.abs() + 1e-30)).float().cpu().numpy()
        out_shift = y_mean.float().cpu().numpy()
        out_scale = torch.maximum(Y.std(0), 1e-6 * y_mean.abs() + 1e-30).float().cpu().numpy()

        def normalise(rows):
            x = torch.zeros(rows.shape[0], ROW, dtype=torch.float32, device=dev)
            x[:, :N_IN] = (rows[:, :N_IN] - torch.as_tensor(in_shift, device=dev)) * torch.as_tensor(in_scale, device=dev)
            y = ((rows[:, N_IN:N_IN + N_OUT] - torch.as_tensor(out_shift, device=dev)) / torch.as_tensor(out_scale, device=dev))
            return x.contiguous(), y.contiguous().float()

        xt, yt = normalise(train)
        n = xt.shape[0]
        batch = min(batch, n)
        theta = torch.as_tensor(init_theta(seed), device=dev)
        mom, vel = torch.zeros_like(theta), torch.zeros_like(theta)
        grad = torch.zeros_like(theta)
        chunk = lib.bk_mlp_grad_chunk()
        z = torch.empty(4 * batch * WIDTH, device=dev)
        d = torch.empty(4 * batch * WIDTH, device=dev)
        y = torch.empty(batch * N_OUT, device=dev)
        d_out = torch.empty(batch * N_OUT, device=dev)
        partial = torch.empty(((batch + chunk - 1) // chunk) * THETA_COUNT, device=dev)
        loss = torch.zeros(1, dtype=torch.float64, device=dev)
        steps_per_epoch = n // batch
        total = epochs * steps_per_epoch
        step = 0
        history = []
        for epoch in range(epochs):
            order = torch.randperm(n, generator=gen).to(dev)
            loss.zero_()
            for k in range(steps_per_epoch):
                idx = order[k * batch:(k + 1) * batch]
                xb, yb = xt[idx].contiguous(), yt[idx].contiguous()
                _native.check(lib.bk_mlp_gradient(_native.ptr(theta), _native.ptr(xb), _native.ptr(yb), batch, _native.ptr(z),
                                                  _native.ptr(y), _native.ptr(d), _native.ptr(d_out), _native.ptr(partial),
                                                  _native.ptr(grad), _native.ptr(loss), _native.stream_ptr()), "bk_mlp_gradient")
                step += 1
                rate = final_lr + 0.5 * (lr - final_lr) * (1.0 + math.cos(math.pi * step / total))
                _native.check(lib.bk_adam_step(_native.ptr(theta), _native.ptr(grad), _native.ptr(mom), _native.ptr(vel),
                                               THETA_COUNT, rate, 0.9, 0.999, 1e-8, step, _native.stream_ptr()), "bk_adam_step")
            history.append(float(loss.item()) / max(steps_per_epoch, 1))
            if log:
                log(f"epoch {epoch + 1}/{epochs}: train mse {history[-1]:.3e} (normalised)")
        model = cls(theta.cpu().numpy(), in_shift, in_scale, out_shift, out_scale,
                    dict(meta or {}, epochs=epochs, batch=batch, lr=lr, seed=seed, samples=int(m), holdout=int(n_hold),
                         train_mse=history))
        report = model.evaluate(hold)
        model.meta["holdout"] = report
        return model, report

    def evaluate(self, rows):
        """Held-out error table of the BF16 inference against the exact outputs of emitted rows."""
        import torch
        x = torch.zeros(rows.shape[0], ROW, dtype=torch.float32, device=rows.device)
        x[:, :N_IN] = rows[:, :N_IN]
        pred = self.predict(x.contiguous()).double()
        exact = rows[:, N_IN:N_IN + N_OUT].double()
        err = (pred - exact).abs()
        names = ("eps_ret_1", "eps_ret_2", "eps_ret_3", "porosity_increment")
        table = {}
        for k, name in enumerate(names):
            scale = float(exact[:, k].abs().max())
            table[name] = dict(mean_abs=float(err[:, k].mean()), p99_abs=float(torch.quantile(err[:, k].float(), 0.99)),
                               max_abs=float(err[:, k].max()), exact_max_abs=scale, exact_std=float(exact[:, k].std()))
        table["rows"] = int(rows.shape[0])
        return table


# ---------------------------------------------------------------- simulation hooks
def collect_samples(sim, steps, capacity, stride=1, log=None):
    """Run ``steps`` substeps of an MpmSimulator or BatchSimulator and return the rows its exact coupled
    return maps emitted (torch (rows, 12) float32), every ``stride``-th substep, at most ``capacity`` rows.

    The device flags are read after every substep.  Collection stops at the first substep that raises
    one, such as the gas-acoustic CFL bound, and that substep's rows are dropped, so every row comes
    from a state the simulator's own checks accept."""
    import torch
    lib = _native.load()
    read_flags = getattr(sim, "_read_flags", None) or sim.ReadFlags
    buf = torch.empty(capacity * SAMPLE_ROW, dtype=torch.float32, device="cuda")
    stats = (ctypes.c_int64 * 5)()
    emitted = 0
    for step in range(steps):
        on = step % stride == 0 and emitted < capacity
        if on:
            _native.check(lib.bk_workspace_set_samples(sim._workspace, _native.ptr(buf[emitted * SAMPLE_ROW:]), capacity - emitted),
                          "bk_workspace_set_samples")
        sim._launch_substeps(1)
        before = emitted
        if on:
            _native.check(lib.bk_workspace_stats(sim._workspace, ctypes.cast(stats, ctypes.c_void_p), 0, _native.stream_ptr()),
                          "bk_workspace_stats")
            emitted = min(capacity, emitted + int(stats[4]))
            _native.check(lib.bk_workspace_set_samples(sim._workspace, None, 0), "bk_workspace_set_samples")
        flags = np.asarray(read_flags())
        if flags.any():
            emitted = before
            if log:
                log(f"substep {step}: device flags {flags.reshape(-1).tolist()} raised; collection stops with {emitted} rows")
            break
        if on and log:
            log(f"substep {step}: {emitted} rows")
    torch.cuda.synchronize()
    return buf[:emitted * SAMPLE_ROW].view(emitted, SAMPLE_ROW).clone()


def _simulators(sim):
    return getattr(sim, "doughs", None) or [sim]


MODES = {"guided": 1, "standin": 2}


def attach(sim, model, strict=True, same_material=False, mode="standin"):
    """Use ``model`` in the coupled solves of ``sim`` (MpmSimulator or BatchSimulator).

    ``mode`` "standin": the network's returned porosity, corrected by one Newton step of the porosity equation,
    is the result when that step is at most 2e-5 (then within about a float ulp of the exact root); otherwise
    the solve runs exactly, starting from the prediction.  "guided": every solve starts from the prediction and
    converges to the exact root, so results equal the exact return map's.
    Every dough with a coupled (gas) Ding return map uses it (``strict``: raise if a dough has none).
    ``same_material`` restricts it to doughs whose material parameters match the training run's.  Returns the
    number of doughs using it.
    """
    if mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    lib = _native.load()
    trained = model.meta.get("signature")
    using = []
    for s in _simulators(sim):
        ok = s.current_state.material_state is not None and s.material.ding_config.gas_enabled
        if strict and not ok:
            raise ValueError("the dough has no coupled (gas) return map to start from the surrogate")
        if ok and same_material and trained is not None:
            ok = material_signature(s) == trained
        using.append(MODES[mode] if ok else 0)
    sim._surrogate_blob = model.blob_tensor("cuda")
    _native.check(lib.bk_workspace_set_surrogate(sim._workspace, _native.ptr(sim._surrogate_blob)), "bk_workspace_set_surrogate")
    _set_surrogate_flags(sim, using)
    return sum(1 for u in using if u)


def detach(sim):
    lib = _native.load()
    _set_surrogate_flags(sim, [0] * len(_simulators(sim)))
    _native.check(lib.bk_workspace_set_surrogate(sim._workspace, None), "bk_workspace_set_surrogate")
    sim._surrogate_blob = None


def _set_surrogate_flags(sim, using):
    if hasattr(sim, "doughs"):
        sim.SetSurrogate(using)
    else:
        sim.SetSurrogate(using[0])


def stats(sim, reset=False):
    """{evaluations, accepted, rejected, return_maps, emitted} of the coupled solves since the last reset."""
    lib = _native.load()
    out = (ctypes.c_int64 * 5)()
    _native.check(lib.bk_workspace_stats(sim._workspace, ctypes.cast(out, ctypes.c_void_p), int(reset), _native.stream_ptr()),
                  "bk_workspace_stats")
    return dict(evaluations=out[0], accepted=out[1], rejected=out[2], return_maps=out[3], emitted=out[4])
