"""End-state porosity consistency of the Ding return map (reference ``physics/materials/porous_return.py:1-100``).

``solve_porosity`` evaluates the device solver (trial early exit, else 64
bisection steps over an fp64 Newton overstress solve) for arrays of inputs.
"""
import torch

from ...launch import Kernel

_SolvePorosity = Kernel("bk_solve_porosity", ["f32", "f64", "f32", "f64", "i"])


def solve_porosity(q_value, p_value, dev_factor, mean_scale, exponent, plastic_volume, initial_porosity, solver="bisection"):
    """Return (porosity float32, porosity_prediction(porosity) float64) for broadcast tensor inputs.

    solver="bisection" runs the reference's 64 bisection steps; "bracketed" the
    safeguarded Newton solve of the same equation used by the fused substep.
    """
    cols = torch.broadcast_tensors(*[torch.as_tensor(v, dtype=torch.float32, device="cuda")
                                     for v in (q_value, p_value, dev_factor, mean_scale, exponent, initial_porosity)])
    args = torch.stack([c.reshape(-1) for c in cols], dim=1).contiguous()
    pv = torch.as_tensor(plastic_volume, dtype=torch.float64, device="cuda").expand(args.shape[0]).contiguous()
    out = torch.empty(args.shape[0], dtype=torch.float32, device="cuda")
    pred = torch.empty(args.shape[0], dtype=torch.float64, device="cuda")
    _SolvePorosity(args.shape[0], args, pv, out, pred, {"bisection": 0, "bracketed": 1}[solver])
    return out, pred
