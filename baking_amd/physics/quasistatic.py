"""Read-only settling diagnostics (reference ``physics/quasistatic.py:1-21``).

``SettlingMetrics`` reduces, over the particles, the maximum speed, the
maximum norm of the symmetric affine velocity gradient and the maximum
relative plastic-gradient rate; the simulator's settling controller runs the
same reduction fused with the rest/bake decision after every resting substep.
"""
from __future__ import annotations

from ..launch import Kernel

# inputs: velocities (N,3) f32, affine (N,3,3) f32, old_plastic (N,3,3) f32, new_plastic (N,3,3) f32,
# metrics (3,) f32 running maxima, time_step (elapsed time since old_plastic)
SettlingMetrics = Kernel("bk_SettlingMetrics", ["f32", "f32", "f32", "f32", "f32", "f"])
