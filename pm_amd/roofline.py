"""Frozen algorithmic FLOP counts per contribution and MI355X peaks.

A contribution is one (trace, output sample) pair that passes the
time-dependent aperture and the sample-bounds test (E).  Counting convention:
FMA = 2, add/sub/mul = 1, sqrt/rsq/rcp/log/exp = 1, compares/floor/convert = 0.

  straight ray, spreading + obliquity: 2 FMA radicands (4) + 2 sqrt (2) + add
    (1) + FMA sample index (2) + frac (1) + lerp sub+FMA (3) + rcp(t) (1) +
    2 mul for t0/(v t^2) (2) + mul pair weight (1) + FMA accumulate (2) = 19;
    with both weights off 16.
  curved ray 25, VTI 31, triangle anti-aliasing 47, MLP 5-64-64-64-64-2
    25,472 per evaluation: fixed before their first measurement and kept
    regardless of how the kernels evaluate them.
  learned VTI traveltime, VALU part of a contribution: the learned factor
    of each leg from the cubic of its lattice segment (Horner, 3 FMA per leg:
    12) + per leg h s (1), 1 + eta g (2), FMA radicand (2) (10) + 2 sqrt +
    add (3) + the straight-ray tail (sample index FMA 2, frac 1, lerp 3,
    rcp 1, t0/(v t^2) 2, pair weight 1, accumulate FMA 2: 12) = 37; the
    network evaluations on the lattice are MFMA work, counted apart.
"""

from __future__ import annotations

FLOPS_PER_CONTRIBUTION = {
    "straight": 19,
    "straight_noweights": 16,
    "curved": 25,
    "vti": 31,
    "triangle_aa": 47,
    "vti_learned": 37,
}
MLP_FLOPS_PER_EVAL = 2 * (5 * 64 + 3 * 64 * 64 + 64 * 2)  # 25,472

MI355X_FP32_VALU_TFLOPS = 157.3        # packed FP32 VALU peak
MI355X_FP32_VALU_UNPACKED_TFLOPS = 78.6
MI355X_BF16_MFMA_TFLOPS = 2500.0       # dense
MI355X_HBM_GBS = 8000.0


def flops_per_contribution(traveltime: str, antialias: str, weights: bool) -> int:
    if antialias == "triangle":
        return FLOPS_PER_CONTRIBUTION["triangle_aa"]
    if traveltime == "curved_ray":
        return FLOPS_PER_CONTRIBUTION["curved"]
    if traveltime == "anisotropic_vti":
        return FLOPS_PER_CONTRIBUTION["vti"]
    return FLOPS_PER_CONTRIBUTION["straight" if weights else "straight_noweights"]
