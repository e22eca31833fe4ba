"""Frozen algorithmic-operation model of the transport (the roofline's numerator).

Each IEEE operation and each transcendental counts as one lane-op; the weights are the
operation counts of the reference's per-event work (src/philox.jl:10-28 for a Philox
block, src/geometry.jl:19-32 + src/transport.jl:127-159 for a step, src/optics.jl for
the samplers).  They are fixed here and never tuned to the kernel.

  ops = 98 N_block + 4 N_draw + 31 N_step + 26 N_fresnel + 64 N_surface_try
        + 10 N_lambert + 60 N_rayleigh_try + 14 N_photon + 10 N_detect
"""
from __future__ import annotations

# index order of the device/oracle op counters
COUNTER_NAMES = ("n_photon", "n_draw", "n_block", "n_step", "n_fresnel", "n_surface_try",
                 "n_lambert", "n_rayleigh_try", "n_detect", "n_nscat")

WEIGHTS = {"n_block": 98, "n_draw": 4, "n_step": 31, "n_fresnel": 26, "n_surface_try": 64,
           "n_lambert": 10, "n_rayleigh_try": 60, "n_photon": 14, "n_detect": 10}

# MI355X VALU issue ceiling for one-op-per-lane code: 256 CU x 64 lanes x 2.4 GHz
VALU_PEAK_OPS = 256 * 64 * 2.4e9
FP32_NAMEPLATE_FLOPS = 157.3e12


def algorithmic_ops(counts) -> int:
    """counts: mapping name -> value, or a sequence in COUNTER_NAMES order."""
    if not hasattr(counts, "keys"):
        counts = dict(zip(COUNTER_NAMES, (int(v) for v in counts)))
    return int(sum(w * int(counts[k]) for k, w in WEIGHTS.items()))
