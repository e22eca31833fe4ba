"""CryspLight optical-photon transport on AMD MI355X (gfx950).

The compute path is libcrysplight_hip.so (hand-written HIP kernels behind the C-ABI of
include/crysplight.h); this package is the host side that mirrors the reference's
Julia entry points (run_photons_ka!, run_photons!, run_events!, load_setup,
run_from_config).
"""
from .params import (Accumulator, Box, Emission, OpticalParams, Readout, SipmGrid,
                     TimeBinning, readout, deg2rad, STATUS_DETECTED, STATUS_ABS_BULK,
                     STATUS_ABS_WALL, STATUS_ABS_SIPM, STATUS_CAP, STATUS_ABS_SUR,
                     FACE_XP, FACE_XM, FACE_YP, FACE_YM, FACE_BACK, FACE_FRONT)
from .transport import (Records, run_photons, run_photons_ka, transport_records,
                        transport_accumulate, reduce_records, debug_draws)
from ._lib import Context, CryspError, default_context

__all__ = [n for n in dir() if not n.startswith("_")]
