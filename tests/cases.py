"""Parameter sets shared by the tests.

Reference fixtures: GRID/TB (test/runtests.jl:8-9), ideal() (:18-21), the all-processes
configuration of the KA parity test (:227-232).  C1 is the point-source workload of the
benchmark configuration list (CsI(Tl) 25.6x25.6x20 mm, Teflon, grease, 8x8 SiPMs)."""
import math

import numpy as np

from crysplight_amd.params import (Box, OpticalParams, SipmGrid, TimeBinning, readout,
                                   deg2rad)

INF = float("inf")
GRID = SipmGrid(6.0, 8, 8)                  # test/runtests.jl:8
TB = TimeBinning(100.0, 40)                 # test/runtests.jl:9
BOX_CRYSP = Box((48.0, 48.0, 37.2))
POS_CRYSP = (24.0, 24.0, 18.6)


def ideal(n_crystal, wrap_R=1.0, spec=False, abs_len=INF, ray=INF, gap=True, sigma=0.0,
          pde=1.0):
    """test/runtests.jl:18-21"""
    return OpticalParams(n_crystal, 1.45, abs_len, ray, wrap_R, spec, gap, sigma, pde)


# every process on at once (test/runtests.jl:229-231)
OP_ALL = OpticalParams(1.95, 1.45, 500.0, 400.0, 0.98, False, True, deg2rad(6), 0.4)
RO_ALL = readout(GRID, center=(24, 24), radius=20, sur_R=0.9)

# C1: point source in CsI(Tl) 25.6x25.6x20, n=1.79, abs 400 mm, Teflon R=0.98 back-painted
# (sigma_alpha 1.3 deg), grease n=1.5, 8x8 SiPMs at 3.2 mm, PDE 0.3, tau 1000 ns
C1_BOX = Box((25.6, 25.6, 20.0))
C1_OP = OpticalParams(1.79, 1.5, 400.0, INF, 0.98, False, True, deg2rad(1.3), 0.3)
C1_RO = readout(SipmGrid(3.2, 8, 8))
C1_TB = TimeBinning(100.0, 80)
C1_POS = (12.8, 12.8, 10.0)
C1_TAU = 1000.0
C1_SEED = 1234

# C2 optics: BGO 25.6x25.6x20, n=2.15, abs 300 mm, ESR R=0.98 specular, back-painted
C2_BOX = Box((25.6, 25.6, 20.0))
C2_OP = OpticalParams(2.15, 1.5, 300.0, INF, 0.98, True, True, deg2rad(1.3), 0.3)
C2_TAU = 300.0
