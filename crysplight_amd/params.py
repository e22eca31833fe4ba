"""Parameter types of the optical transport, byte-compatible with the reference's
isbits structs and with include/crysplight.h.

Reference: src/transport.jl:14-62 (OpticalParams, SipmGrid, Readout, TimeBinning),
src/geometry.jl:11-13 (Box), src/transport.jl:64-94 (Accumulator, merge!,
total_terminated).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

# photon fates, src/transport.jl:6-12
STATUS_DETECTED = 1
STATUS_ABS_BULK = 2
STATUS_ABS_WALL = 3
STATUS_ABS_SIPM = 4
STATUS_CAP = 5
STATUS_ABS_SUR = 6

# face codes, src/geometry.jl:4-9
FACE_XP, FACE_XM, FACE_YP, FACE_YM, FACE_BACK, FACE_FRONT = 1, 2, 3, 4, 5, 6

C_MM_NS = 299.792458  # src/transport.jl:4


class Box(C.Structure):
    """Crystal [0,Lx]x[0,Ly]x[0,Lz] mm; SiPM plane at z = Lz (src/geometry.jl:11-13)."""
    _fields_ = [("L", C.c_float * 3)]

    def __init__(self, L):
        super().__init__((C.c_float * 3)(*[float(v) for v in L]))

    def __repr__(self):
        return f"Box(L={tuple(self.L)})"


class OpticalParams(C.Structure):
    """src/transport.jl:14-24 -- positional order as the Julia constructor."""
    _fields_ = [
        ("n_crystal", C.c_float), ("n_coupling", C.c_float), ("abs_len_mm", C.c_float),
        ("rayleigh_mm", C.c_float), ("wrap_R", C.c_float), ("wrap_specular", C.c_bool),
        ("air_gap", C.c_bool), ("sigma_alpha", C.c_float), ("pde", C.c_float),
    ]

    def __init__(self, n_crystal, n_coupling, abs_len_mm, rayleigh_mm, wrap_R,
                 wrap_specular, air_gap, sigma_alpha, pde):
        super().__init__(float(n_crystal), float(n_coupling), float(abs_len_mm),
                         float(rayleigh_mm), float(wrap_R), bool(wrap_specular),
                         bool(air_gap), float(sigma_alpha), float(pde))

    def __repr__(self):
        return "OpticalParams(" + ", ".join(f"{n}={getattr(self, n)!r}" for n, _ in self._fields_) + ")"


class SipmGrid(C.Structure):
    """src/transport.jl:26-30."""
    _fields_ = [("pitch_mm", C.c_float), ("nx", C.c_int32), ("ny", C.c_int32)]

    def __init__(self, pitch_mm, nx, ny):
        super().__init__(float(pitch_mm), int(nx), int(ny))


class Readout(C.Structure):
    """src/transport.jl:41-50.  Build it with `readout(grid, ...)`, the keyword
    constructor of src/transport.jl:52-57."""
    _fields_ = [
        ("grid", SipmGrid), ("disc", C.c_bool), ("cx", C.c_float), ("cy", C.c_float),
        ("r2", C.c_float), ("sur_R", C.c_float), ("sur_specular", C.c_bool),
        ("two_sided", C.c_bool),
    ]


def readout(grid: SipmGrid, center=None, radius=None, sur_R=0.0, sur_specular=False,
            two_sided=False) -> Readout:
    """Readout(grid; center, radius, sur_R, sur_specular, two_sided)
    (src/transport.jl:52-57): whole face active unless a disc radius is given."""
    r = Readout()
    r.grid = grid
    r.two_sided = bool(two_sided)
    if radius is None:
        r.disc = False
        r.cx = r.cy = r.r2 = r.sur_R = 0.0
        r.sur_specular = False
    else:
        r.disc = True
        r.cx, r.cy = float(center[0]), float(center[1])
        rf = np.float32(radius)
        r.r2 = float(rf * rf)
        r.sur_R = float(sur_R)
        r.sur_specular = bool(sur_specular)
    return r


class TimeBinning(C.Structure):
    """src/transport.jl:59-62: nbins + 1 bins, the last is the overflow bin."""
    _fields_ = [("bin_ns", C.c_float), ("nbins", C.c_int32)]

    def __init__(self, bin_ns, nbins):
        super().__init__(float(bin_ns), int(nbins))


class Emission(C.Structure):
    """Emission-time model: decay (the reference's tau_ns) plus an optional rise time,
    an extension with no reference counterpart (src/generation.jl:12-13 is decay-only)."""
    _fields_ = [("tau_decay", C.c_float), ("tau_rise", C.c_float)]

    def __init__(self, tau_decay=0.0, tau_rise=0.0):
        super().__init__(float(tau_decay), float(tau_rise))


# byte sizes the Julia isbits layout fixes (checked by the tests)
LAYOUT_BYTES = {"Box": 12, "OpticalParams": 32, "SipmGrid": 12, "Readout": 36,
                "TimeBinning": 8, "Emission": 8}


@dataclass
class Accumulator:
    """src/transport.jl:64-80.  counts is (nx, ny, nbins+1) in Fortran order so that
    counts[i-1, j-1, b-1] is the reference's counts[i, j, b]."""
    counts: np.ndarray
    first_ns: np.ndarray
    ndet: int = 0
    ndet_front: int = 0
    nabs_bulk: int = 0
    nabs_wall: int = 0
    nabs_sipm: int = 0
    ncap: int = 0
    nabs_sur: int = 0
    nscat: int = 0
    sum_bounces_det: int = 0

    @classmethod
    def empty(cls, grid: SipmGrid, tb: TimeBinning) -> "Accumulator":
        return cls(np.zeros((grid.nx, grid.ny, tb.nbins + 1), np.uint32, order="F"),
                   np.full((grid.nx, grid.ny), np.inf, np.float32, order="F"))

    COUNTER_NAMES = ("ndet", "ndet_front", "nabs_bulk", "nabs_wall", "nabs_sipm", "ncap",
                     "nabs_sur", "nscat", "sum_bounces_det")

    def merge(self, other: "Accumulator") -> "Accumulator":
        """Base.merge! (src/transport.jl:82-91)."""
        self.counts += other.counts
        np.minimum(self.first_ns, other.first_ns, out=self.first_ns)
        for n in self.COUNTER_NAMES:
            setattr(self, n, getattr(self, n) + getattr(other, n))
        return self

    def total_terminated(self) -> int:
        """src/transport.jl:93-94."""
        return (self.ndet + self.nabs_bulk + self.nabs_wall + self.nabs_sipm + self.ncap
                + self.nabs_sur)


def deg2rad(x) -> float:
    return float(np.float32(math.radians(x)))
