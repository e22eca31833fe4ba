"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the crysplight_amd package.  Parameter arguments are any
objects with the reference's field names (the crysplight_amd ctypes structs work).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None
_xlib = None


class Box(C.Structure):
    _fields_ = [("L", C.c_float * 3)]


class Optical(C.Structure):
    _fields_ = [("n_crystal", C.c_float), ("n_coupling", C.c_float), ("abs_len_mm", C.c_float),
                ("rayleigh_mm", C.c_float), ("wrap_R", C.c_float), ("wrap_specular", C.c_uint8),
                ("air_gap", C.c_uint8), ("pad0", C.c_uint8), ("pad1", C.c_uint8),
                ("sigma_alpha", C.c_float), ("pde", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("pitch_mm", C.c_float), ("nx", C.c_int32), ("ny", C.c_int32)]


class ReadoutS(C.Structure):
    _fields_ = [("grid", Grid), ("disc", C.c_uint8), ("pad0", C.c_uint8 * 3), ("cx", C.c_float),
                ("cy", C.c_float), ("r2", C.c_float), ("sur_R", C.c_float),
                ("sur_specular", C.c_uint8), ("two_sided", C.c_uint8), ("pad1", C.c_uint8 * 2)]


class TBin(C.Structure):
    _fields_ = [("bin_ns", C.c_float), ("nbins", C.c_int32)]


class OpCount(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_photon", "n_draw", "n_block", "n_step", "n_fresnel",
                                         "n_surface_try", "n_lambert", "n_rayleigh_try",
                                         "n_detect", "n_nscat")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


OPCOUNT_NAMES = [n for n, _ in OpCount._fields_]


def lib():
    global _lib
    if _lib is None:
        so = HERE / "liboracle.so"
        if not so.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), str(so)], check=True)
        _lib = C.CDLL(str(so))
        _lib.or_philox4x32_10.argtypes = [C.c_void_p] * 3
        _lib.or_layout.argtypes = [C.c_void_p]
        _lib.or_wall_hit.argtypes = [C.c_void_p] * 4
        _lib.or_max_threads.restype = C.c_int32
    return _lib


def xlib():
    """rocRAND cross-check library (host build of rocRAND's header engine)."""
    global _xlib
    if _xlib is None:
        try:
            import torch  # noqa: F401  (share torch's HIP runtime, see crysplight_amd._lib)
        except Exception:
            pass
        so = HERE / "librocrand_xcheck.so"
        if not so.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), str(so)], check=True)
        _xlib = C.CDLL(str(so))
    return _xlib


def _box(b):
    return Box((C.c_float * 3)(*[float(v) for v in b.L]))


def _op(o):
    return Optical(o.n_crystal, o.n_coupling, o.abs_len_mm, o.rayleigh_mm, o.wrap_R,
                   int(bool(o.wrap_specular)), int(bool(o.air_gap)), 0, 0, o.sigma_alpha, o.pde)


def _grid(g):
    return Grid(g.pitch_mm, g.nx, g.ny)


def _ro(r):
    if not hasattr(r, "grid"):          # a bare grid: whole face active
        return ReadoutS(_grid(r), 0, (C.c_uint8 * 3)(), 0, 0, 0, 0, 0, 0)
    return ReadoutS(_grid(r.grid), int(bool(r.disc)), (C.c_uint8 * 3)(), r.cx, r.cy, r.r2,
                    r.sur_R, int(bool(r.sur_specular)), int(bool(r.two_sided)))


def _tb(t):
    return TBin(t.bin_ns, t.nbins)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def rocrand_block(seed: int, pid: int, block: int) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    f = xlib().xcheck_philox_block
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]
    f(seed, pid, block, _p(out))
    return out


def debug_draws(seed: int, n_pids: int, n_draws: int, pid0: int = 0) -> np.ndarray:
    out = np.zeros((n_pids, n_draws), np.uint32)
    f = lib().or_debug_draws
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p]
    f(seed, pid0, n_pids, n_draws, _p(out))
    return out


def u32_to_unit(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint32)
    u = np.zeros(x.shape, np.float32)
    f = lib().or_u32_to_unit
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    f(_p(x), x.size, _p(u))
    return u


def wall_hit(p, box):
    pa = np.asarray(p, np.float32)
    L = np.asarray(box.L, np.float32)
    d = C.c_float()
    face = C.c_int32()
    lib().or_wall_hit(_p(pa), _p(L), C.byref(d), C.byref(face))
    return d.value, face.value


def fresnel(n1, n2, cosi):
    R = C.c_float()
    f = lib().or_fresnel
    f.argtypes = [C.c_float, C.c_float, C.c_float, C.c_void_p]
    tir = f(n1, n2, cosi, C.byref(R))
    return bool(tir), R.value


def sample(kind: str, seed: int, pid: int, n: int, face: int = 0, args=()) -> np.ndarray:
    """n successive draws of a sampler from one stream: isotropic, lambertian, rayleigh,
    surface, emission.  Returns (n, 4) float32."""
    code = {"isotropic": 0, "lambertian": 1, "rayleigh": 2, "surface": 3, "emission": 4}[kind]
    a = np.zeros(8, np.float32)
    a[:len(args)] = args
    out = np.zeros((n, 4), np.float32)
    f = lib().or_sample
    f.argtypes = [C.c_int32, C.c_uint64, C.c_uint64, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
    f(code, seed, pid, n, face, _p(a), _p(out))
    return out


def propagate_one(box, op, ro, seed, pid, start, max_bounces=100_000):
    """propagate_photon! from a given (x, y, z, ux, uy, uz, t0) on stream (seed, pid)."""
    s = np.asarray(start, np.float32)
    st, ix, t, b = C.c_uint8(), C.c_int16(), C.c_float(), C.c_int32()
    f = lib().or_propagate_one
    f.argtypes = [C.c_void_p] * 3 + [C.c_uint64, C.c_uint64, C.c_void_p, C.c_int32] + [C.c_void_p] * 4
    bx, o, r = _box(box), _op(op), _ro(ro)
    f(C.byref(bx), C.byref(o), C.byref(r), seed, pid, _p(s), max_bounces, C.byref(st),
      C.byref(ix), C.byref(t), C.byref(b))
    return st.value, ix.value, t.value, b.value


def transport_records(box, op, ro, *, n_photons, seed, pos=None, positions=None, t0s=None,
                      pid_offset=0, tau_ns=0.0, tau_rise_ns=0.0, max_bounces=100_000,
                      threads=0, count_ops=False):
    """Records of src/kernel.jl:159-305 for n photons.  Returns a dict of arrays
    (+ 'ops' when count_ops)."""
    if positions is None:
        positions = [pos]
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    px, py, pz = (np.ascontiguousarray(p[:, a]) for a in range(3))
    pt = None if t0s is None else np.ascontiguousarray(np.asarray(t0s, np.float32))
    n = int(n_photons)
    st = np.zeros(n, np.uint8)
    ix = np.zeros(n, np.int16)
    t = np.zeros(n, np.float32)
    b = np.zeros(n, np.int32)
    nscat = C.c_int64()
    oc = OpCount()
    f = lib().or_transport_records
    f.argtypes = [C.c_void_p] * 3 + [C.c_void_p] * 4 + [C.c_int32, C.c_int64, C.c_uint64,
                                                        C.c_uint64, C.c_int32, C.c_float,
                                                        C.c_float] + [C.c_void_p] * 6 + [C.c_int32]
    bx, o, r = _box(box), _op(op), _ro(ro)
    f(C.byref(bx), C.byref(o), C.byref(r), _p(px), _p(py), _p(pz), _p(pt), len(px), n, seed,
      pid_offset, max_bounces, tau_ns, tau_rise_ns, _p(st), _p(ix), _p(t), _p(b),
      C.byref(nscat), C.byref(oc) if count_ops else None, threads)
    out = {"status": st, "idx": ix, "t": t, "bounces": b, "nscat": nscat.value}
    if count_ops:
        out["ops"] = oc.as_dict()
    return out


COUNTER_NAMES = ("ndet", "ndet_front", "nabs_bulk", "nabs_wall", "nabs_sipm", "ncap",
                 "nabs_sur", "nscat", "sum_bounces_det")


def reduce_records(rec, grid, tb):
    """src/kernel.jl:338-365: counts (nx, ny, nbins+1) F-order, first_ns, counters."""
    n = rec["status"].shape[0]
    counts = np.zeros((grid.nx, grid.ny, tb.nbins + 1), np.uint32, order="F")
    first = np.zeros((grid.nx, grid.ny), np.float32, order="F")
    ctr = np.zeros(9, np.int64)
    f = lib().or_reduce_records
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + [C.c_void_p] * 7
    g, t = _grid(grid), _tb(tb)
    f(C.byref(g), C.byref(t), n, _p(rec["status"]), _p(rec["idx"]), _p(rec["t"]),
      _p(rec["bounces"]), _p(counts), _p(first), _p(ctr))
    res = dict(zip(COUNTER_NAMES, (int(v) for v in ctr)))
    res["nscat"] = int(rec.get("nscat", 0))
    res["counts"] = counts
    res["first_ns"] = first
    return res


def run_photons(box, op, ro, tb, **kw):
    """run_photons! equivalent: records then the reference host reduction."""
    rec = transport_records(box, op, ro, **kw)
    grid = ro.grid if hasattr(ro, "grid") else ro
    return reduce_records(rec, grid, tb)


def run_events(box, op, ro, dep_xyzt, dep_nph, dep_event, nev, *, seed, pid0=0,
               max_bounces=100_000, tau_ns=0.0, tau_rise_ns=0.0, threads=0,
               return_records=False, count_ops=False):
    dep = np.ascontiguousarray(np.asarray(dep_xyzt, np.float32).reshape(-1, 4))
    nph = np.ascontiguousarray(np.asarray(dep_nph, np.int32))
    dev = np.ascontiguousarray(np.asarray(dep_event, np.int32))
    grid = ro.grid if hasattr(ro, "grid") else ro
    nxy = grid.nx * grid.ny
    maps = np.zeros((nev, nxy), np.int32)
    tmin = np.zeros((nev, nxy), np.float32)
    npe = np.zeros(nev, np.int32)
    ntot = int(nph.astype(np.int64).sum())
    rs = ri = rt = None
    if return_records:
        rs, ri, rt = np.zeros(ntot, np.uint8), np.zeros(ntot, np.int16), np.zeros(ntot, np.float32)
    f = lib().or_run_events
    f.argtypes = [C.c_void_p] * 6 + [C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, C.c_int32,
                                     C.c_float, C.c_float] + [C.c_void_p] * 7 + [C.c_int32]
    bx, o, r = _box(box), _op(op), _ro(ro)
    oc = np.zeros(len(OPCOUNT_NAMES), np.int64) if count_ops else None
    f(C.byref(bx), C.byref(o), C.byref(r), _p(dep), _p(nph), _p(dev), dep.shape[0], nev, seed,
      pid0, max_bounces, tau_ns, tau_rise_ns, _p(maps), _p(tmin), _p(npe), _p(rs), _p(ri),
      _p(rt), _p(oc), threads)
    out = {"maps": maps, "tmin": tmin, "npe": npe}
    if return_records:
        out.update(status=rs, idx=ri, t=rt)
    if count_ops:
        out["ops"] = dict(zip(OPCOUNT_NAMES, (int(v) for v in oc)))
    return out


def math(which: str, x: np.ndarray) -> np.ndarray:
    """The oracle's Float32 log / sin / cos ((float)f((double)x), glibc)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    f = lib().or_math
    f.argtypes = [C.c_int32, C.c_void_p, C.c_int64, C.c_void_p]
    f({"log": 0, "sin": 1, "cos": 2}[which], _p(x), x.size, _p(y))
    return y


def correctly_rounded(which: str, x: np.ndarray) -> np.ndarray:
    """float32 rounding of the function evaluated in x87 extended precision (64-bit
    significand): the correctly rounded binary32 value except within ~2^-63 relative
    of a rounding boundary."""
    xl = np.asarray(x, np.float32).astype(np.longdouble)
    return getattr(np, which)(xl).astype(np.float32)


def uniform_domain() -> np.ndarray:
    """Every value randu can return: k * 2^-24, k = 1..2^24 (src/philox.jl:62)."""
    return (np.arange(1, 2**24 + 1, dtype=np.float64) * 2.0**-24).astype(np.float32)


def phase_domain() -> np.ndarray:
    """Every azimuth the samplers form: 2f0*Float32(pi) * u, u in uniform_domain()."""
    return np.float32(6.2831855) * uniform_domain()


def tilt_domain(step: int = 7) -> np.ndarray:
    """Every step-th binary32 value in [0, 1.2): the facet tilts alpha < 1.2."""
    hi = np.float32(1.2).view(np.uint32)
    return np.arange(0, hi, step, dtype=np.uint32).view(np.float32)


def layout_sizes():
    s = np.zeros(6, np.int32)
    lib().or_layout(_p(s))
    return dict(zip(["Box", "OpticalParams", "SipmGrid", "Readout", "TimeBinning", "OpCount"],
                    (int(v) for v in s)))


def max_threads() -> int:
    return int(lib().or_max_threads())
