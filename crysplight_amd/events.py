"""Event pipeline on the HIP library: gamma deposits -> optical photons -> per-event maps.

run_events         <- run_events!  (src/events.jl:39-125)
transport_events   <- the flush!/per-event reduction of run_events! (src/events.jl:65-91)
GammaMaterial      <- Gamma.make_material / sigma_macro / mfp (src/gamma/gmaterials.jl)
gamma_events       <- the event loop's gamma stage (src/events.jl:93-117)
rand_poisson       <- rand_poisson (src/events.jl:11-22), on Philox streams

All photons of all events go through ONE persistent launch fed by a device work
queue, so the reference's batch_photons knob only bounds host memory here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from .params import Box, Emission, OpticalParams, Readout, SipmGrid, TimeBinning, readout

DATA_DIR = Path(__file__).resolve().parent / "data"

DEP_PHOTO, DEP_COMPTON, DEP_BELOWCUT = 0, 1, 2          # src/gamma_interface.jl:11-13


def _vp(a):
    return None if a is None else C.c_void_p(_lib.ptr(a))


def _bind():
    lib = _lib.load()
    if getattr(lib, "_gamma_bound", False):
        return lib
    vp = C.c_void_p
    lib.crysp_gamma_material_load.restype = vp
    lib.crysp_gamma_material_load.argtypes = [C.c_char_p, C.c_double]
    lib.crysp_gamma_material_create.restype = vp
    lib.crysp_gamma_material_create.argtypes = [vp, C.c_int32, C.c_double]
    lib.crysp_gamma_material_free.argtypes = [vp]
    lib.crysp_gamma_last_error.restype = C.c_char_p
    lib.crysp_gamma_sigma.argtypes = [vp, C.c_double, vp]
    lib.crysp_gamma_history.argtypes = [vp, C.POINTER(Box), C.c_double, vp, vp, C.c_uint64,
                                        C.c_uint64, vp, C.c_int32, vp]
    lib.crysp_gamma_events.argtypes = [vp, C.POINTER(Box), C.c_int64, C.c_uint64, C.c_double,
                                       C.c_double, C.c_int32] + [vp] * 14
    lib.crysp_synthetic_deposits.argtypes = [C.c_int64, C.c_uint64, vp, vp, C.c_float,
                                             C.c_float, C.c_double, C.c_int32, vp, vp, vp]
    lib.crysp_poisson_samples.argtypes = [C.c_uint64, C.c_double, C.c_int64, vp]
    lib._gamma_bound = True
    return lib


class GammaMaterial:
    """Photon cross sections of a crystal from a NIST XCOM table and a density
    (Gamma.make_material, src/gamma/gmaterials.jl:22-31)."""

    def __init__(self, xcom: str | Path, density: float, name: str = ""):
        lib = _bind()
        path = Path(xcom)
        if not path.is_absolute() and not path.exists():
            path = DATA_DIR / path
        h = lib.crysp_gamma_material_load(str(path).encode(), float(density))
        if not h:
            raise _lib.CryspError(lib.crysp_gamma_last_error().decode())
        self._h = h
        self.name = name
        self.density = float(density)

    @property
    def handle(self):
        return self._h

    def sigma_macro(self, e_mev: float):
        """(Sigma_Compton, Sigma_photo, Sigma_pair) [1/cm] (src/gamma/gmaterials.jl:45-55)."""
        out = np.zeros(3)
        _bind().crysp_gamma_sigma(self._h, float(e_mev), out.ctypes.data)
        return tuple(float(v) for v in out)

    def mfp(self, e_mev: float) -> float:
        """Mean free path [cm] (src/gamma/gmaterials.jl:58-61)."""
        s = sum(self.sigma_macro(e_mev))
        return 1.0 / s if s > 0 else float("inf")

    def history(self, box: Box, e0_mev: float, p0_cm, dir0, seed: int, stream: int = 1,
                cap: int = 64) -> np.ndarray:
        """One gamma history (Gamma.propagate_photon, src/gamma/gtransport.jl:44-81):
        rows (x, y, z [cm, box-centred], e_in, e_dep [MeV], process code)."""
        recs = np.zeros((cap, 6))
        n = C.c_int32()
        p = np.asarray(p0_cm, np.float64)
        d = np.asarray(dir0, np.float64)
        rc = _bind().crysp_gamma_history(self._h, C.byref(box), float(e0_mev), p.ctypes.data,
                                         d.ctypes.data, int(seed), int(stream),
                                         recs.ctypes.data, cap, C.byref(n))
        _lib.check(rc, "crysp_gamma_history")
        return recs[:n.value]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _bind().crysp_gamma_material_free(self._h)
                self._h = None
        except Exception:
            pass


@dataclass
class Deposits:
    """A deposit table: deposit d at xyzt[d] = (x, y, z [mm], t0 [ns]) with nph[d]
    photons, belonging to event event[d]."""
    xyzt: np.ndarray        # (ndep, 4) float32
    nph: np.ndarray         # (ndep,) int32
    event: np.ndarray       # (ndep,) int32
    e_kev: np.ndarray       # (ndep,) float32
    n_events: int

    @property
    def n_photons(self) -> int:
        return int(self.nph.astype(np.int64).sum())


def gamma_events(box: Box, mat: GammaMaterial, yield_per_mev: float, *, n_events: int,
                 seed: int, e0_mev: float = 0.511, max_dep_per_event: int = 32):
    """The gamma stage of run_events! for n_events gammas entering z = 0 along +z:
    returns (Deposits, truth dict with edep, xyz1, e1, xyz2, e2, er, int_type, n_int)."""
    n = int(n_events)
    K = int(max_dep_per_event)
    dep = np.zeros((n * K, 4), np.float32)
    nph = np.zeros(n * K, np.int32)
    ev = np.zeros(n * K, np.int32)
    ek = np.zeros(n * K, np.float32)
    proc = np.zeros(n * K, np.int8)
    first = np.zeros(n + 1, np.int64)
    truth = {"edep": np.zeros(n, np.float32), "xyz1": np.zeros((n, 3), np.float32),
             "e1": np.zeros(n, np.float32), "xyz2": np.zeros((n, 3), np.float32),
             "e2": np.zeros(n, np.float32), "er": np.zeros(n, np.float32),
             "int_type": np.zeros(n, np.int8), "n_int": np.zeros(n, np.int16)}
    nd = C.c_int64()
    rc = _bind().crysp_gamma_events(
        mat.handle, C.byref(box), n, int(seed), float(e0_mev), float(yield_per_mev), K,
        *[_vp(a) for a in (dep, nph, ev, ek, proc, first, truth["edep"], truth["xyz1"],
                           truth["e1"], truth["xyz2"], truth["e2"], truth["er"],
                           truth["int_type"], truth["n_int"])], C.byref(nd))
    if rc != 0:
        raise _lib.CryspError("crysp_gamma_events: " + _bind().crysp_gamma_last_error().decode())
    m = nd.value
    d = Deposits(dep[:m].copy(), nph[:m].copy(), ev[:m].copy(), ek[:m].copy(), n)
    truth["dep_proc"] = proc[:m].copy()
    truth["ev_first"] = first
    return d, truth


def synthetic_deposits(box: Box, *, n_events: int, seed: int, e_kev=(10.0, 511.0),
                       yield_per_mev: float = 0.0, n_fixed: int = 0, lo=None, hi=None) -> Deposits:
    """Single-deposit events, position uniform in [lo, hi] (default the whole crystal),
    energy uniform in e_kev, photons Poisson(yield * E / 1000) or n_fixed each."""
    n = int(n_events)
    lo = np.asarray(lo if lo is not None else (0.0, 0.0, 0.0), np.float32)
    hi = np.asarray(hi if hi is not None else tuple(box.L), np.float32)
    dep = np.zeros((n, 4), np.float32)
    ek = np.zeros(n, np.float32)
    nph = np.zeros(n, np.int32)
    rc = _bind().crysp_synthetic_deposits(n, int(seed), lo.ctypes.data, hi.ctypes.data,
                                          float(e_kev[0]), float(e_kev[1]), float(yield_per_mev),
                                          int(n_fixed), dep.ctypes.data, ek.ctypes.data,
                                          nph.ctypes.data)
    _lib.check(rc, "crysp_synthetic_deposits")
    return Deposits(dep, nph, np.arange(n, dtype=np.int32), ek, n)


def rand_poisson(lam: float, n: int, seed: int = 1) -> np.ndarray:
    """n draws of rand_poisson(lam) (src/events.jl:11-22) from independent streams."""
    out = np.zeros(int(n), np.int64)
    _lib.check(_bind().crysp_poisson_samples(int(seed), float(lam), int(n), out.ctypes.data),
               "crysp_poisson_samples")
    return out


def _check_event_arrays(ctx, ndep: int, nev: int, nxy: int, inputs, outputs) -> None:
    """Types and sizes of an event call's arrays (numpy or torch, host or device), then
    wait for torch's queued work on the device ones (the library runs on its own stream)."""
    xyzt, nph, ev = inputs[:3]
    _lib.check_array(xyzt, "xyzt", np.float32, (ndep, 4), device=ctx.device)
    _lib.check_array(nph, "nph", np.int32, (ndep,), device=ctx.device)
    _lib.check_array(ev, "event", np.int32, (ndep,), device=ctx.device)
    if len(inputs) > 3:
        _lib.check_array(inputs[3], "dep_box", np.int32, (ndep,), device=ctx.device)
    maps, tmin, npe = outputs
    _lib.check_array(maps, "maps", np.int32, numel=nev * nxy, device=ctx.device)
    _lib.check_array(tmin, "tmin", np.float32, numel=nev * nxy, device=ctx.device)
    _lib.check_array(npe, "npe", np.int32, numel=nev, device=ctx.device)
    _lib.ready(*inputs, *outputs)


@dataclass
class EventMaps:
    maps: np.ndarray       # (n_events, ny, nx) int32 view of the (nx, ny, n) column-major array
    tmin: np.ndarray       # (n_events, ny, nx) float32, +inf where no hit
    npe: np.ndarray        # (n_events,) int32
    n_photons: int


def transport_events(box: Box, op: OpticalParams, ro, dep: Deposits, *, seed: int,
                     tau_ns: float, tau_rise_ns: float = 0.0, pid0: int = 0,
                     max_bounces: int = 100_000, ctx=None, device_arrays=None,
                     out=None) -> EventMaps:
    """All photons of the deposit table through one persistent launch: per-event maps.
    device_arrays: optional (xyzt, nph, event) torch tensors already on the GPU;
    out: optional (maps, tmin, npe) torch tensors on the GPU (then returned as-is)."""
    ctx = ctx or _lib.default_context()
    ro = readout(ro) if isinstance(ro, SipmGrid) else ro
    nxy = ro.grid.nx * ro.grid.ny
    nev = int(dep.n_events)
    if device_arrays is not None:
        xyzt, nph, ev = device_arrays
    else:
        xyzt = np.ascontiguousarray(dep.xyzt, np.float32)
        nph = np.ascontiguousarray(dep.nph, np.int32)
        ev = np.ascontiguousarray(dep.event, np.int32)
    if out is not None:
        maps, tmin, npe = out
    else:
        maps = np.zeros((nev, ro.grid.ny, ro.grid.nx), np.int32)
        tmin = np.zeros((nev, ro.grid.ny, ro.grid.nx), np.float32)
        npe = np.zeros(nev, np.int32)
    em = Emission(tau_ns, tau_rise_ns)
    ntot = C.c_int64()
    ndep = int(dep.xyzt.shape[0]) if device_arrays is None else int(xyzt.shape[0])
    _check_event_arrays(ctx, ndep, nev, nxy, (xyzt, nph, ev), (maps, tmin, npe))
    rc = _lib.load().crysp_run_events(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), _lib.ptr(xyzt), _lib.ptr(nph),
        _lib.ptr(ev), ndep, nev, int(seed), int(pid0), int(max_bounces), C.byref(em),
        _lib.ptr(maps), _lib.ptr(tmin), _lib.ptr(npe), C.byref(ntot))
    _lib.check(rc, "crysp_run_events")
    return EventMaps(maps, tmin, npe, ntot.value)


def run_events(box: Box, op: OpticalParams, ro, tb: TimeBinning, mat: GammaMaterial,
               yield_per_mev: float, *, n_events: int, seed: int, tau_ns: float,
               tau_rise_ns: float = 0.0, e0_mev: float = 0.511, max_bounces: int = 100_000,
               batch_photons: int | None = None, ctx=None) -> dict:
    """run_events! (src/events.jl:39-125): gamma deposits, Poisson photon yields, one
    batched transport, per-event outputs and truth.  Returns a dict with the reference's
    fields; maps are Int32 (n_events, ny, nx) (the reference's UInt16 (nx, ny, n) read
    in row-major order), tmin likewise.  batch_photons is accepted for interface parity:
    the device work queue takes every event in one launch."""
    dep, truth = gamma_events(box, mat, yield_per_mev, n_events=n_events, seed=seed,
                              e0_mev=e0_mev)
    res = transport_events(box, op, ro, dep, seed=seed, tau_ns=tau_ns,
                           tau_rise_ns=tau_rise_ns, max_bounces=max_bounces, ctx=ctx)
    out = {"edep": truth["edep"], "npe": res.npe, "maps": res.maps, "tmin": res.tmin,
           "xyz1": truth["xyz1"], "e1": truth["e1"], "xyz2": truth["xyz2"],
           "e2": truth["e2"], "er": truth["er"], "int_type": truth["int_type"],
           "n_int": truth["n_int"], "n_photons": res.n_photons, "deposits": dep}
    return out


def autotune(box: Box, op: OpticalParams, ro, dep: Deposits, *, seed: int, tau_ns: float,
             tau_rise_ns: float = 0.0, max_bounces: int = 100_000, ctx=None,
             wg_sizes=(64, 128, 256, 512, 1024), ppts=(1, 2, 3, 4, 5, 6, 7, 8), reps: int = 2):
    """Sweep workgroup size x photons per thread on this deposit table (event mode),
    keep the fastest schedule in the context.  Returns ((wg, ppt), times_ms[n_wg, n_ppt])."""
    ctx = ctx or _lib.default_context()
    ro = readout(ro) if isinstance(ro, SipmGrid) else ro
    wg = np.ascontiguousarray(wg_sizes, np.int32)
    pp = np.ascontiguousarray(ppts, np.int32)
    times = np.zeros((wg.size, pp.size), np.float32)
    bw, bp = C.c_int32(), C.c_int32()
    em = Emission(tau_ns, tau_rise_ns)
    xyzt = np.ascontiguousarray(dep.xyzt, np.float32)
    nph = np.ascontiguousarray(dep.nph, np.int32)
    ev = np.ascontiguousarray(dep.event, np.int32)
    rc = _lib.load().crysp_autotune_events(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), xyzt.ctypes.data, nph.ctypes.data,
        ev.ctypes.data, xyzt.shape[0], int(dep.n_events), int(seed), int(max_bounces),
        C.byref(em), wg.ctypes.data, wg.size, pp.ctypes.data, pp.size, int(reps),
        times.ctypes.data, C.byref(bw), C.byref(bp))
    _lib.check(rc, "crysp_autotune_events")
    return (bw.value, bp.value), times


def autotune_queue(box: Box, op: OpticalParams, ro, dep: Deposits, *, seed: int, tau_ns: float,
                   tau_rise_ns: float = 0.0, max_bounces: int = 100_000, ctx=None,
                   wg_sizes=(128, 256, 512), ppts=(4,), chunks=(64, 128, 256, 512),
                   refill_mins=(2, 4, 8, 12), reps: int = 2):
    """Sweep every launch knob -- workgroup size, photons per thread, the work queue's grab
    size (chunk) and refill batch (refill_min) -- on this deposit table (event mode) and keep
    the fastest in the context.  Returns (best (wg, ppt, chunk, refill_min),
    times_ms[n_wg, n_ppt, n_chunk, n_refill])."""
    ctx = ctx or _lib.default_context()
    ro = readout(ro) if isinstance(ro, SipmGrid) else ro
    lists = [np.ascontiguousarray(v, np.int32) for v in (wg_sizes, ppts, chunks, refill_mins)]
    times = np.zeros(tuple(v.size for v in lists), np.float32)
    best = np.zeros(4, np.int32)
    em = Emission(tau_ns, tau_rise_ns)
    xyzt = np.ascontiguousarray(dep.xyzt, np.float32)
    nph = np.ascontiguousarray(dep.nph, np.int32)
    ev = np.ascontiguousarray(dep.event, np.int32)
    rc = _lib.load().crysp_autotune_events_queue(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), xyzt.ctypes.data, nph.ctypes.data,
        ev.ctypes.data, xyzt.shape[0], dep.n_events, int(seed), int(max_bounces), C.byref(em),
        *[a for v in lists for a in (v.ctypes.data, v.size)], int(reps), times.ctypes.data,
        best.ctypes.data)
    _lib.check(rc, "crysp_autotune_events_queue")
    return tuple(int(v) for v in best), times


def transport_events_multibox(boxes, dep_box, op: OpticalParams, ro, dep: Deposits, *, seed: int,
                              tau_ns: float, tau_rise_ns: float = 0.0, pid0: int = 0,
                              max_bounces: int = 100_000, ctx=None, device_arrays=None,
                              out=None) -> EventMaps:
    """Event mode with deposits in several crystals in one queue: deposit d lies in
    boxes[dep_box[d]].  device_arrays: optional (xyzt, nph, event, dep_box) torch tensors
    already on the GPU (dep_box is then ignored); out: optional (maps, tmin, npe) torch
    tensors on the GPU (then returned as-is)."""
    ctx = ctx or _lib.default_context()
    ro = readout(ro) if isinstance(ro, SipmGrid) else ro
    nev = int(dep.n_events)
    arr = (Box * len(boxes))(*[Box(tuple(b.L)) for b in boxes])
    if device_arrays is not None:
        xyzt, nph, ev, db = device_arrays
        ndep = int(xyzt.shape[0])
    else:
        db = np.ascontiguousarray(dep_box, np.int32)
        xyzt = np.ascontiguousarray(dep.xyzt, np.float32)
        nph = np.ascontiguousarray(dep.nph, np.int32)
        ev = np.ascontiguousarray(dep.event, np.int32)
        ndep = int(xyzt.shape[0])
    if out is not None:
        maps, tmin, npe = out
    else:
        maps = np.zeros((nev, ro.grid.ny, ro.grid.nx), np.int32)
        tmin = np.zeros((nev, ro.grid.ny, ro.grid.nx), np.float32)
        npe = np.zeros(nev, np.int32)
    em = Emission(tau_ns, tau_rise_ns)
    ntot = C.c_int64()
    _check_event_arrays(ctx, ndep, nev, ro.grid.nx * ro.grid.ny, (xyzt, nph, ev, db),
                        (maps, tmin, npe))
    lib = _lib.load()
    f = lib.crysp_run_events_multibox
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(OpticalParams),
                  C.POINTER(Readout), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                  C.c_uint64, C.c_uint64, C.c_int32, C.POINTER(Emission), C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p]
    rc = f(ctx.handle, C.addressof(arr), len(boxes), _lib.ptr(db), C.byref(op), C.byref(ro),
           _lib.ptr(xyzt), _lib.ptr(nph), _lib.ptr(ev), ndep, nev, int(seed), int(pid0),
           int(max_bounces), C.byref(em), _lib.ptr(maps), _lib.ptr(tmin), _lib.ptr(npe),
           C.byref(ntot))
    _lib.check(rc, "crysp_run_events_multibox")
    return EventMaps(maps, tmin, npe, ntot.value)


def concat_deposits(parts) -> tuple[Deposits, np.ndarray, np.ndarray]:
    """Concatenate deposit tables (events renumbered in order).  Returns the table, the
    part index of every deposit, and the event offsets of every part."""
    xyzt, nph, ev, ek, part = [], [], [], [], []
    off = 0
    ev_off = []
    for i, d in enumerate(parts):
        xyzt.append(d.xyzt); nph.append(d.nph); ev.append(d.event + off); ek.append(d.e_kev)
        part.append(np.full(d.nph.shape[0], i, np.int32))
        ev_off.append(off)
        off += d.n_events
    return (Deposits(np.concatenate(xyzt), np.concatenate(nph), np.concatenate(ev).astype(np.int32),
                     np.concatenate(ek), off), np.concatenate(part), np.asarray(ev_off))
