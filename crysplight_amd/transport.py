"""Host mirror of the reference's transport operators, running on the HIP library.

run_photons_ka  <- run_photons_ka! (src/kernel.jl:312-367)
run_photons     <- run_photons!    (src/transport.jl:236-273)
debug_draws     <- the first K words of PhiloxStream(seed, pid) (src/philox.jl:43-57)

Arrays may be numpy (host) or torch tensors on the GPU; results are numpy unless
`out=` device tensors are supplied by the caller.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .params import (Accumulator, Box, Emission, OpticalParams, Readout, SipmGrid,
                     TimeBinning, readout, STATUS_DETECTED, STATUS_ABS_BULK,
                     STATUS_ABS_WALL, STATUS_ABS_SIPM, STATUS_ABS_SUR)


@dataclass
class Records:
    """Per-photon records of transport_kernel! (src/kernel.jl:300-303)."""
    status: np.ndarray   # uint8
    idx: np.ndarray      # int16, SiPM linear index, negative on the z = 0 face
    t: np.ndarray        # float32, time at the end of the photon [ns]
    bounces: np.ndarray  # int32


def _positions(pos, positions, t0s):
    if positions is None:
        if pos is None:
            raise ValueError("pass pos or positions")
        positions = [pos]
    p = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    px, py, pz = (np.ascontiguousarray(p[:, a]) for a in range(3))
    pt = None
    if t0s is not None:
        pt = np.ascontiguousarray(np.asarray(t0s, dtype=np.float32))
        if pt.shape[0] != p.shape[0]:
            raise ValueError("t0s must have one entry per position")
    return px, py, pz, pt


def _as_readout(ro) -> Readout:
    return readout(ro) if isinstance(ro, SipmGrid) else ro


def transport_records(box: Box, op: OpticalParams, ro: Readout, *, n_photons: int,
                      seed: int, pos=None, positions=None, t0s=None, pid_offset: int = 0,
                      tau_ns: float = 0.0, tau_rise_ns: float = 0.0,
                      max_bounces: int = 100_000, ctx=None) -> Records:
    """Records mode: the bit-for-bit replacement of transport_kernel!."""
    ctx = ctx or _lib.default_context()
    ro = _as_readout(ro)
    px, py, pz, pt = _positions(pos, positions, t0s)
    n = int(n_photons)
    rec = Records(np.zeros(n, np.uint8), np.zeros(n, np.int16), np.zeros(n, np.float32),
                  np.zeros(n, np.int32))
    em = Emission(tau_ns, tau_rise_ns)
    rc = _lib.load().crysp_transport_records(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), px.ctypes.data, py.ctypes.data,
        pz.ctypes.data, None if pt is None else pt.ctypes.data, len(px), n, int(seed),
        int(pid_offset), int(max_bounces), C.byref(em), rec.status.ctypes.data,
        rec.idx.ctypes.data, rec.t.ctypes.data, rec.bounces.ctypes.data)
    _lib.check(rc, "crysp_transport_records")
    return rec


def reduce_records(rec: Records, grid: SipmGrid, tb: TimeBinning) -> Accumulator:
    """Host reduction of records into an Accumulator (src/kernel.jl:338-365)."""
    acc = Accumulator.empty(grid, tb)
    s = rec.status
    det = s == STATUS_DETECTED
    lin = rec.idx[det].astype(np.int64)
    acc.ndet_front = int(np.count_nonzero(lin < 0))
    lin = np.abs(lin)
    ii = (lin - 1) % grid.nx
    jj = (lin - 1) // grid.nx
    t = rec.t[det]
    q = (t / np.float32(tb.bin_ns)).astype(np.float32)
    b = np.where(q >= tb.nbins, tb.nbins, np.trunc(np.maximum(q, 0)).astype(np.int64))
    np.add.at(acc.counts, (ii, jj, b), 1)
    flat = acc.first_ns.reshape(-1, order="F")
    np.minimum.at(flat, ii + grid.nx * jj, t)
    acc.first_ns = flat.reshape(grid.nx, grid.ny, order="F")
    acc.ndet = int(det.sum())
    acc.sum_bounces_det = int(rec.bounces[det].astype(np.int64).sum())
    acc.nabs_bulk = int(np.count_nonzero(s == STATUS_ABS_BULK))
    acc.nabs_wall = int(np.count_nonzero(s == STATUS_ABS_WALL))
    acc.nabs_sipm = int(np.count_nonzero(s == STATUS_ABS_SIPM))
    acc.nabs_sur = int(np.count_nonzero(s == STATUS_ABS_SUR))
    acc.ncap = int(s.shape[0]) - acc.ndet - acc.nabs_bulk - acc.nabs_wall - acc.nabs_sipm \
        - acc.nabs_sur
    return acc


def transport_accumulate(box: Box, op: OpticalParams, ro: Readout, tb: TimeBinning, *,
                         n_photons: int, seed: int, pos=None, positions=None, t0s=None,
                         pid_offset: int = 0, tau_ns: float = 0.0, tau_rise_ns: float = 0.0,
                         max_bounces: int = 100_000, ctx=None) -> Accumulator:
    """Accumulate mode: the photons reduced on the device into an Accumulator."""
    ctx = ctx or _lib.default_context()
    ro = _as_readout(ro)
    px, py, pz, pt = _positions(pos, positions, t0s)
    acc = Accumulator.empty(ro.grid, tb)
    counters = np.zeros(9, np.int64)
    em = Emission(tau_ns, tau_rise_ns)
    rc = _lib.load().crysp_transport_accumulate(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), C.byref(tb), px.ctypes.data,
        py.ctypes.data, pz.ctypes.data, None if pt is None else pt.ctypes.data, len(px),
        int(n_photons), int(seed), int(pid_offset), int(max_bounces), C.byref(em),
        acc.counts.ctypes.data, acc.first_ns.ctypes.data, counters.ctypes.data)
    _lib.check(rc, "crysp_transport_accumulate")
    for name, v in zip(Accumulator.COUNTER_NAMES, counters):
        setattr(acc, name, int(v))
    return acc


def autotune_accumulate(box: Box, op: OpticalParams, ro, tb: TimeBinning, *, n_photons: int,
                        seed: int, pos=None, positions=None, t0s=None, tau_ns: float = 0.0,
                        tau_rise_ns: float = 0.0, max_bounces: int = 100_000, ctx=None,
                        wg_sizes=(128, 256, 512), ppts=(1, 2, 4, 8), chunks=(32, 64, 128, 256, 512),
                        refill_mins=(2, 4, 8, 12), reps: int = 2):
    """Sweep every launch knob (workgroup size, photons per thread, queue chunk, refill
    batch) on this point-source run (accumulate mode) and keep the fastest in the context.
    Returns (best (wg, ppt, chunk, refill_min), times_ms[n_wg, n_ppt, n_chunk, n_refill])."""
    ctx = ctx or _lib.default_context()
    ro = _as_readout(ro)
    px, py, pz, pt = _positions(pos, positions, t0s)
    lists = [np.ascontiguousarray(v, np.int32) for v in (wg_sizes, ppts, chunks, refill_mins)]
    times = np.zeros(tuple(v.size for v in lists), np.float32)
    best = np.zeros(4, np.int32)
    em = Emission(tau_ns, tau_rise_ns)
    rc = _lib.load().crysp_autotune_accumulate(
        ctx.handle, C.byref(box), C.byref(op), C.byref(ro), C.byref(tb), px.ctypes.data,
        py.ctypes.data, pz.ctypes.data, None if pt is None else pt.ctypes.data, len(px),
        int(n_photons), int(seed), int(max_bounces), C.byref(em),
        *[a for v in lists for a in (v.ctypes.data, v.size)], int(reps), times.ctypes.data,
        best.ctypes.data)
    _lib.check(rc, "crysp_autotune_accumulate")
    return tuple(int(v) for v in best), times


def run_photons_ka(box: Box, op: OpticalParams, ro, tb: TimeBinning, *, n_photons: int,
                   seed: int, pos=None, positions=None, t0s=None, pid_offset: int = 0,
                   tau_ns: float, tau_rise_ns: float = 0.0, max_bounces: int = 100_000,
                   return_records: bool = False, ctx=None):
    """run_photons_ka! (src/kernel.jl:312-367) on the MI355X.  Returns the Accumulator,
    or (acc, records) with return_records=True (then reduced on the host from the
    records exactly as the reference does; otherwise reduced on the device)."""
    ro = _as_readout(ro)
    kw = dict(n_photons=n_photons, seed=seed, pos=pos, positions=positions, t0s=t0s,
              pid_offset=pid_offset, tau_ns=tau_ns, tau_rise_ns=tau_rise_ns,
              max_bounces=max_bounces, ctx=ctx)
    if return_records:
        rec = transport_records(box, op, ro, **kw)
        return reduce_records(rec, ro.grid, tb), rec
    return transport_accumulate(box, op, ro, tb, **kw)


def run_photons(box: Box, op: OpticalParams, ro, tb: TimeBinning, *, n_photons: int,
                seed: int, pos=None, positions=None, tau_ns: float,
                tau_rise_ns: float = 0.0, max_bounces: int = 100_000, ctx=None) -> Accumulator:
    """run_photons! (src/transport.jl:236-273): pid = 1..n, positions round-robin."""
    return transport_accumulate(box, op, _as_readout(ro), tb, n_photons=n_photons, seed=seed,
                                pos=pos, positions=positions, tau_ns=tau_ns,
                                tau_rise_ns=tau_rise_ns, max_bounces=max_bounces, ctx=ctx)


def debug_draws(seed: int, n_pids: int, n_draws: int, pid0: int = 0, ctx=None) -> np.ndarray:
    """First n_draws raw words of the device Philox streams pid0+1 .. pid0+n_pids."""
    ctx = ctx or _lib.default_context()
    out = np.zeros((int(n_pids), int(n_draws)), np.uint32)
    _lib.check(_lib.load().crysp_debug_draws(ctx.handle, int(seed), int(pid0), int(n_pids),
                                             int(n_draws), out.ctypes.data),
               "crysp_debug_draws")
    return out


def debug_math(which: str, x: np.ndarray, ctx=None) -> np.ndarray:
    """Device Float32 log / sin / cos / sqrt as the transport evaluates them, and the
    hardware log2 estimate of the bulk-absorption decision ("log2_est"), and the largest
    gap between the estimated and the exact Fresnel reflectance ("fresnel_gap")."""
    ctx = ctx or _lib.default_context()
    code = {"log": 0, "sin": 1, "cos": 2, "sqrt": 3, "cos1": 4, "sin_tilt": 5, "cos_tilt": 6,
            "log2_est": 7, "fresnel_gap": 8}[which]
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    _lib.check(_lib.load().crysp_debug_math(ctx.handle, code, x.ctypes.data, x.size,
                                            y.ctypes.data), "crysp_debug_math")
    return y
