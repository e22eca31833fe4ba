"""Evaluation of the light-response surrogate against the Monte Carlo.

train_for_box     : the training workload (uniform points x photons each -> MLP)
validate          : per-SiPM count error against FP64-accumulated MC means on held-out points
compare_events    : gamma events in several crystal thicknesses through (a) the plain MC in
                    one batched queue and (b) the surrogate (Poisson maps from the baked
                    network) with MC fallback for deposits near a face; events/s of both and
                    the per-SiPM relative error and energy-sum bias of (b) against (a)
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from . import surrogate as S
from .events import (Deposits, GammaMaterial, concat_deposits, gamma_events,
                     transport_events_multibox)
from .params import Box


def _sync():
    import torch
    torch.cuda.synchronize()


def train_for_box(box, op, ro, *, tau_ns, tau_rise_ns=0.0, n_samples=1_000_000,
                  photons=10_000, epochs=300, batch=8192, seed=11, log_depth_mm=1.0, fold=True,
                  ctx=None):
    ctx = ctx or _lib.default_context()
    sur = S.LightSurrogate(box, op, ro, tau_ns=tau_ns, tau_rise_ns=tau_rise_ns, seed=seed,
                           log_depth_mm=log_depth_mm, fold=fold)
    t0 = time.perf_counter()
    ts = sur.generate(n_samples, photons, seed=seed, ctx=ctx)
    _sync()
    t1 = time.perf_counter()
    losses = sur.fit(ts, epochs=epochs, batch=batch, lr=3e-3, lr_final=5e-5, seed=seed)
    _sync()
    t2 = time.perf_counter()
    info = {"n_samples": n_samples, "photons_per_sample": photons, "generate_s": t1 - t0,
            "fit_s": t2 - t1, "epochs": epochs, "batch": batch, "adam_steps": int(losses.size),
            "symmetry_folded": sur.fold,
            "final_loss": float(losses[-min(50, losses.size):].mean()),
            "train_flops": float(3.0 * sur.net.flops_per_row * (losses.size * batch))}
    info["train_tflops_per_s"] = info["train_flops"] / info["fit_s"] / 1e12
    return sur, info


def validate(sur, *, n_points=1000, photons=2_000_000, seed=77, fallback_mm=1.0, ctx=None):
    """Count-weighted mean per-SiPM error against MC means (exact integer counts) of
    n_points held-out positions, and its breakdown by distance to the readout face."""
    ctx = ctx or _lib.default_context()
    L = np.asarray(sur.box.L, np.float64)
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n_points, 3)) * L
    e = rng.uniform(*sur.e_range, n_points)
    c, tsum = S.mc_light_maps(sur.box, sur.op, sur.ro, xyz, photons, seed=seed,
                              tau_ns=sur.tau_ns, tau_rise_ns=sur.tau_rise_ns, pid0=1 << 44, ctx=ctx)
    p_true = c / float(photons)
    p, tm = sur.predict(xyz, e)
    inner = np.all((xyz >= fallback_mm) & (xyz <= L - fallback_mm), axis=1)
    depth = L[2] - xyz[:, 2]
    out = {"points": n_points, "photons_per_point": photons,
           "count_error_all": S.count_error(p, p_true),
           "count_error_beyond_fallback": S.count_error(p[inner], p_true[inner]),
           "truth_relative_noise": float(np.mean(np.sqrt(c.sum(1)) / np.maximum(c.sum(1), 1)))}
    bins = {}
    for a, b in ((1, 2), (2, 4), (4, 8), (8, 1e9)):
        m = inner & (depth >= a) & (depth < b)
        if m.any():
            bins[f"readout_distance_{a}-{b if b < 1e9 else 'inf'}mm"] = S.count_error(p[m], p_true[m])
    out["by_readout_distance"] = bins
    with np.errstate(invalid="ignore", divide="ignore"):
        tt = tsum / np.maximum(c, 1)
    m = (c > 1000) & inner[:, None]
    out["time_mean_abs_err_ns"] = float(np.abs(tm[m] - tt[m]).mean()) if m.any() else float("nan")
    return out


def gamma_mix(thicknesses, footprint, mat, yield_per_mev, n_events, seed, max_dep=3):
    """n_events gamma events with 1..max_dep deposits, split evenly over the thicknesses."""
    parts, boxes = [], []
    per = [n_events // len(thicknesses) + (i < n_events % len(thicknesses))
           for i in range(len(thicknesses))]
    for i, (t, n) in enumerate(zip(thicknesses, per)):
        box = Box((footprint, footprint, float(t)))
        boxes.append(box)
        got = None
        k = 0
        while got is None or got.n_events < n:
            d, truth = gamma_events(box, mat, yield_per_mev, n_events=int(n * 1.5) + 16,
                                    seed=seed + 1000 * i + k)
            nint = truth["n_int"].astype(np.int64)
            first = truth["ev_first"]
            ok = (nint >= 1) & (nint <= max_dep) & (np.diff(first) == nint)
            keep_ev = np.flatnonzero(ok)[: n if got is None else n - got.n_events]
            remap = -np.ones(d.n_events, np.int64)
            remap[keep_ev] = np.arange(keep_ev.size)
            sel = remap[d.event] >= 0
            part = Deposits(d.xyzt[sel], d.nph[sel], remap[d.event[sel]].astype(np.int32),
                            d.e_kev[sel], int(keep_ev.size))
            got = part if got is None else concat_deposits([got, part])[0]
            k += 1
        parts.append(got)
    dep, part_of, ev_off = concat_deposits(parts)
    return dep, part_of, ev_off, boxes


def compare_events(surrogates, boxes, op, ro, dep, part_of, *, yield_per_mev, tau_ns,
                   tau_rise_ns=0.0, seed=5, fallback_mm=1.0, grid_shape=(64, 64), ctx=None):
    """(a) plain MC of every deposit in one multi-crystal queue; (b) surrogate Poisson maps
    for deposits at least fallback_mm from every face, MC for the rest.  Both paths start
    from deposit tables resident on the device and leave their maps there (the maps are
    copied back after the timed regions), like the transport's photons/s."""
    ctx = ctx or _lib.default_context()
    nev = dep.n_events
    import torch
    dev = torch.device("cuda", 0)
    Ls = np.array([b.L for b in boxes], np.float32)[part_of]
    x = dep.xyzt[:, :3]
    use = np.all((x >= fallback_mm) & (x <= Ls - fallback_mm), axis=1)

    def on_dev(d, part):
        return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                     for a in (d.xyzt, d.nph, d.event, part.astype(np.int32)))

    grid = ro.grid if hasattr(ro, "grid") else ro

    def map_out():
        return (torch.zeros((nev, grid.ny, grid.nx), dtype=torch.int32, device=dev),
                torch.zeros((nev, grid.ny, grid.nx), dtype=torch.float32, device=dev),
                torch.zeros(nev, dtype=torch.int32, device=dev))

    mc_in, mc_out = on_dev(dep, part_of), map_out()
    # (a) plain MC
    _sync()
    t0 = time.perf_counter()
    mc = transport_events_multibox(boxes, part_of, op, ro, dep, seed=seed, tau_ns=tau_ns,
                                   tau_rise_ns=tau_rise_ns, ctx=ctx, device_arrays=mc_in,
                                   out=mc_out)
    _sync()
    t_mc = time.perf_counter() - t0
    # (b) surrogate + fallback
    grids = [s.bake((grid_shape[0], grid_shape[1], max(8, int(round(b.L[2] * 2)) + 1)))
             for s, b in zip(surrogates, boxes)]
    ev_first = np.zeros(nev + 1, np.int64)
    np.add.at(ev_first, dep.event + 1, 1)
    ev_first = np.cumsum(ev_first)
    g_d = [torch.from_numpy(g).to(dev) for g in grids]
    xyzt_d = torch.from_numpy(dep.xyzt).to(dev)
    ek_d = torch.from_numpy(dep.e_kev).to(dev)
    evf_d = torch.from_numpy(ev_first).to(dev)
    uses_d = [torch.from_numpy((use & (part_of == i)).astype(np.uint8)).to(dev)
              for i in range(len(boxes))]
    maps_d = [torch.zeros((nev, 64), dtype=torch.int32, device=dev) for _ in boxes]
    tm_d = [torch.zeros((nev, 64), dtype=torch.float32, device=dev) for _ in boxes]
    fb = Deposits(dep.xyzt[~use], dep.nph[~use], dep.event[~use], dep.e_kev[~use], nev)
    fb_in, fb_out = on_dev(fb, part_of[~use]), map_out()
    _sync()
    t0 = time.perf_counter()
    for i, b in enumerate(boxes):
        S.sample_events(g_d[i], b, Deposits(xyzt_d, None, None, ek_d, nev), evf_d, uses_d[i],
                        yield_per_mev, seed=seed + 17 * i, out=(maps_d[i], tm_d[i]))
    _sync()
    t_sur_only = time.perf_counter() - t0
    transport_events_multibox(boxes, part_of[~use], op, ro, fb, seed=seed + 99, tau_ns=tau_ns,
                              tau_rise_ns=tau_rise_ns, pid0=1 << 40, ctx=ctx,
                              device_arrays=fb_in, out=fb_out)
    sur_d = fb_out[0].reshape(nev, -1)
    for m in maps_d:
        sur_d = sur_d + m
    _sync()
    t_sur = time.perf_counter() - t0
    sur_maps = sur_d.cpu().numpy().astype(np.int64)
    mc_maps = mc.maps.cpu().numpy().reshape(nev, -1).astype(np.int64)
    tot_mc = mc_maps.sum(0).astype(np.float64)
    tot_sur = sur_maps.sum(0).astype(np.float64)
    rel = (tot_sur - tot_mc) / np.maximum(tot_mc, 1)
    out = {"events": nev, "deposits": int(dep.xyzt.shape[0]), "photons_mc": int(mc.n_photons),
           "io": "device-resident deposit tables and maps on both paths",
           "fallback_deposit_fraction": float(1 - use.mean()),
           "fallback_photon_fraction": float(dep.nph[~use].sum() / max(dep.nph.sum(), 1)),
           "mc_s": t_mc, "surrogate_s": t_sur, "surrogate_sampling_only_s": t_sur_only,
           "mc_events_per_s": nev / t_mc, "surrogate_events_per_s": nev / t_sur,
           "surrogate_sampling_only_events_per_s": nev / t_sur_only,
           "speedup": t_mc / t_sur, "speedup_sampling_only": t_mc / t_sur_only,
           "per_sipm_rel_error_mean": float(np.mean(np.abs(rel))),
           "per_sipm_rel_error_max": float(np.max(np.abs(rel))),
           "energy_sum_bias": float(sur_maps.sum() / max(mc_maps.sum(), 1) - 1.0)}
    per_t = {}
    for i, b in enumerate(boxes):
        evs = np.unique(dep.event[part_of == i])
        a, s_ = mc_maps[evs].sum(0).astype(np.float64), sur_maps[evs].sum(0).astype(np.float64)
        per_t[f"{b.L[2]:g}mm"] = {"per_sipm_rel_error_mean": float(np.mean(np.abs((s_ - a) / np.maximum(a, 1)))),
                                  "energy_sum_bias": float(s_.sum() / max(a.sum(), 1) - 1.0)}
    out["by_thickness"] = per_t
    return out
