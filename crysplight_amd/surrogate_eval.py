"""Evaluation of the light-response surrogate against the Monte Carlo.

train_for_box          : the training workload (uniform points x photons each -> MLP)
validate               : per-SiPM count errors against FP64-accumulated MC means, on held-out
                         points uniform in the crystal and in near-face regions (bands of
                         distance to each face type, edges, corners), with the MC noise
                         subtracted
fallback_from_regions  : the narrowest MC-fallback rule whose complement keeps every measured
                         region within the error bar
compare_events         : gamma events in several crystal thicknesses through (a) the plain MC
                         in one batched queue and (b) the surrogate (Poisson maps from the
                         network) with MC fallback by one or more rules; events/s of both, and
                         the per-SiPM errors of (b) against (a), binned by event position
"""
from __future__ import annotations

import math
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
                  photons=10_000, epochs=1200, batch=65536, seed=11, log_depth_mm=1.0, fold=True,
                  direct=True, images=True, near_readout_frac=0.0, near_readout_mm=2.0, ctx=None):
    ctx = ctx or _lib.default_context()
    sur = S.LightSurrogate(box, op, ro, tau_ns=tau_ns, tau_rise_ns=tau_rise_ns, seed=seed,
                           log_depth_mm=log_depth_mm, fold=fold, direct=direct, images=images)
    t0 = time.perf_counter()
    ts = sur.generate(n_samples, photons, seed=seed, near_readout_frac=near_readout_frac,
                      near_readout_mm=near_readout_mm, ctx=ctx)
    _sync()
    t1 = time.perf_counter()
    # Adam's step size grows as the square root of the batch beyond 8192 rows (3e-3 there),
    # decaying 60-fold over the epochs; batch 65536 at 1200 epochs fits more accurately than
    # 8192 at 300 with the same data, and its steps run at about twice the MFMA rate
    lr = 3e-3 * math.sqrt(max(1.0, batch / 8192))
    losses = sur.fit(ts, epochs=epochs, batch=batch, lr=lr, lr_final=lr / 60, seed=seed)
    _sync()
    t2 = time.perf_counter()
    info = {"n_samples": n_samples, "photons_per_sample": photons, "generate_s": t1 - t0,
            "fit_s": t2 - t1, "epochs": epochs, "batch": batch, "lr": lr, "adam_steps": int(losses.size),
            "symmetry_folded": sur.fold, "direct_light_term": sur.direct,
            "mirror_image_terms": sur.images,
            "near_readout_samples": [near_readout_frac, near_readout_mm],
            "final_loss": float(losses[-min(50, losses.size):].mean()),
            "train_flops": float(3.0 * sur.net.flops_per_row * (losses.size * batch))}
    info["train_tflops_per_s"] = info["train_flops"] / info["fit_s"] / 1e12
    # the Adam steps alone (the fit also reshuffles the training set every epoch)
    info["step_us"] = sur.fit_step_s / max(1, losses.size) * 1e6
    info["step_tflops_per_s"] = info["train_flops"] / sur.fit_step_s / 1e12
    info["step_frac_bf16_dense_peak"] = info["step_tflops_per_s"] / 2500.0
    return sur, info


def map_metrics(p, counts, n_photons):
    """Per-SiPM errors of predicted per-photon maps p [n, ncell] against MC counts of
    n_photons photons per point, per point then averaged over points:
      count_error_raw   sum_c |p - t| / sum_c t (t = counts / n_photons; includes the MC noise)
      noise_floor       the value count_error_raw takes for a perfect p (E|t - p| of Poisson t)
      rms_rel_debiased  sqrt(sum_c [(p - t)^2 - t / N] / t / sum_c t): count-weighted RMS
                        relative error with the MC noise subtracted
      unweighted_rms_rel_debiased  the same with every cell weighted equally."""
    N = float(n_photons)
    t = counts / N
    tsum = np.maximum(t.sum(1), 1e-30)
    raw = np.abs(p - t).sum(1) / tsum
    floor = (np.sqrt(2 / np.pi) * np.sqrt(t / N)).sum(1) / tsum
    ok = t > 0
    ts = np.where(ok, t, 1.0)
    sq = np.where(ok, ((p - t) ** 2 - t / N) / ts, 0.0).sum(1) / tsum
    sq_u = np.where(ok, ((p - t) ** 2 - t / N) / ts ** 2, 0.0).sum(1) / np.maximum(ok.sum(1), 1)
    rms = np.sqrt(np.maximum(sq, 0.0))
    return {"points": int(p.shape[0]), "count_error_raw": float(raw.mean()),
            "noise_floor": float(floor.mean()), "rms_rel_debiased": float(rms.mean()),
            "rms_rel_debiased_p90": float(np.quantile(rms, 0.9)),
            "unweighted_rms_rel_debiased": float(np.sqrt(np.maximum(sq_u, 0.0)).mean())}


def region_points(L, rng, n):
    """Held-out points by region: bands of distance to the readout face, to the front face
    and to a side face, and the crystal's edges and corners (two or three faces close)."""
    Lx, Ly, Lz = (float(v) for v in L)
    mid = (1.0, Lx - 1.0)

    def pts(xr, yr, zr):
        return np.column_stack([rng.uniform(*xr, n), rng.uniform(*yr, n), rng.uniform(*zr, n)])
    out = {}
    for a, b in [(0, 0.05), (0.05, 0.1), (0.1, 0.25), (0.25, 0.5), (0.5, 1.0), (1.0, 2.0)]:
        out[f"readout_{a}-{b}mm"] = pts(mid, mid, (Lz - b, Lz - a))
    for a, b in [(0, 0.1), (0.1, 1.0)]:
        out[f"front_{a}-{b}mm"] = pts(mid, mid, (a, b))
        out[f"side_{a}-{b}mm"] = pts((a, b), mid, (1.0, Lz - 1.0))
    for a, b in [(0, 0.25), (0.25, 0.5), (0.5, 1.0)]:
        out[f"edge_side_side_{a}-{b}mm"] = pts((a, b), (a, b), (1.0, Lz - 1.0))
        out[f"edge_side_readout_{a}-{b}mm"] = pts((a, b), mid, (Lz - b, Lz - a))
        out[f"edge_side_front_{a}-{b}mm"] = pts((a, b), mid, (a, b))
        out[f"corner_readout_{a}-{b}mm"] = pts((a, b), (a, b), (Lz - b, Lz - a))
        out[f"corner_front_{a}-{b}mm"] = pts((a, b), (a, b), (a, b))
    return out


def validate(sur, *, n_points=10_000, photons=2_000_000, seed=77, region_points_n=100,
             region_photons=10_000_000, ctx=None):
    """The surrogate (network plus direct light, as sampled) against FP64-accumulated MC
    means (exact integer counts): n_points held-out positions uniform in the crystal with
    `photons` each, and region_points_n positions in each near-face region with
    region_photons each (map_metrics per set)."""
    ctx = ctx or _lib.default_context()
    L = np.asarray(sur.box.L, np.float64)
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, 1, (n_points, 3)) * L
    e = rng.uniform(*sur.e_range, n_points)
    c, tsum = S.mc_light_maps(sur.box, sur.op, sur.ro, xyz, photons, seed=seed,
                              tau_ns=sur.tau_ns, tau_rise_ns=sur.tau_rise_ns, pid0=1 << 44, ctx=ctx)
    p, tm = sur.predict(xyz, e)
    out = {"photons_per_point": photons, "uniform": map_metrics(p.astype(np.float64), c, photons)}
    with np.errstate(invalid="ignore", divide="ignore"):
        tt = tsum / np.maximum(c, 1)
    m = c > 1000
    out["time_mean_abs_err_ns"] = float(np.abs(tm[m] - tt[m]).mean()) if m.any() else float("nan")
    if region_points_n:
        reg = {}
        for k, (name, q) in enumerate(region_points(L, rng, region_points_n).items()):
            cr, _ = S.mc_light_maps(sur.box, sur.op, sur.ro, q, region_photons, seed=seed + 1 + k,
                                    tau_ns=sur.tau_ns, tau_rise_ns=sur.tau_rise_ns,
                                    pid0=1 << 45, ctx=ctx)
            pr, _ = sur.predict(q, rng.uniform(*sur.e_range, q.shape[0]))
            reg[name] = map_metrics(pr.astype(np.float64), cr, region_photons)
        out["regions"] = {"photons_per_point": region_photons, "by_region": reg}
    return out


def fallback_from_regions(validations, bar=0.02):
    """The narrowest fallback rule (near_face_mask keywords) whose complement has every
    measured region within `bar` (noise-subtracted count-weighted RMS relative error) in
    every crystal: readout_mm from the readout-face bands, edge_mm from the edge and corner
    regions.  validations: the validate() results of the crystals."""
    def worst(prefix, lo):
        v = 0.0
        for val in validations:
            for name, m in val["regions"]["by_region"].items():
                if name.startswith(prefix):
                    a = float(name[len(prefix):].split("-")[0])
                    if a >= lo:
                        v = max(v, m["rms_rel_debiased"])
        return v
    readout_mm = next((r for r in (0.0, 0.05, 0.1, 0.25, 0.5, 1.0, 2.0)
                       if worst("readout_", r) <= bar), 2.0)
    edge_mm = next((r for r in (0.0, 0.25, 0.5, 1.0)
                    if all(worst(p_, r) <= bar for p_ in ("edge_side_side_", "edge_side_readout_",
                                                          "edge_side_front_"))), 1.0)
    # corners are also edges: a corner region within edge_mm is covered already
    corner_mm = next((r for r in (0.0, 0.25, 0.5, 1.0)
                      if all(worst(p_, max(r, edge_mm)) <= bar
                             for p_ in ("corner_readout_", "corner_front_"))), 1.0)
    return {"readout_mm": readout_mm, "edge_mm": edge_mm, "corner_mm": max(corner_mm, edge_mm)}


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


def near_face_mask(xyz, L, *, any_face_mm=0.0, edge_mm=0.0, corner_mm=0.0, readout_mm=0.0):
    """Deposits for the Monte Carlo fallback: within any_face_mm of any face, within
    edge_mm of two or more faces (crystal edges), within corner_mm of three faces (corners),
    or within readout_mm of the readout face z = Lz.  xyz [n, 3], L [n, 3] (each deposit's
    crystal)."""
    xyz = np.asarray(xyz, np.float64)
    L = np.asarray(L, np.float64)
    dist = np.minimum(xyz, L - xyz)
    m = np.zeros(xyz.shape[0], bool)
    if any_face_mm > 0:
        m |= (dist < any_face_mm).any(1)
    if edge_mm > 0:
        m |= (dist < edge_mm).sum(1) >= 2
    if corner_mm > 0:
        m |= (dist < corner_mm).sum(1) >= 3
    if readout_mm > 0:
        m |= (L[:, 2] - xyz[:, 2]) < readout_mm
    return m


def binned_map_errors(mc_maps, sur_maps, dep, part_of, boxes, nbin=(4, 4, 4)):
    """Per-SiPM errors of the surrogate event maps against the plain MC's on the same
    events, binned by each event's photon-weighted deposit position (nbin per axis per
    crystal): the maps of a bin are summed on both sides and compared cell by cell.
    Returns the mean absolute relative error (count-weighted within a bin, then averaged
    over bins), its Poisson noise floor, the noise-subtracted count-weighted RMS relative
    error ((B - A)^2 - (A + B)) / A^2, and the unweighted mean over cells."""
    nev = mc_maps.shape[0]
    w = dep.nph.astype(np.float64)
    pos = np.zeros((nev, 3))
    wsum = np.zeros(nev)
    np.add.at(pos, dep.event, dep.xyzt[:, :3] * w[:, None])
    np.add.at(wsum, dep.event, w)
    ev_part = np.zeros(nev, np.int64)
    ev_part[dep.event] = part_of
    pos /= np.maximum(wsum, 1)[:, None]
    L = np.array([b.L for b in boxes], np.float64)[ev_part]
    idx = np.minimum((pos / L * np.asarray(nbin)).astype(np.int64), np.asarray(nbin) - 1)
    key = ((ev_part * nbin[0] + idx[:, 0]) * nbin[1] + idx[:, 1]) * nbin[2] + idx[:, 2]
    key = np.where(wsum > 0, key, -1)
    nb = len(boxes) * nbin[0] * nbin[1] * nbin[2]
    A = np.zeros((nb, mc_maps.shape[1]))
    B = np.zeros_like(A)
    ok = key >= 0
    np.add.at(A, key[ok], mc_maps[ok])
    np.add.at(B, key[ok], sur_maps[ok])
    use = A.sum(1) > 0
    A, B = A[use], B[use]
    good = A > 0
    Asafe = np.where(good, A, 1.0)
    rel = np.where(good, np.abs(B - A) / Asafe, 0.0)
    raw = (rel * A).sum(1) / A.sum(1)
    floor = (np.where(good, np.sqrt(2 / np.pi) * np.sqrt(A + B) / Asafe, 0.0) * A).sum(1) / A.sum(1)
    sq = (np.where(good, ((B - A) ** 2 - (A + B)) / Asafe ** 2, 0.0) * A).sum(1) / A.sum(1)
    unw = np.where(good, rel, 0.0).sum(1) / good.sum(1)
    return {"bins": int(use.sum()), "events_per_bin_mean": float(ok.sum() / max(use.sum(), 1)),
            "mean_abs_rel": float(raw.mean()), "noise_floor": float(floor.mean()),
            "rms_rel_debiased": float(np.sqrt(np.maximum(sq, 0)).mean()),
            "rms_rel_debiased_max_bin": float(np.sqrt(np.maximum(sq, 0)).max()),
            "unweighted_mean_abs_rel": float(unw.mean())}


def compare_events(surrogates, boxes, op, ro, dep, part_of, *, yield_per_mev, tau_ns,
                   tau_rise_ns=0.0, seed=5, rules=None, ctx=None):
    """(a) plain MC of every deposit in one multi-crystal queue; (b) for each fallback rule
    (name -> near_face_mask keywords), the surrogate's Poisson maps from the network for the
    other deposits plus the MC for the rule's deposits.  Both paths start from deposit
    tables resident on the device and leave their maps there (they are copied back after
    the timed regions), like the transport's photons/s.  The surrogate's per-event cost
    includes the direct-light term and the network evaluation of every deposit."""
    import torch
    ctx = ctx or _lib.default_context()
    rules = rules or {"any_face_1mm": {"any_face_mm": 1.0}}
    nev = dep.n_events
    dev = torch.device("cuda", 0)
    Ls = np.array([b.L for b in boxes], np.float32)[part_of]
    grid = ro.grid if hasattr(ro, "grid") else ro
    ncell = grid.nx * grid.ny

    def on_dev(d, part):
        return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                     for a in (d.xyzt, d.nph, d.event, part.astype(np.int32)))

    def map_out():
        return (torch.zeros((nev, grid.ny, grid.nx), dtype=torch.int32, device=dev),
                torch.zeros((nev, grid.ny, grid.nx), dtype=torch.float32, device=dev),
                torch.zeros(nev, dtype=torch.int32, device=dev))

    # (a) plain MC
    mc_in, mc_out = on_dev(dep, part_of), map_out()
    _sync()
    t0 = time.perf_counter()
    mc = transport_events_multibox(boxes, part_of, op, ro, dep, seed=seed, tau_ns=tau_ns,
                                   tau_rise_ns=tau_rise_ns, ctx=ctx, device_arrays=mc_in,
                                   out=mc_out)
    _sync()
    t_mc = time.perf_counter() - t0
    mc_maps = mc.maps.cpu().numpy().reshape(nev, -1).astype(np.int64)
    del mc_in

    # deposits and events of each crystal are contiguous (gamma_mix concatenates them)
    ev_first = np.zeros(nev + 1, np.int64)
    np.add.at(ev_first, dep.event + 1, 1)
    ev_first = np.cumsum(ev_first)
    parts = []
    for i in range(len(boxes)):
        ds = np.flatnonzero(part_of == i)
        evs = np.unique(dep.event[ds])
        dlo, dhi, elo, ehi = int(ds[0]), int(ds[-1]) + 1, int(evs[0]), int(evs[-1]) + 1
        assert dhi - dlo == ds.size and ehi - elo == evs.size
        parts.append((dlo, dhi, elo, ehi))
    xyzt_d = torch.from_numpy(dep.xyzt).to(dev)
    ek_d = torch.from_numpy(dep.e_kev).to(dev)
    evf_d = [torch.from_numpy(ev_first[elo:ehi + 1] - dlo).to(dev) for dlo, dhi, elo, ehi in parts]
    out = {"events": nev, "deposits": int(dep.xyzt.shape[0]), "photons_mc": int(mc.n_photons),
           "io": "device-resident deposit tables and maps on both paths",
           "mc_s": t_mc, "mc_events_per_s": nev / t_mc, "lines": {}}
    for r_i, (name, kw) in enumerate(rules.items()):
        fbm = near_face_mask(dep.xyzt[:, :3], Ls, **kw)
        use_d = torch.from_numpy((~fbm).astype(np.uint8)).to(dev)
        fb = Deposits(dep.xyzt[fbm], dep.nph[fbm], dep.event[fbm], dep.e_kev[fbm], nev)
        fb_in, fb_out = on_dev(fb, part_of[fbm]), map_out()
        sur_maps = torch.zeros((nev, ncell), dtype=torch.int32, device=dev)
        sur_tm = torch.zeros((nev, ncell), dtype=torch.float32, device=dev)
        _sync()
        t0 = time.perf_counter()
        for i, (s_, (dlo, dhi, elo, ehi)) in enumerate(zip(surrogates, parts)):
            s_.sample(Deposits(xyzt_d[dlo:dhi], None, None, ek_d[dlo:dhi], ehi - elo), evf_d[i],
                      use_d[dlo:dhi], yield_per_mev, seed=seed + 17 * i + 1000 * r_i,
                      out=(sur_maps[elo:ehi], sur_tm[elo:ehi]))
        _sync()
        t_samp = time.perf_counter() - t0
        if fbm.any():
            transport_events_multibox(boxes, part_of[fbm], op, ro, fb, seed=seed + 99, tau_ns=tau_ns,
                                      tau_rise_ns=tau_rise_ns, pid0=1 << 40, ctx=ctx,
                                      device_arrays=fb_in, out=fb_out)
            sur_maps += fb_out[0].reshape(nev, -1)
        _sync()
        t_sur = time.perf_counter() - t0
        sm = sur_maps.cpu().numpy().astype(np.int64)
        tot_mc, tot_sur = mc_maps.sum(0).astype(np.float64), sm.sum(0).astype(np.float64)
        rel = (tot_sur - tot_mc) / np.maximum(tot_mc, 1)
        line = {"fallback_rule": kw,
                "fallback_deposit_fraction": float(fbm.mean()),
                "fallback_photon_fraction": float(dep.nph[fbm].sum() / max(dep.nph.sum(), 1)),
                "surrogate_s": t_sur, "surrogate_sampling_only_s": t_samp,
                "surrogate_events_per_s": nev / t_sur,
                "speedup": t_mc / t_sur, "speedup_sampling_only": t_mc / t_samp,
                "grand_sum_per_sipm_rel_error_mean": float(np.mean(np.abs(rel))),
                "energy_sum_bias": float(sm.sum() / max(mc_maps.sum(), 1) - 1.0),
                "position_binned": binned_map_errors(mc_maps, sm, dep, part_of, boxes)}
        per_t = {}
        for i, b in enumerate(boxes):
            dlo, dhi, elo, ehi = parts[i]
            a_, s2 = mc_maps[elo:ehi].sum(0).astype(np.float64), sm[elo:ehi].sum(0).astype(np.float64)
            per_t[f"{b.L[2]:g}mm"] = {"energy_sum_bias": float(s2.sum() / max(a_.sum(), 1) - 1.0)}
        line["by_thickness"] = per_t
        out["lines"][name] = line
        del fb_in, fb_out, sur_maps, sur_tm
    return out
