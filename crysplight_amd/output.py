"""Run outputs and summaries.

write_output       <- write_output (src/output.jl:4-52): counts, first_time_ns, attributes
symmetry_residual  <- symmetry_residual (src/output.jl:58-62)
write_events       <- the events.h5 writer of scripts/shoot_gammas.jl:12-17,60-90

Both files are HDF5 (written by crysplight_amd.h5) under the reference's names and
dtypes, with the arrays in the order h5py presents the reference's Julia (column-major)
files: light counts as (nbins+1, ny, nx), first_time_ns (ny, nx); events maps/tmin
(n, ny, nx), xyz (n, 3).  Root attributes carry the reference's keys.  The event
datasets are chunked along the event axis, 4096 events per chunk with the full other
dimensions (scripts/shoot_gammas.jl:11-17).  maps are narrowed to UInt16 modulo 2^16:
the reference counts each photoelectron with `maps[ii, jj, ev] += UInt16(1)`
(src/events.jl:49,81), and Julia's UInt16 addition wraps.
"""
from __future__ import annotations

import shutil
from pathlib import Path

import numpy as np

from . import h5


def symmetry_residual(m: np.ndarray) -> float:
    """Max relative deviation of a map from its 8-fold symmetrization (x-mirror, y-mirror,
    transpose): 0 for a symmetric map."""
    m = np.asarray(m, np.float64)
    ms = (m + m[::-1, :] + m[:, ::-1] + m[::-1, ::-1]
          + m.T + m.T[::-1, :] + m.T[:, ::-1] + m.T[::-1, ::-1]) / 8
    return float(np.max(np.abs(m - ms) / np.maximum(ms, 1)))


def write_output(st, acc, elapsed: float, write: bool = True, verbose: bool = True) -> dict:
    n = st.n_photons
    fdet = acc.ndet / n
    if write:
        st.outdir.mkdir(parents=True, exist_ok=True)
        shutil.copy(st.cfg_path, st.outdir / st.cfg_path.name)
        surf = st.cfg.get("surface", {})
        attrs = {"tag": st.tag, "n_photons": n, "seed": st.seed, "bin_ns": np.float32(st.tb.bin_ns),
                 "nbins": np.int32(st.tb.nbins), "crystal": st.cfg["crystal"]["material"],
                 "wrap": st.cfg["wrap"]["material"], "pde": np.float32(st.op.pde), "ndet": acc.ndet,
                 "nabs_bulk": acc.nabs_bulk, "nabs_wall": acc.nabs_wall,
                 "nabs_sipm": acc.nabs_sipm, "ncap": acc.ncap, "nabs_surround": acc.nabs_sur,
                 "nscat": acc.nscat, "finish": surf.get("finish", "backpainted"),
                 "sigma_alpha_deg": float(surf.get("sigma_alpha_deg", 0.0))}
        h5.write(st.outdir / "light.h5",
                 {"counts": np.asarray(acc.counts, np.uint32).transpose(2, 1, 0),
                  "first_time_ns": np.asarray(acc.first_ns, np.float32).T}, attrs)
    m = acc.counts.sum(axis=2)
    sym = symmetry_residual(m)
    mean_b = acc.sum_bounces_det / acc.ndet if acc.ndet else float("nan")
    summary = {"tag": st.tag, "n_photons": n, "detected_fraction": fdet,
               "abs_bulk_fraction": acc.nabs_bulk / n, "abs_wall_fraction": acc.nabs_wall / n,
               "abs_sipm_fraction": acc.nabs_sipm / n, "cap_fraction": acc.ncap / n,
               "mean_bounces_detected": mean_b, "symmetry_residual": sym, "elapsed_s": elapsed}
    if verbose:
        print(f"{st.tag:<22s}  detected {100 * fdet:6.2f}%   bulk {100 * acc.nabs_bulk / n:6.2f}%   "
              f"wall {100 * acc.nabs_wall / n:6.2f}%   cap {acc.ncap / n:.2e}")
        print(f"{'':<22s}  <bounces|det> {mean_b:.1f}   symmetry residual {sym:.3f}   "
              f"{elapsed:.3f}s (MI355X)")
    return summary


EVENT_CHUNK = 4096


def uint16_maps(maps) -> np.ndarray:
    """Per-event SiPM counts as the reference's UInt16 matrix holds them: modulo 2^16."""
    maps = np.asarray(maps)
    if maps.dtype.kind not in "iu" or (maps.dtype.kind == "i" and maps.size and maps.min() < 0):
        raise ValueError("maps must hold non-negative integer counts")
    return (maps.astype(np.uint64) & 0xFFFF).astype(np.uint16)


def write_events(c, r: dict, elapsed: float, outdir, write: bool = True, verbose: bool = True) -> dict:
    nev = c.n_events
    maps = np.asarray(r["maps"])
    if write:
        outdir = Path(outdir)
        outdir.mkdir(parents=True, exist_ok=True)
        shutil.copy(c.cfg_path, outdir / c.cfg_path.name)
        attrs = {"tag": c.tag, "n_events": nev, "seed": c.seed,
                 "crystal": c.cfg["crystal"]["material"], "pde": np.float32(c.op.pde), "backend": "hip-gfx950",
                 "int_type_key": "-1 = no interaction, 0 = direct photoelectric, X>=1 = X Compton scatters",
                 "tmin_nohit": "Inf"}
        h5.write(outdir / "events.h5", {
            "edep_kev": np.asarray(r["edep"], np.float32), "npe": np.asarray(r["npe"], np.int32),
            "maps": uint16_maps(maps), "tmin_ns": np.asarray(r["tmin"], np.float32),
            "xyz1_mm": np.asarray(r["xyz1"], np.float32), "e1_kev": np.asarray(r["e1"], np.float32),
            "xyz2_mm": np.asarray(r["xyz2"], np.float32), "e2_kev": np.asarray(r["e2"], np.float32),
            "er_kev": np.asarray(r["er"], np.float32), "int_type": np.asarray(r["int_type"], np.int8),
            "n_int": np.asarray(r["n_int"], np.int16)}, attrs, chunk_events=EVENT_CHUNK)
    edep = np.asarray(r["edep"])
    npe = np.asarray(r["npe"])
    inter = edep > 0
    peak = edep > 510.0
    summ = {"tag": c.tag, "n_events": nev, "elapsed_s": elapsed,
            "interacted_fraction": float(inter.mean()), "photopeak_fraction": float(peak.mean()),
            "n_photons": int(r.get("n_photons", 0))}
    if peak.any():
        m, s = float(npe[peak].mean()), float(npe[peak].std(ddof=1)) if peak.sum() > 1 else 0.0
        summ.update(photopeak_pe_mean=m, photopeak_fwhm_over_mean=2.355 * s / m if m else float("nan"))
    if verbose:
        print(f"{c.tag:<14s} {nev} events ({elapsed:.2f}s, hip-gfx950): interacted "
              f"{100 * summ['interacted_fraction']:.1f}%, photopeak {100 * summ['photopeak_fraction']:.1f}%")
        if peak.any():
            print(f"{'':<14s} photopeak: {summ['photopeak_pe_mean']:.0f} pe mean, FWHM/mean "
                  f"{100 * summ['photopeak_fwhm_over_mean']:.2f}% (photostatistics+transport only)")
    return summ
