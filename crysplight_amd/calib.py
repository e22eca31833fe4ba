"""Calibration campaigns: parameter sweeps over the accumulate path on the MI355X.

csitl_ladder   <- scripts/csitl_calib.jl (Brose wrapping ladder, disc readout + surround)
ding_cube      <- scripts/ding_calib.jl (BGO 6 mm cube, two-sided readout)
soleti_crysp   <- scripts/soleti_calib.jl (cold CsI 3x3x20 mm, full-face readout)
fit_R          <- fit_R (scripts/csitl_calib.jl:57-68)

Each campaign reads a run file of the reference's layout (crysplight_amd/runs/*_calib.toml),
runs one accumulate-mode transport per grid point through crysp_transport_accumulate,
prints the reference's table and writes output/<tag>/summary.csv with its columns.
The parameters are built with the reference's Float32/Float64 promotions (for example
Λ_abs = Float32(Float32(Λ_ext) / (1 − f)) in the CsI(Tl) ladder), so that a grid point
gives the same OpticalParams bits as the reference's script.

    python -m crysplight_amd.calib crysplight_amd/runs/csitl_calib.toml [more...]
"""
from __future__ import annotations

import math
import shutil
import sys
from pathlib import Path

import numpy as np

from .config import _toml, load_material
from .params import Box, OpticalParams, SipmGrid, TimeBinning, readout
from .transport import transport_accumulate

f32 = np.float32


def _deg2rad32(deg) -> float:
    return float(f32(float(deg) * (math.pi / 180.0)))


def _load(cfg_path):
    cfg_path = Path(cfg_path)
    with open(cfg_path, "rb") as f:
        return cfg_path, _toml.load(f)


def _outdir(cfg_path: Path, tag: str) -> Path:
    return cfg_path.resolve().parent.parent / "output" / tag


def _num(x) -> str:
    """A float as Julia prints it in a CSV row (shortest round-trip form, NaN as NaN)."""
    if isinstance(x, float) and math.isnan(x):
        return "NaN"
    return repr(float(x)) if isinstance(x, (float, np.floating)) else str(x)


def _write_csv(cfg_path: Path, tag: str, header: str, rows) -> Path:
    out = _outdir(cfg_path, tag)
    out.mkdir(parents=True, exist_ok=True)
    shutil.copy(cfg_path, out / cfg_path.name)
    p = out / "summary.csv"
    with open(p, "w") as f:
        f.write(header + "\n")
        for r in rows:
            f.write(",".join(_num(v) for v in r) + "\n")
    return p


def _collection(box, op, ro, tb, positions, nph, seed, maxb, ctx) -> float:
    acc = transport_accumulate(box, op, ro, tb, n_photons=nph, seed=seed, positions=positions,
                               tau_ns=0.0, max_bounces=maxb, ctx=ctx)
    if acc.total_terminated() != nph:
        raise RuntimeError(f"{acc.total_terminated()} of {nph} photons terminated")
    return acc.ndet / nph


def fit_R(Rs, fracs, base, target) -> float:
    """The R at which frac(R)/base crosses target, by linear interpolation between the first
    bracketing pair of scan points; NaN when the scan does not bracket it."""
    ratios = [fr / base for fr in fracs]
    for i in range(len(Rs) - 1):
        lo, hi = ratios[i], ratios[i + 1]
        if (lo - target) * (hi - target) <= 0:
            w = (target - lo) / (hi - lo)
            return Rs[i] + w * (Rs[i + 1] - Rs[i])
    return math.nan


def stratified_depths(att_mm: float, length_mm: float, n: int) -> list:
    """Equal-weight quantiles (k − 1/2)/n of an exponential depth profile (attenuation
    att_mm) truncated at length_mm, as Float32 (scripts/soleti_calib.jl:27-31)."""
    cdf_end = 1 - math.exp(-length_mm / att_mm)
    return [float(f32(-att_mm * math.log(1 - ((k - 0.5) / n) * cdf_end))) for k in range(1, n + 1)]


def csitl_ladder(cfg_path, ctx=None, data_dir=None, write=True, verbose=True,
                 lambda_ext_mm=None) -> list:
    """scripts/csitl_calib.jl: for each (f_split, sigma_alpha) the bare collection and,
    per Teflon case, the wrap reflectivity that reproduces the measured LY ratio.

    The extinction that the grid splits is recombined from the material's absorption and
    Rayleigh lengths, as the reference's script does; `lambda_ext_mm` sets it instead (the
    measured 311.5 mm, where the material's rounded split recombines to 311.41 mm)."""
    cfg_path, cfg = _load(cfg_path)
    camp = cfg["campaign"]
    mat = load_material(cfg["crystal"]["material"], data_dir, cfg_path)
    sz = [float(f32(v)) for v in cfg["crystal"]["size_mm"]]
    box = Box(sz)
    # The grid splits the measured extinction; recombine it from the material's split.
    if lambda_ext_mm is None:
        lam_ext = f32(1 / (1 / float(mat["abs_length_mm"]) + 1 / float(mat.get("rayleigh_mm", math.inf))))
    else:
        lam_ext = f32(lambda_ext_mm)
    n_crystal = float(f32(mat["n"]))
    rd = cfg["readout"]
    grid = SipmGrid(float(f32(sz[0]) / f32(8)), 8, 8)      # light-map binning only
    tb = TimeBinning(100.0, 40)
    src = cfg["source"]
    xy = [float(f32(v)) for v in src["xy_mm"]]
    positions = [(xy[0], xy[1], float(f32(z))) for z in src["positions_z_mm"]]
    nph, seed, maxb = int(camp["n_photons"]), int(camp["seed"]), int(camp["max_bounces"])

    def collection(f, sigma_deg, wrap_R, sur_R):
        lam_abs = float(lam_ext) / (1 - f) if f < 1 else math.inf
        lam_ray = float(lam_ext) / f if f > 0 else math.inf
        op = OpticalParams(n_crystal, float(f32(rd["coupling_n"])), float(f32(lam_abs)),
                           float(f32(lam_ray)), float(f32(wrap_R)), False, True,
                           _deg2rad32(sigma_deg), float(f32(rd["pde"])))
        ro = readout(grid, center=tuple(rd["disc_center_mm"]), radius=rd["disc_radius_mm"],
                     sur_R=float(f32(sur_R)))
        return _collection(box, op, ro, tb, positions, nph, seed, maxb, ctx)

    cases = cfg["cases"]
    bare = [c for c in cases if c["name"] == "bare"]
    if len(bare) != 1:
        raise ValueError("the campaign needs exactly one case named 'bare'")
    bare = bare[0]
    teflon = [c for c in cases if "wrap_R_scan" in c]
    ly_bare = bare["target_ly"]
    results = []
    for f in cfg["grid"]["f_split"]:
        for sig in cfg["grid"]["sigma_alpha_deg"]:
            fb = collection(f, sig, bare["wrap_R"], bare["surround_R"])
            if verbose:
                print(f"\nf = {f:.2f}  sigma_alpha = {sig:.1f} deg   bare collection = {fb:.4f}")
            row = {"f": f, "sigma": sig, "bare_frac": fb, "fracs": {}}
            for c in teflon:
                Rs = [float(R) for R in c["wrap_R_scan"]]
                fr = [collection(f, sig, R, R if c.get("surround", "") == "wrap" else c.get("surround_R", 0.0))
                      for R in Rs]
                target = c["target_ly"] / ly_bare
                Rfit = fit_R(Rs, fr, fb, target)
                row[c["name"]] = Rfit
                row["fracs"][c["name"]] = fr
                if verbose:
                    print(f"  {c['name']:<18s} target ratio {target:5.2f}   fracs "
                          + " ".join(f"{x:.3f}" for x in fr) + "   R_eff = "
                          + ("out of scan range" if math.isnan(Rfit) else f"{Rfit:.3f}"))
            results.append(row)
    names = [c["name"] for c in teflon]
    if write:
        p = _write_csv(cfg_path, camp["tag"], "f_split,sigma_alpha_deg,bare_frac,"
                       + ",".join("Reff_" + n for n in names),
                       [[r["f"], r["sigma"], r["bare_frac"]] + [r.get(n, math.nan) for n in names]
                        for r in results])
    if verbose:
        print("\n==== summary (R_eff fitted to each measured LY ratio) ====")
        print(f"{'f':<6s} {'sigma':<10s} {'bare_frac':<10s} " + "".join(f"{n:>20s}" for n in names))
        for r in results:
            print(f"{r['f']:<6.2f} {r['sigma']:<10.1f} {r['bare_frac']:<10.4f} "
                  + "".join(f"{('—' if math.isnan(r[n]) else f'{r[n]:.3f}'):>20s}" for n in names))
        if write:
            print(f"\nwrote {p}")
    return results


def _eps_grid_params(n_crystal, coupling, lam_ext, f, wrap_R, sigma):
    lam_abs = lam_ext / (1 - f) if f < 1 else math.inf
    lam_ray = lam_ext / f if f > 0 else math.inf
    return OpticalParams(n_crystal, coupling, float(f32(lam_abs)), float(f32(lam_ray)),
                         float(f32(wrap_R)), False, True, sigma, 1.0)


def ding_cube(cfg_path, ctx=None, data_dir=None, write=True, verbose=True, extra_R=()) -> list:
    """scripts/ding_calib.jl: detection efficiency of the two-sided 6 mm BGO cube over
    (sigma_alpha, Λ_ext) × wrap_R.  `extra_R` appends reflectivities to the scan."""
    cfg_path, cfg = _load(cfg_path)
    camp = cfg["campaign"]
    mat = load_material(cfg["crystal"]["material"], data_dir, cfg_path)
    sz = [float(f32(v)) for v in cfg["crystal"]["size_mm"]]
    box = Box(sz)
    n_crystal = float(f32(mat["n"]))
    ro = readout(SipmGrid(sz[0], 1, 1), two_sided=bool(cfg["readout"]["two_sided"]))
    tb = TimeBinning(100.0, 80)
    src = cfg["source"]
    yz = [float(f32(v)) for v in src["yz_mm"]]
    positions = [(x, yz[0], yz[1]) for x in stratified_depths(float(src["att_mm"]), sz[0], int(src["n_points"]))]
    nph, seed, maxb = int(camp["n_photons"]), int(camp["seed"]), int(camp["max_bounces"])
    coupling = float(f32(cfg["readout"]["coupling_n"]))
    f = float(cfg["grid"]["f_split"])
    Rs = [float(R) for R in cfg["grid"]["wrap_R"]] + [float(R) for R in extra_R]
    if verbose:
        t = cfg["targets"]
        print(f"target: eps = {t['eps_measured']:.3f} +- 0.08 (measured, PDE band); their Geant4 MC "
              f"{t['eps_their_mc']:.2f}\n")
        print(f"{'sigma_alpha':<12s} {'Lam_ext':<10s} | " + "   ".join(f"R={R:.2f}" for R in Rs))
    results = []
    for sig in cfg["grid"]["sigma_alpha_deg"]:
        for lam in cfg["grid"]["lambda_ext_mm"]:
            eps = [_collection(box, _eps_grid_params(n_crystal, coupling, float(lam), f, R, _deg2rad32(sig)),
                               ro, tb, positions, nph, seed, maxb, ctx) for R in Rs]
            results.append((float(sig), float(lam), eps))
            if verbose:
                print(f"{sig:<12.0f} {lam:<10.0f} | " + "    ".join(f"{e:.3f}" for e in eps))
    if write:
        p = _write_csv(cfg_path, camp["tag"], "sigma_alpha_deg,lambda_ext_mm,"
                       + ",".join(f"eps_R{R!r}" for R in Rs), [[s, l] + e for s, l, e in results])
        if verbose:
            print(f"\nwrote {p}")
    return results


def soleti_crysp(cfg_path, ctx=None, data_dir=None, write=True, verbose=True) -> list:
    """scripts/soleti_calib.jl: detection efficiency of the 3x3x20 mm cold-CsI crystal over
    (Λ_ext, f) × wrap_R with one SiPM over the whole z = L face."""
    cfg_path, cfg = _load(cfg_path)
    camp = cfg["campaign"]
    mat = load_material(cfg["crystal"]["material"], data_dir, cfg_path)
    sz = [float(f32(v)) for v in cfg["crystal"]["size_mm"]]
    box = Box(sz)
    n_crystal = float(f32(mat["n"]))
    ro = readout(SipmGrid(sz[0], 1, 1))
    tb = TimeBinning(100.0, 80)
    src = cfg["source"]
    xy = [float(f32(v)) for v in src["xy_mm"]]
    positions = [(xy[0], xy[1], z) for z in stratified_depths(float(src["att_mm"]), sz[2], int(src["n_points"]))]
    nph, seed, maxb = int(camp["n_photons"]), int(camp["seed"]), int(camp["max_bounces"])
    coupling = float(f32(cfg["readout"]["coupling_n"]))
    sigma = _deg2rad32(cfg["surface"]["sigma_alpha_deg"])
    Rs = [float(R) for R in cfg["grid"]["wrap_R"]]
    if verbose:
        t1, t2 = cfg["targets"]["eps_measured"]
        print(f"targets: eps = {t1:.3f} (crystal 1) / {t2:.3f} (crystal 2), their LUT MC "
              f"{cfg['targets']['eps_their_mc']:.2f}\n")
        print(f"{'Lambda_ext':<14s} {'f':<6s} | " + "   ".join(f"R={R:.2f}" for R in Rs))
    results = []
    for lam in cfg["grid"]["lambda_ext_mm"]:
        for f in cfg["grid"]["f_split"]:
            eps = [_collection(box, _eps_grid_params(n_crystal, coupling, float(lam), float(f), R, sigma),
                               ro, tb, positions, nph, seed, maxb, ctx) for R in Rs]
            results.append((float(lam), float(f), eps))
            if verbose:
                print(f"{lam:<14.0f} {f:<6.1f} | " + "    ".join(f"{e:.3f}" for e in eps))
    if write:
        p = _write_csv(cfg_path, camp["tag"], "lambda_ext_mm,f_split,"
                       + ",".join(f"eps_R{R!r}" for R in Rs), [[l, f] + e for l, f, e in results])
        if verbose:
            print(f"\nwrote {p}")
    return results


CAMPAIGNS = {"csitl_calib": csitl_ladder, "ding_calib": ding_cube, "soleti_calib": soleti_crysp}


def run_campaign(cfg_path, **kw):
    """Dispatch on the campaign tag (csitl_calib, ding_calib or soleti_calib)."""
    _, cfg = _load(cfg_path)
    tag = cfg["campaign"]["tag"]
    for key, fn in CAMPAIGNS.items():
        if tag.startswith(key):
            return fn(cfg_path, **kw)
    raise ValueError(f"unknown campaign tag {tag!r}: expected one of {sorted(CAMPAIGNS)}")


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m crysplight_amd.calib <campaign.toml> [more...]", file=sys.stderr)
        return 2
    for p in argv:
        run_campaign(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
