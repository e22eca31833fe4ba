"""Surface/bulk realism scan for CsI + Teflon, the reference's scripts/study_surface.jl on
the MI355X transport.

Centred point source in a 48x48x37.2 mm CsI crystal (n 1.95, coupling 1.45, Teflon
Lambertian R 0.99, PDE 1, delta emission), 1e6 photons with Philox seed 1 per case, over
nine surface/bulk models (scripts/study_surface.jl:8-26).  The reference publishes the
resulting fractions (latex/crysp_light_metal.tex:610-631, table tab:surfscan).

    python -m crysplight_amd.surface_scan
"""
from __future__ import annotations

import math

from .params import Box, OpticalParams, SipmGrid, TimeBinning, deg2rad, readout

BOX = Box((48.0, 48.0, 37.2))
POS = (24.0, 24.0, 18.6)
GRID = SipmGrid(6.0, 8, 8)
TB = TimeBinning(100.0, 40)
N_PHOTONS = 1_000_000
SEED = 1


def optics(abs_mm=500.0, ray_mm=math.inf, gap=True, sigma_deg=0.0) -> OpticalParams:
    """mk(...) of scripts/study_surface.jl:10-12."""
    return OpticalParams(1.95, 1.45, abs_mm, ray_mm, 0.99, False, gap, deg2rad(sigma_deg), 1.0)


CASES = [
    ("gap, polished (v0 baseline)", optics()),
    ("gap, sigma_alpha 1.3 deg", optics(sigma_deg=1.3)),
    ("gap, sigma_alpha 6 deg", optics(sigma_deg=6.0)),
    ("gap, sigma_alpha 12 deg", optics(sigma_deg=12.0)),
    ("contact (frontpainted) Lambertian", optics(gap=False)),
    ("gap, polished + Rayleigh 300 mm", optics(ray_mm=300.0)),
    ("gap, 1.3 deg + Rayleigh 300 mm", optics(sigma_deg=1.3, ray_mm=300.0)),
    ("gap, polished, Lambda_abs 5 m", optics(abs_mm=5000.0)),
    ("gap, 1.3 deg, Lambda_abs 5 m", optics(abs_mm=5000.0, sigma_deg=1.3)),
]


def run(ctx=None, n_photons: int = N_PHOTONS, verbose: bool = True) -> list[dict]:
    """Every case through the device accumulate mode; fractions in percent."""
    from .transport import run_photons
    out = []
    for name, op in CASES:
        acc = run_photons(BOX, op, readout(GRID), TB, n_photons=n_photons, seed=SEED, pos=POS,
                          tau_ns=0.0, ctx=ctx)
        row = {"case": name, "detected": 100.0 * acc.ndet / n_photons,
               "bulk": 100.0 * acc.nabs_bulk / n_photons, "wall": 100.0 * acc.nabs_wall / n_photons,
               "bounces_per_detected": acc.sum_bounces_det / max(acc.ndet, 1)}
        out.append(row)
        if verbose:
            print(f"{name:<36s} det {row['detected']:6.2f}%  bulk {row['bulk']:6.2f}%  "
                  f"wall {row['wall']:6.2f}%  <b|det> {row['bounces_per_detected']:5.1f}")
    return out


if __name__ == "__main__":
    run()
