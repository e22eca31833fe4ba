"""Calibration campaigns (scripts/{csitl,ding,soleti}_calib.jl) on the MI355X.

1. One grid point of each campaign, GPU accumulate against the oracle, bit for bit: the
   disc readout with a reflecting surround, Rayleigh scattering, nine source points (CsI(Tl)
   ladder); the two-sided readout on rough faces (BGO cube); the full-face readout of a
   long thin crystal (cold CsI).
2. The CsI(Tl) wrapping-ladder table the reference publishes
   (latex/csitl_calib.tex:227-238): bare collection and fitted Teflon reflectivities for
   six (f, sigma_alpha) grid points, from runs/csitl_calib.toml (4e5 photons, seed 1).
   The table was made with the measured extinction, 311.5 mm (the campaign's header and
   latex/csitl_calib.tex:275), before the fitted split was written into data/csitl.toml,
   whose rounded lengths (3115 mm, 346 mm) recombine to 311.41 mm.  At 311.5 mm all 13
   fitted reflectivities and all 10 out-of-range marks print as published, and five of the
   six bare fractions do; the sixth, (0.95, 8.5 deg), is 0.06046 against a printed 0.061:
   16 photons of 4e5 below the rounding edge, 0.1 sigma.  At 311.41 mm two reflectivities
   move by 0.001 (0.868, 0.753 against 0.867, 0.754).
3. The BGO-cube and cold-CsI statements of latex/csitl_calib.tex:360-369,436-443.
"""
import math

import numpy as np
import pytest

import oracle
from crysplight_amd import calib
from crysplight_amd.config import HERE
from crysplight_amd.params import Box, OpticalParams, SipmGrid, TimeBinning, readout
from crysplight_amd.transport import transport_accumulate

pytestmark = pytest.mark.gpu
f32 = np.float32
RUNS = HERE / "runs"


def _same(acc, ref):
    assert np.array_equal(acc.counts, ref["counts"])
    assert np.array_equal(acc.first_ns.view(np.uint32), ref["first_ns"].view(np.uint32))
    for k in oracle.COUNTER_NAMES:
        assert getattr(acc, k) == ref[k], k


POINTS = {
    # CsI(Tl) ladder, f = 0.9, sigma_alpha 1.3 deg, Teflon R = 0.95 on the wrap and surround
    "csitl_ladder": (Box((60.0, 55.0, 360.0)),
                     OpticalParams(1.79, 1.0, float(f32(311.4 / 0.1)), float(f32(311.4 / 0.9)), 0.95,
                                   False, True, calib._deg2rad32(1.3), 1.0),
                     readout(SipmGrid(7.5, 8, 8), center=(30.0, 27.5), radius=25.0, sur_R=0.95),
                     TimeBinning(100.0, 40),
                     [(30.0, 27.5, float(z)) for z in range(20, 360, 40)]),
    # BGO cube, sigma_alpha 12 deg, Lambda_ext 2 m split f = 0.9, R = 0.97, two-sided
    "ding_cube": (Box((6.0, 6.0, 6.0)),
                  OpticalParams(2.15, 1.0, float(f32(2000.0 / 0.1)), float(f32(2000.0 / 0.9)), 0.97,
                                False, True, calib._deg2rad32(12.0), 1.0),
                  readout(SipmGrid(6.0, 1, 1), two_sided=True), TimeBinning(100.0, 80),
                  [(x, 3.0, 3.0) for x in calib.stratified_depths(0.6, 6.0, 16)]),
    # cold CsI, Lambda_ext 1 m split f = 0.5, R = 0.95
    "soleti_crysp": (Box((3.0, 3.0, 20.0)),
                     OpticalParams(1.95, 1.0, float(f32(1000.0 / 0.5)), float(f32(1000.0 / 0.5)), 0.95,
                                   False, True, calib._deg2rad32(1.3), 1.0),
                     readout(SipmGrid(3.0, 1, 1)), TimeBinning(100.0, 80),
                     [(1.5, 1.5, z) for z in calib.stratified_depths(23.0, 20.0, 32)]),
}


@pytest.mark.parametrize("name", sorted(POINTS))
def test_campaign_point_matches_oracle(name, ctx):
    box, op, ro, tb, positions = POINTS[name]
    kw = dict(n_photons=20_000, seed=1, positions=positions, tau_ns=0.0, max_bounces=100_000)
    acc = transport_accumulate(box, op, ro, tb, ctx=ctx, **kw)
    ref = oracle.run_photons(box, op, ro, tb, **kw)
    _same(acc, ref)
    assert acc.total_terminated() == 20_000


# (f, sigma_alpha) -> bare collection, R_eff for 1x38um, 3x38um, 4x38um, 200um mono;
# ">" / "<": the target ratio lies above the scan's ceiling / below its floor.
LADDER = {
    (0.80, 1.3): ("0.054", "0.867", ">", ">", "0.966"),
    (0.80, 8.5): ("0.054", "0.778", "0.963", ">", "0.868"),
    (0.90, 1.3): ("0.060", "0.754", "0.934", "0.965", "0.840"),
    (0.90, 8.5): ("0.058", "<", "0.880", "0.910", "0.793"),
    (0.95, 1.3): ("0.064", "<", "0.866", "0.896", "0.781"),
    (0.95, 8.5): ("0.061", "<", "<", "<", "<"),
}
CASES = ("teflon_1x38um", "teflon_3x38um", "teflon_4x38um", "teflon_200um_mono")


@pytest.fixture(scope="module")
def ladder(ctx):
    return calib.csitl_ladder(RUNS / "csitl_calib.toml", ctx=ctx, write=False, verbose=True,
                              lambda_ext_mm=311.5)


def _fitted(row, case, targets):
    R = row[case]
    if not math.isnan(R):
        return f"{R:.3f}"
    ratios = [fr / row["bare_frac"] for fr in row["fracs"][case]]
    return ">" if max(ratios) < targets[case] else "<"


def test_csitl_ladder_reproduces_published_table(ladder):
    targets = {"teflon_1x38um": 34 / 15, "teflon_3x38um": 53 / 15, "teflon_4x38um": 59 / 15,
               "teflon_200um_mono": 41 / 15}
    got = {(r["f"], r["sigma"]): tuple(_fitted(r, c, targets) for c in CASES) for r in ladder}
    bare = {(r["f"], r["sigma"]): r["bare_frac"] for r in ladder}
    print(got, bare)
    assert set(got) == set(LADDER)
    for key, want in LADDER.items():
        assert got[key] == want[1:], (key, got[key], want)
        # printed to 3 decimals: inside the rounding interval, widened by one binomial sigma
        p, sig = float(want[0]), math.sqrt(bare[key] * (1 - bare[key]) / 400_000)
        assert abs(bare[key] - p) <= 0.0005 + sig, (key, bare[key], want[0])
    assert sum(f"{bare[k]:.3f}" == LADDER[k][0] for k in LADDER) >= 5


def test_ding_cube_statements(ctx):
    """latex/csitl_calib.tex:438-443: sigma_alpha and Lambda_ext move eps by < 0.5%; the
    measured 0.825 is reproduced at R = 0.90 and the authors' Monte Carlo 0.80 at R = 0.88."""
    extra = (0.86, 0.87, 0.88, 0.89, 0.90, 0.91, 0.92)
    rows = calib.ding_cube(RUNS / "ding_calib.toml", ctx=ctx, write=False, verbose=True, extra_R=extra)
    eps = np.array([e for _, _, e in rows])
    grid = eps[:, :3]                                     # the campaign's own R = 0.95, 0.97, 0.99
    assert np.all(grid.max(axis=0) - grid.min(axis=0) < 0.005 * grid.mean(axis=0))
    Rs = [0.95, 0.97, 0.99] + list(extra)
    order = np.argsort(Rs)
    Rs = [Rs[i] for i in order]
    for e in eps:
        e = [e[i] for i in order]
        assert round(calib.fit_R(Rs, e, 1.0, 0.825), 2) == 0.90
        assert round(calib.fit_R(Rs, e, 1.0, 0.80), 2) == 0.88


def test_soleti_degeneracy_line(ctx):
    """latex/csitl_calib.tex:360-369: the measured eps = 0.529 is reproduced at
    (0.3 m, f 0.9, R 0.93), (1 m, f 0.5, R 0.95-0.97) and (3 m, f 0, R 0.93)."""
    rows = {(lam, f): e for lam, f, e in calib.soleti_crysp(RUNS / "soleti_calib.toml", ctx=ctx,
                                                           write=False, verbose=True)}
    Rs = [0.93, 0.95, 0.97]
    for key in ((300.0, 0.9), (3000.0, 0.0)):
        e = rows[key]
        assert int(np.argmin([abs(x - 0.529) for x in e])) == 0, (key, e)
    e = rows[(1000.0, 0.5)]
    assert e[1] <= 0.529 <= e[2], e
    assert Rs == [0.93, 0.95, 0.97]


@pytest.mark.parametrize("lam,f,published", [(300.0, 0.9, 82.9), (1000.0, 0.5, 63.9), (3000.0, 0.0, 62.2)])
def test_soleti_solutions_in_the_crysp_geometry(lam, f, published, ctx):
    """latex/csitl_calib.tex:363-369, last column: the CRYSP-geometry collection of each
    degeneracy solution, with the CRYSP detector's saturated Teflon (R = 0.99) and the
    cold-CsI run file otherwise (runs/point_csi_teflon.toml: 1e6 photons, seed 1234).
    Ours: 82.946, 63.921, 62.303 %; the first two print as published, the third is 0.1 point
    (2 sigma at 1e6 photons) above its printed 62.2."""
    from crysplight_amd.config import load_setup
    st = load_setup(RUNS / "crysp_csi_teflon.toml")
    o = st.op
    lab = lam / (1 - f) if f < 1 else math.inf
    lr = lam / f if f > 0 else math.inf
    op = OpticalParams(o.n_crystal, o.n_coupling, float(f32(lab)), float(f32(lr)), o.wrap_R,
                       o.wrap_specular, o.air_gap, o.sigma_alpha, o.pde)
    assert abs(o.wrap_R - 0.99) < 1e-7
    acc = transport_accumulate(st.box, op, readout(st.grid), st.tb, n_photons=st.n_photons,
                               seed=st.seed, pos=st.pos, tau_ns=st.tau_ns, ctx=ctx)
    pct = 100 * acc.ndet / st.n_photons
    print(lam, f, f"{pct:.3f}")
    assert abs(pct - published) <= 0.15, (pct, published)
