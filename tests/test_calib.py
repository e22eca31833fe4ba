"""Calibration-campaign helpers (scripts/{csitl,ding,soleti}_calib.jl), on the CPU."""
import math

import numpy as np
import pytest

from crysplight_amd import calib
from crysplight_amd.config import HERE


def test_fit_R_recovers_a_linear_crossing():
    Rs = [0.80, 0.85, 0.90, 0.95]
    fr = [0.10 + 2.0 * (R - 0.80) for R in Rs]          # frac(R) linear; base 0.05
    # ratio(R) = (0.10 + 2 (R - 0.8)) / 0.05 crosses 4.6 at R = 0.865
    assert calib.fit_R(Rs, fr, 0.05, 4.6) == pytest.approx(0.865, abs=1e-12)
    assert calib.fit_R(Rs, fr, 0.05, 2.0) == pytest.approx(0.80, abs=1e-12)   # at a scan point
    assert math.isnan(calib.fit_R(Rs, fr, 0.05, 1.0))                       # below the floor
    assert math.isnan(calib.fit_R(Rs, fr, 0.05, 9.0))                       # above the ceiling


def test_fit_R_takes_the_first_bracketing_pair():
    Rs = [0.1, 0.2, 0.3, 0.4]
    fr = [1.0, 3.0, 1.0, 3.0]                                              # crosses 2 three times
    assert calib.fit_R(Rs, fr, 1.0, 2.0) == pytest.approx(0.15)


def test_stratified_depths_are_the_truncated_exponential_quantiles():
    att, L, n = 23.0, 20.0, 32
    z = np.array(calib.stratified_depths(att, L, n), np.float64)
    assert z.dtype == np.float64 and np.all(np.diff(z) > 0) and 0 < z[0] and z[-1] < L
    cdf = (1 - np.exp(-z / att)) / (1 - np.exp(-L / att))
    assert np.allclose(cdf, (np.arange(1, n + 1) - 0.5) / n, rtol=0, atol=1e-6)
    assert all(float(np.float32(v)) == v for v in z)                        # Float32 values


@pytest.mark.parametrize("name,fn", [("csitl_calib", calib.csitl_ladder), ("ding_calib", calib.ding_cube),
                                     ("soleti_calib", calib.soleti_crysp)])
def test_run_files_dispatch(name, fn, monkeypatch):
    seen = {}
    monkeypatch.setitem(calib.CAMPAIGNS, name, lambda p, **kw: seen.setdefault("p", p))
    calib.run_campaign(HERE / "runs" / f"{name}.toml")
    assert seen["p"] == HERE / "runs" / f"{name}.toml"


def test_unknown_campaign_is_refused(tmp_path):
    p = tmp_path / "x.toml"
    p.write_text('[campaign]\ntag = "nothing_calib"\n')
    with pytest.raises(ValueError):
        calib.run_campaign(p)


def test_csv_numbers_print_as_julia_does():
    assert calib._num(0.8) == "0.8" and calib._num(math.nan) == "NaN"
    assert calib._num(0.054123) == "0.054123" and calib._num(12) == "12"
