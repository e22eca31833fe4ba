"""Run-file parsing (src/config.jl:146-194, scripts/shoot_gammas.jl:19-47) and the summary
helpers (src/output.jl:58-62), on the CPU."""
import math

import numpy as np
import pytest

from crysplight_amd import config as CF
from crysplight_amd.output import symmetry_residual

RUNS = CF.HERE / "runs"

# a run file in the reference's key layout (runs/point_csitl_teflon.toml's structure)
REF_STYLE = """
[run]
n_photons = 1000
seed = 5
max_bounces = 777

[crystal]
material = "csitl_c1"
size_mm = [48.0, 48.0, 37.2]

[wrap]
material = "teflon_098"

[sipm]
material = "sipm_8x8_3p2"
pde_override = 1.0
coupling_n = 1.45

[surface]
finish = "frontpainted"
sigma_alpha_deg = 6.0

[source]
type = "point"
position_mm = [24.0, 24.0, 18.6]
time_profile = "delta"

[binning]
bin_ns = 100.0
window_ns = 8000.0
"""


def test_surface_has_gap():
    """test/runtests.jl:11-15"""
    assert CF.surface_has_gap("backpainted")
    assert not CF.surface_has_gap("frontpainted")
    with pytest.raises(ValueError):
        CF.surface_has_gap("back-painted")


def test_load_setup_c1_run_file():
    st = CF.load_setup(RUNS / "point_csitl_c1.toml")
    assert st.tag == "point_csitl_c1" and st.n_photons == 1_000_000 and st.seed == 1234
    assert tuple(st.box.L) == pytest.approx((25.6, 25.6, 20.0))
    assert st.op.n_crystal == pytest.approx(1.79) and st.op.abs_len_mm == 400.0
    assert math.isinf(st.op.rayleigh_mm) and st.op.air_gap and not st.op.wrap_specular
    assert st.op.sigma_alpha == pytest.approx(math.radians(1.3), rel=1e-6)
    assert st.op.pde == pytest.approx(0.3) and st.op.n_coupling == 1.5
    assert st.tb.nbins == 80 and st.tau_ns == 1000.0 and st.tau_rise_ns == 20.0
    assert st.max_bounces == 100_000


def test_load_setup_reference_layout(tmp_path):
    (tmp_path / "runs").mkdir()
    p = tmp_path / "runs" / "my_run.toml"
    p.write_text(REF_STYLE)
    st = CF.load_setup(p)
    assert st.tag == "my_run" and st.max_bounces == 777
    assert st.op.pde == 1.0 and not st.op.air_gap            # pde_override, frontpainted
    assert st.tau_ns == 0.0                                  # delta profile
    assert st.outdir == tmp_path / "output" / "my_run"
    with pytest.raises(FileNotFoundError):
        CF.load_material("no_such_material")


def test_gamma_campaign_file():
    c = CF.load_gamma_campaign(RUNS / "gammas_bgo_c2.toml")
    assert c.n_events == 10_000 and c.yield_per_mev == 8500 and c.tau_ns == 300.0
    assert c.op.wrap_specular and c.op.n_crystal == pytest.approx(2.15) and c.density == 7.13


def test_symmetry_residual():
    m = np.ones((8, 8))
    assert symmetry_residual(m) == 0.0
    i, j = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    assert symmetry_residual(((i - 3.5) ** 2 + (j - 3.5) ** 2) + 5) < 1e-12
    m2 = m.copy()
    m2[0, 0] = 9.0          # one corner off: symmetrized value (9 + 3) / 4 ... per class
    assert symmetry_residual(m2) > 0.5
