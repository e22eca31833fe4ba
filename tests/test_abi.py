"""The C-ABI library builds, loads without a GPU and exports every symbol of
include/crysplight.h; the parameter structs have the reference's isbits layout; host-side
logic (reduction of records, gamma stage, Poisson feed) against the reference's anchors."""
import ctypes as C
import math
import re
import subprocess

import numpy as np
import pytest

from crysplight_amd import _lib, events as E, params as P, transport as T
from tests import cases as K


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (crysp_\w+)", out))
    assert set(syms) <= exported


def test_abi_version_and_no_device_error():
    lib = _lib.load()
    assert lib.crysp_abi_version() == 1
    # a context on a machine without a HIP device fails cleanly with a message
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert not lib.crysp_create(0)
        assert lib.crysp_last_error().decode()


def test_struct_layout_matches_julia_isbits():
    """OpticalParams 32 B with bools at 20/21 and sigma at 24; Readout 36 B
    (src/transport.jl:14-50)."""
    for name, size in P.LAYOUT_BYTES.items():
        assert C.sizeof(getattr(P, name)) == size, name
    assert P.OpticalParams.wrap_specular.offset == 20
    assert P.OpticalParams.air_gap.offset == 21
    assert P.OpticalParams.sigma_alpha.offset == 24
    assert P.OpticalParams.pde.offset == 28
    assert P.Readout.disc.offset == 12 and P.Readout.cx.offset == 16
    assert P.Readout.sur_specular.offset == 32 and P.Readout.two_sided.offset == 33
    hdr = (_lib.HEADER).read_text()
    assert "uint8_t pad_[2];" in hdr


def test_readout_constructor_semantics():
    """Readout(grid; center, radius, ...) (src/transport.jl:52-57)."""
    r = P.readout(K.GRID)
    assert not r.disc and r.r2 == 0 and r.sur_R == 0
    r = P.readout(K.GRID, center=(24, 24), radius=10, sur_R=0.9, sur_specular=True)
    assert r.disc and r.cx == 24 and r.r2 == 100 and math.isclose(r.sur_R, 0.9, rel_tol=1e-7)
    assert r.sur_specular and not r.two_sided


def test_host_record_reduction_matches_oracle():
    """transport.reduce_records (host mirror of src/kernel.jl:338-365) equals the
    oracle's reduction on the same records."""
    import oracle
    rec = oracle.transport_records(K.BOX_CRYSP, K.OP_ALL, K.RO_ALL, n_photons=20_000, seed=8,
                                   pos=K.POS_CRYSP, tau_ns=800.0)
    a = T.reduce_records(T.Records(rec["status"], rec["idx"], rec["t"], rec["bounces"]),
                         K.GRID, K.TB)
    b = oracle.reduce_records(rec, K.GRID, K.TB)
    assert np.array_equal(a.counts, b["counts"]) and np.array_equal(a.first_ns, b["first_ns"])
    for n in ("ndet", "ndet_front", "nabs_bulk", "nabs_wall", "nabs_sipm", "ncap", "nabs_sur",
              "sum_bounces_det"):
        assert getattr(a, n) == b[n], n


def test_accumulator_merge():
    """merge! (src/transport.jl:82-91) and total_terminated (:93-94)."""
    a = P.Accumulator.empty(K.GRID, K.TB)
    b = P.Accumulator.empty(K.GRID, K.TB)
    a.counts[1, 2, 3] = 4
    b.counts[1, 2, 3] = 5
    a.first_ns[0, 0] = 7.0
    b.first_ns[0, 0] = 3.0
    a.ndet, b.ndet, b.ncap = 4, 5, 2
    a.merge(b)
    assert a.counts[1, 2, 3] == 9 and a.first_ns[0, 0] == 3.0
    assert a.total_terminated() == 11


# ----------------------------------------------------------- gamma stage ---------
@pytest.fixture(scope="module")
def csi():
    return E.GammaMaterial("xcom_CSI.csv", 4.51)


@pytest.fixture(scope="module")
def bgo():
    return E.GammaMaterial("xcom_BGO.csv", 7.13)


def test_gamma_cross_sections(csi, bgo):
    """test/runtests.jl:258-272"""
    assert math.isclose(csi.mfp(0.511), 2.44, rel_tol=0.02)
    assert bgo.mfp(0.511) < csi.mfp(0.511)
    assert csi.sigma_macro(0.03325)[1] > 2 * csi.sigma_macro(0.03310)[1]
    assert csi.sigma_macro(0.511)[2] == 0.0
    assert csi.sigma_macro(10.0)[2] > 0.0


@pytest.mark.parametrize("name,L_mm", [("csi", 37.2), ("bgo", 22.4)])
def test_gamma_interaction_probability_closed_form(name, L_mm, request):
    """Gammas entering a slab along its normal interact (deposit energy) with probability
    1 - exp(-mu L), mu = (incoherent + photoelectric) at 511 keV: coherent scattering is
    not simulated (src/gamma/Gamma.jl:7).  Within 4 binomial sigma over 40k events."""
    mat = request.getfixturevalue(name)
    inc, pe, pair = mat.sigma_macro(0.511)
    p = 1.0 - math.exp(-(inc + pe + pair) * L_mm / 10.0)
    dep, truth = E.gamma_events(P.Box((48.0, 48.0, L_mm)), mat, 1000.0, n_events=40_000, seed=21)
    got = float((truth["edep"] > 0).mean())
    assert abs(got - p) < 4 * math.sqrt(p * (1 - p) / 40_000), (got, p)


def test_gamma_loglog_interpolation_hits_table(csi):
    """At a grid energy the interpolation returns the tabulated mu/rho * rho; between
    two rows it is the straight line in log-log (src/gamma/nist_data.jl:59-71)."""
    rows = np.array([[float(v) for v in line.split()[:8]]
                     for line in (E.DATA_DIR / "xcom_CSI.csv").read_text().splitlines()
                     if line.strip() and line.strip()[0].isdigit()])
    e, inc = rows[15, 0], rows[15, 2]
    assert math.isclose(csi.sigma_macro(e)[0], inc * 4.51, rel_tol=1e-12)
    e0, e1 = rows[15, 0], rows[16, 0]
    em = math.sqrt(e0 * e1)
    want = math.sqrt(rows[15, 2] * rows[16, 2]) * 4.51
    assert math.isclose(csi.sigma_macro(em)[0], want, rel_tol=1e-9)


def test_gamma_vacuum_and_deposits(csi):
    """test/runtests.jl:274-299 (vacuum: one escape record with the full energy;
    crystal: deposits inside, energy <= 511 keV, times < 1 ns, process tags)."""
    vac = E.GammaMaterial.__new__(E.GammaMaterial)
    rows = np.zeros((2, 8))
    rows[:, 0] = [0.01, 10.0]
    lib = E._bind()
    vac._h = lib.crysp_gamma_material_create(rows.ctypes.data, 2, 0.0)
    L = (48.0, 48.0, 37.2)
    box = P.Box(L)
    recs = vac.history(box, 0.511, (0.0, 0.0, -1.86), (0.0, 0.0, 1.0), seed=1)
    assert recs.shape[0] == 1 and recs[0, 5] == 4 and recs[0, 3] == 0.511
    dep, truth = E.gamma_events(box, csi, 54_000.0, n_events=400, seed=3)
    nin = int((truth["int_type"] >= 0).sum())
    assert nin > 100
    x = dep.xyzt
    assert np.all((x[:, 0] >= 0) & (x[:, 0] <= 48) & (x[:, 1] >= 0) & (x[:, 1] <= 48)
                  & (x[:, 2] >= 0) & (x[:, 2] <= 37.2))
    assert np.all((x[:, 3] >= 0) & (x[:, 3] < 1.0))
    first = truth["ev_first"]
    proc = truth["dep_proc"]
    for ev in range(400):
        a, b = first[ev], first[ev + 1]
        if a == b:
            continue
        assert dep.e_kev[a:b].sum() <= 511.001
        assert np.all(proc[a:b - 1] == E.DEP_COMPTON)
        assert proc[b - 1] in (E.DEP_PHOTO, E.DEP_COMPTON, E.DEP_BELOWCUT)
    # truth consistency (test/runtests.jl:320-332)
    np.testing.assert_allclose(truth["e1"] + truth["e2"] + truth["er"], 511.0, atol=0.01)
    assert np.array_equal(truth["int_type"] == -1, truth["edep"] == 0)
    direct = truth["int_type"] == 0
    assert np.all(truth["n_int"][direct] == 1)
    np.testing.assert_allclose(truth["e1"][direct], 511.0, atol=0.01)


def test_poisson_moments():
    """test/runtests.jl:301-303, plus the Gaussian branch above lambda = 50."""
    p = E.rand_poisson(20.0, 20_000, seed=1)
    assert math.isclose(p.mean(), 20, rel_tol=0.02) and math.isclose(p.var(), 20, rel_tol=0.05)
    p = E.rand_poisson(4000.0, 20_000, seed=2)
    assert math.isclose(p.mean(), 4000, rel_tol=0.002) and math.isclose(p.var(), 4000, rel_tol=0.05)
    assert np.all(E.rand_poisson(0.0, 10) == 0)


def test_synthetic_deposits_ranges():
    d = E.synthetic_deposits(K.C1_BOX, n_events=5000, seed=3, e_kev=(10, 511),
                             yield_per_mev=54_000)
    assert d.xyzt[:, :3].min() >= 0 and np.all(d.xyzt[:, :3].max(0) <= np.array(K.C1_BOX.L))
    assert 10 <= d.e_kev.min() and d.e_kev.max() <= 511
    lam = 54.0 * d.e_kev
    z = (d.nph - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05


class _NoCtx:
    """Stands in for a context: the argument checks run before any library call."""
    device = 0
    handle = None


def test_event_wrappers_reject_bad_arrays_before_calling():
    dep = E.Deposits(np.zeros((4, 4), np.float32), np.ones(4, np.int32), np.arange(4, dtype=np.int32),
                     np.ones(4, np.float32), 4)
    ro = P.readout(P.SipmGrid(3.2, 8, 8))
    good_in = (dep.xyzt, dep.nph, dep.event)
    good_out = (np.zeros((4, 8, 8), np.int32), np.zeros((4, 8, 8), np.float32), np.zeros(4, np.int32))
    bad_cases = [
        ((dep.xyzt.astype(np.float64), dep.nph, dep.event), good_out, "xyzt"),
        ((dep.xyzt, dep.nph.astype(np.int64), dep.event), good_out, "nph"),
        ((dep.xyzt, dep.nph, dep.event[:3]), good_out, "event"),
        ((np.zeros((4, 8), np.float32)[:, ::2], dep.nph, dep.event), good_out, "xyzt"),
        (good_in, (np.zeros((3, 8, 8), np.int32), good_out[1], good_out[2]), "maps"),
        (good_in, (good_out[0], np.zeros((4, 8, 8), np.float64), good_out[2]), "tmin"),
        (good_in, (good_out[0], good_out[1], np.zeros(5, np.int32)), "npe"),
    ]
    for inp, out, what in bad_cases:
        with pytest.raises(ValueError, match=what):
            E.transport_events(K.C2_BOX, K.C2_OP, ro, dep, seed=1, tau_ns=300.0, ctx=_NoCtx(),
                               device_arrays=inp, out=out)
    with pytest.raises(ValueError, match="dep_box"):
        E.transport_events_multibox([K.C2_BOX], None, K.C2_OP, ro, dep, seed=1, tau_ns=300.0,
                                    ctx=_NoCtx(), device_arrays=good_in + (np.zeros(4, np.int16),),
                                    out=good_out)


def test_check_array_torch_rules():
    torch = pytest.importorskip("torch")
    _lib.check_array(torch.zeros((2, 4)), "x", np.float32, (2, 4))
    with pytest.raises(ValueError, match="contiguous"):
        _lib.check_array(torch.zeros((4, 4))[:, :2], "x", np.float32)
    with pytest.raises(ValueError, match="dtype"):
        _lib.check_array(torch.zeros(3, dtype=torch.int64), "n", np.int32)


def test_every_pointer_taking_entry_point_has_a_signature():
    """ctypes passes an undeclared pointer argument as a 32-bit int: every entry point the
    Python host calls with a context handle is declared in _lib.SIGNATURES."""
    import pathlib
    called = set()
    for f in pathlib.Path(_lib.HERE).glob("*.py"):
        called |= set(re.findall(r"load\(\)\.(crysp_\w+)\(", f.read_text()))
    missing = sorted(n for n in called if n not in _lib.SIGNATURES)
    assert not missing, missing


def test_math_tables_are_the_generator_output_and_correctly_rounded():
    """The device's transcendental tables are what their generator writes.  sin, cos of
    k pi/128 (k = 0..256): every entry the binary64 rounding of the value computed in x87
    extended precision (within half an ulp plus the extended precision's own error), the
    multiples of pi/2 exact.  log: for each of the 129 intervals of m in [sqrt(2)/2,
    sqrt(2)) (low exponent bit and 7 leading mantissa bits 0x35..0xB5), 1/c is the rounded
    inverse of the interval's midpoint (of 1 next to 1), so that |m (1/c) - 1| <= 2^-7 over
    the interval, and the second column is -log(1/c) correctly rounded."""
    import importlib.util
    from pathlib import Path
    csrc = Path(__file__).resolve().parents[1] / "crysplight_amd" / "csrc"
    spec = importlib.util.spec_from_file_location("gen_math_tables", csrc / "gen_math_tables.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert (csrc / "math_tables.h").read_text() == gen.header()
    rows = np.array(gen.sincos_table())
    assert rows.shape == (257, 2)
    k = np.arange(257)
    assert np.array_equal(rows[k % 64 == 0], [[0, 1], [1, 0], [0, -1], [-1, 0], [0, 1]])
    pi_l = np.longdouble(4) * np.arctan(np.longdouble(1))
    a = k.astype(np.longdouble) * pi_l / 128
    for col, f in ((0, np.sin), (1, np.cos)):
        want = f(a)
        ok = k % 64 != 0
        err = np.abs(rows[ok, col].astype(np.longdouble) - want[ok]) / np.spacing(np.abs(rows[ok, col]))
        assert err.max() < 0.5 + 1e-3, (col, float(err.max()))
    lt = np.array(gen.log_table())
    assert lt.shape == (129, 2)
    idx = np.arange(0x35, 0xB6)
    lo = np.where(idx < 128, (128 + (idx & 127)) / 256.0, (128 + (idx & 127)) / 128.0)
    hi = lo + np.where(idx < 128, 1 / 256.0, 1 / 128.0)
    lo = np.maximum(lo, np.sqrt(0.5))
    hi = np.minimum(hi, np.sqrt(2.0))
    one = (idx == 0x7F) | (idx == 0x80)
    assert np.array_equal(lt[one], [[1.0, 0.0], [1.0, 0.0]])
    r = np.maximum(np.abs(lo * lt[:, 0] - 1), np.abs(hi * lt[:, 0] - 1))
    assert r.max() <= 2.0 ** -7 * (1 + 1e-12), r.max()
    want = -np.log(lt[~one, 0].astype(np.longdouble))
    err = np.abs(lt[~one, 1].astype(np.longdouble) - want) / np.spacing(np.abs(lt[~one, 1]))
    assert err.max() < 0.5 + 1e-3, float(err.max())
