"""Pins of the CPU oracle (oracle/crysp_oracle.c), before it is trusted as the checker.

Each test restates an assertion of the reference's test/runtests.jl (cited per test),
a published known answer (Philox4x32-10 KAT vectors), an independent implementation
(rocRAND's header-only Philox engine) or a closed form.  Photon counts are reduced
where the assertion is size-independent, so the CPU suite stays short; the GPU suite
repeats the transport anchors at the reference's sizes on the HIP kernels."""
import math

import numpy as np
import pytest

import oracle
from crysplight_amd.params import (Box, OpticalParams, SipmGrid, TimeBinning, readout,
                                   deg2rad, STATUS_DETECTED, STATUS_ABS_SUR, FACE_XP,
                                   FACE_XM, FACE_YM, FACE_BACK)
from tests import cases as K


# ---------------------------------------------------------------- Philox --------
def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle.philox4x32_10(ctr, key)) == want


def test_philox_stream_matches_rocrand():
    """rocRAND's engine(seed, subsequence=block, offset=4*pid) starts at counter
    (pid lo, pid hi, block, 0): its first four words are block `block` of the stream
    PhiloxStream(seed, pid) (src/philox.jl:43-57)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        seed = int(rng.integers(0, 2**63))
        pid = int(rng.integers(1, 2**40))
        nblk = 6
        draws = oracle.debug_draws(seed, 1, 4 * nblk, pid0=pid - 1)[0]
        for b in range(nblk):
            assert list(draws[4 * b:4 * b + 4]) == list(oracle.rocrand_block(seed, pid, b))


def test_randu_closed_form_exhaustive():
    """randu = ((x >>> 8) + 1) * 2^-24 (src/philox.jl:62) for all 2^24 classes of x."""
    k = np.arange(2**24, dtype=np.uint32)
    low = np.random.default_rng(0).integers(0, 256, size=k.size, dtype=np.uint32)
    u = oracle.u32_to_unit((k << np.uint32(8)) | low)
    assert np.array_equal(u.astype(np.float64), (k.astype(np.float64) + 1) * 2.0**-24)
    assert u.min() > 0 and u.max() == 1.0


def test_philox_rng_properties():
    """test/runtests.jl:23-33"""
    a = oracle.u32_to_unit(oracle.debug_draws(42, 1, 100, pid0=6)[0])
    b = oracle.u32_to_unit(oracle.debug_draws(42, 1, 100, pid0=6)[0])
    assert np.array_equal(a, b)
    s3 = oracle.u32_to_unit(oracle.debug_draws(42, 1, 1, pid0=7))[0, 0]
    s4 = oracle.u32_to_unit(oracle.debug_draws(43, 1, 1, pid0=6))[0, 0]
    assert s3 != s4
    u = oracle.u32_to_unit(oracle.debug_draws(1, 200_000, 1)[:, 0]).astype(np.float64)
    assert np.all((u > 0) & (u <= 1))
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.std(ddof=1) - math.sqrt(1 / 12)) < 3e-3


@pytest.mark.parametrize("fn,domain", [("log", "uniform"), ("sin", "phase"), ("cos", "phase"),
                                       ("sin", "tilt"), ("cos", "tilt")])
def test_oracle_transcendentals_correctly_rounded(fn, domain):
    """The oracle's (float)f((double)x) equals the binary32 rounding of the x87
    extended-precision value on every log / phase argument randu can produce and on a
    sample of the tilt domain; Julia's Float32 log/sin/cos are correctly rounded there
    too, so this is where per-photon parity with the reference is decided."""
    assert np.finfo(np.longdouble).nmant >= 63
    x = {"uniform": oracle.uniform_domain, "phase": oracle.phase_domain,
         "tilt": lambda: oracle.tilt_domain(61)}[domain]()
    assert np.array_equal(oracle.math(fn, x), oracle.correctly_rounded(fn, x))


# --------------------------------------------------------------- geometry -------
def test_wall_hit_slab_distances():
    """test/runtests.jl:35-46"""
    box = K.BOX_CRYSP
    d, f = oracle.wall_hit((24, 24, 18.6, 1, 0, 0), box)
    assert math.isclose(d, 24, rel_tol=3.5e-4) and f == FACE_XP
    d, f = oracle.wall_hit((24, 24, 18.6, 0, -1, 0), box)
    assert math.isclose(d, 24, rel_tol=3.5e-4) and f == FACE_YM
    d, f = oracle.wall_hit((24, 24, 18.6, 0, 0, 1), box)
    assert math.isclose(d, np.float32(37.2) - np.float32(18.6), rel_tol=3.5e-4) and f == FACE_BACK
    u = float(np.float32(1) / np.sqrt(np.float32(2)))
    d, f = oracle.wall_hit((0, 0, 0, u, 0, u), box)
    assert f == FACE_BACK and math.isclose(d, 37.2 * math.sqrt(2), rel_tol=3.5e-4)


def test_wall_hit_tie_order():
    """Ties go x, then y, then z (src/geometry.jl:26-31)."""
    box = Box((10.0, 10.0, 10.0))
    u = float(np.float32(1 / math.sqrt(3)))
    assert oracle.wall_hit((5, 5, 5, u, u, u), box)[1] == FACE_XP
    assert oracle.wall_hit((5, 5, 5, 0, -u, -u), box)[1] == 4      # FACE_YM before FRONT


# ---------------------------------------------------------------- Fresnel -------
def test_fresnel_closed_forms():
    """test/runtests.jl:48-63"""
    tir, R = oracle.fresnel(1.95, 1.45, 1.0)
    assert not tir and math.isclose(R, ((1.95 - 1.45) / (1.95 + 1.45)) ** 2, rel_tol=1e-5)
    assert oracle.fresnel(1.95, 1.0, float(np.cos(np.float32(np.deg2rad(40)))))[0]
    assert not oracle.fresnel(1.95, 1.0, float(np.cos(np.float32(np.deg2rad(20)))))[0]
    assert oracle.fresnel(1.95, 1.45, 1e-4)[1] > 0.999
    _, Ra = oracle.fresnel(1.45, 1.95, 1.0)
    assert math.isclose(Ra, ((1.95 - 1.45) / (1.95 + 1.45)) ** 2, rel_tol=1e-5)


def test_fresnel_against_float64_formula():
    """Unpolarized Fresnel in float64 at random angles: the float32 result is within a
    few ulp, and the TIR boundary sits at the critical angle."""
    rng = np.random.default_rng(3)
    for n1, n2 in [(1.79, 1.5), (2.15, 1.0), (1.5, 1.79)]:
        for c in rng.uniform(0.01, 1.0, 200):
            c = float(np.float32(c))
            tir, R = oracle.fresnel(n1, n2, c)
            s2 = (n1 / n2) ** 2 * (1 - c * c)
            if s2 >= 1.0 + 1e-5:
                assert tir
                continue
            if s2 >= 1.0 - 1e-5:
                continue
            ct = math.sqrt(1 - s2)
            rs = ((n1 * c - n2 * ct) / (n1 * c + n2 * ct)) ** 2
            rp = ((n1 * ct - n2 * c) / (n1 * ct + n2 * c)) ** 2
            assert not tir and abs(R - 0.5 * (rs + rp)) < 1e-5


# ------------------------------------------------------------- samplers ---------
def test_sampler_moments():
    """test/runtests.jl:65-100"""
    d = oracle.sample("isotropic", 7, 1, 200_000).astype(np.float64)
    assert abs(d[:, 2].mean()) < 5e-3
    assert abs((d[:, 2] ** 2).mean() - 1 / 3) < 3e-3
    assert np.all(np.abs((d[:1000, :3] ** 2).sum(1) - 1) < 1e-5)
    lam = oracle.sample("lambertian", 7, 2, 200_000, face=FACE_BACK).astype(np.float64)
    assert np.all(lam[:, 2] < 0)
    assert abs((-lam[:, 2]).mean() - 2 / 3) < 3e-3
    lam = oracle.sample("lambertian", 7, 3, 1000, face=FACE_XM)
    assert np.all(lam[:, 0] > 0)
    dt = oracle.sample("emission", 7, 4, 200_000, args=(800.0, 0.0))[:, 0].astype(np.float64)
    assert abs(dt.mean() - 800) < 8.0
    assert np.all(oracle.sample("emission", 7, 5, 10, args=(0.0, 0.0))[:, 0] == 0)
    mus = oracle.sample("rayleigh", 7, 6, 200_000, args=(0, 0, 1))[:, 2].astype(np.float64)
    assert abs(mus.mean()) < 6e-3
    assert abs((mus ** 2).mean() - 0.4) < 5e-3
    v = oracle.sample("rayleigh", 7, 7, 1000, args=(0.6, 0.64, 0.48)).astype(np.float64)
    assert np.all(np.abs((v[:, :3] ** 2).sum(1) - 1) < 1e-5)


def test_rise_time_moments():
    """Rise-time extension: Exp(td) + Exp(tr) has mean td + tr, variance td^2 + tr^2."""
    td, tr = 1000.0, 20.0
    t = oracle.sample("emission", 11, 1, 400_000, args=(td, tr))[:, 0].astype(np.float64)
    assert abs(t.mean() - (td + tr)) < 4 * math.sqrt(td * td + tr * tr) / math.sqrt(t.size)
    assert abs(t.var() / (td * td + tr * tr) - 1) < 0.02
    # small-t shape: the pdf vanishes at t = 0 (bi-exponential), unlike a pure Exp(td)
    frac_small = np.mean(t < 5.0)
    expect = 1 - (td * math.exp(-5 / td) - tr * math.exp(-5 / tr)) / (td - tr)
    assert abs(frac_small - expect) < 5 * math.sqrt(expect / t.size)


def test_unified_surface_limits():
    """test/runtests.jl:90-99"""
    u = (0.5, 0.5, float(np.sqrt(np.float32(0.5))))
    r = oracle.sample("surface", 7, 8, 1, face=FACE_XP, args=(1.95, 1e-6) + u)[0]
    assert r[3] == 0 and abs(r[0] + u[0]) < 1e-3 and abs(r[1] - u[1]) < 1e-3
    res = oracle.sample("surface", 7, 9, 2000, face=FACE_XP, args=(1.95, deg2rad(12)) + u)
    assert np.all((res[:, 3] == 1) | (res[:, 0] < 0))
    assert np.any(res[:, 3] == 1)


# ------------------------------------------------------------ transport ---------
def _run(box, op, ro, n, seed, pos, tau=0.0, maxb=100_000, tb=K.TB):
    return oracle.run_photons(box, op, ro, tb, n_photons=n, seed=seed, pos=pos, tau_ns=tau,
                              max_bounces=maxb)


def _closed(a, n):
    return a["ndet"] + a["nabs_bulk"] + a["nabs_wall"] + a["nabs_sipm"] + a["ncap"] + a["nabs_sur"] == n


def test_conservation_physical():
    """test/runtests.jl:106-111"""
    op = K.ideal(1.95, wrap_R=0.99, abs_len=500.0, pde=0.4)
    a = _run(K.BOX_CRYSP, op, K.GRID, 20_000, 1, K.POS_CRYSP, tau=800.0)
    assert _closed(a, 20_000) and a["counts"].sum() == a["ndet"]


def test_idealized_teflon_escape_cone():
    """test/runtests.jl:113-124: nothing absorbed, >= the direct grease escape cone."""
    nl = 4000
    a = _run(K.BOX_CRYSP, K.ideal(1.95), K.GRID, nl, 2, K.POS_CRYSP, maxb=20_000)
    assert a["ndet"] + a["ncap"] == nl
    assert a["nabs_bulk"] == a["nabs_wall"] == a["nabs_sipm"] == 0
    p = 1 - math.cos(math.asin(1.45 / 1.95))
    assert a["ndet"] / nl > p - 4 * math.sqrt(p * (1 - p) / nl)
    assert a["ndet"] < nl


def test_specular_box_exact_escape_cone():
    """test/runtests.jl:126-135: a specular box detects exactly 1 - cos(asin(1.45/2.15))."""
    n = 20_000
    a = _run(Box((48.0, 48.0, 22.4)), K.ideal(2.15, spec=True), K.GRID, n, 3,
             (24.0, 24.0, 11.2), maxb=2_000)
    p = 1 - math.cos(math.asin(1.45 / 2.15))
    assert abs(a["ndet"] / n - p) < 4 * math.sqrt(p * (1 - p) / n)
    assert a["ncap"] + a["ndet"] == n


def test_black_walls_close():
    """test/runtests.jl:137-143"""
    a = _run(K.BOX_CRYSP, K.ideal(1.95, wrap_R=0.0, abs_len=100.0), K.GRID, 20_000, 4,
             K.POS_CRYSP)
    assert _closed(a, 20_000) and a["nabs_bulk"] > 0 and a["nabs_wall"] > 0


def test_rayleigh_ergodic():
    """test/runtests.jl:145-150"""
    a = _run(K.BOX_CRYSP, K.ideal(1.95, ray=100.0), K.GRID, 3_000, 5, K.POS_CRYSP)
    assert a["ndet"] == 3_000 and a["nscat"] > 0


def test_roughness_raises_collection():
    """test/runtests.jl:152-159"""
    n0 = _run(K.BOX_CRYSP, K.ideal(1.95), K.GRID, 3_000, 6, K.POS_CRYSP, maxb=5_000)["ndet"]
    n1 = _run(K.BOX_CRYSP, K.ideal(1.95, sigma=deg2rad(6)), K.GRID, 3_000, 6, K.POS_CRYSP,
              maxb=5_000)["ndet"]
    assert n1 > n0


def test_contact_lambertian_detects_all():
    """test/runtests.jl:161-165"""
    a = _run(K.BOX_CRYSP, K.ideal(1.95, gap=False), K.GRID, 5_000, 7, K.POS_CRYSP)
    assert a["ndet"] == 5_000


def test_all_processes_conserve():
    """test/runtests.jl:167-173"""
    a = _run(K.BOX_CRYSP, OpticalParams(1.95, 1.45, 500.0, 400.0, 0.98, False, True,
                                        deg2rad(6), 0.4), K.GRID, 20_000, 8, K.POS_CRYSP,
             tau=800.0)
    assert _closed(a, 20_000) and a["counts"].sum() == a["ndet"]


# ---------------------------------------------------------------- readout -------
def test_readout_disc_and_surround():
    """test/runtests.jl:176-197"""
    ro = readout(K.GRID, center=(24, 24), radius=10, sur_R=0)
    op = K.ideal(1.95)
    st = oracle.propagate_one(K.BOX_CRYSP, op, ro, 21, 1, (24, 24, 0, 0, 0, 1, 0))[0]
    assert st == STATUS_DETECTED
    st = oracle.propagate_one(K.BOX_CRYSP, op, ro, 21, 2, (4, 4, 0, 0, 0, 1, 0))[0]
    assert st == STATUS_ABS_SUR
    ro = readout(K.GRID, center=(24, 24), radius=10, sur_R=0.95)
    opb = K.ideal(1.79, wrap_R=0.9, abs_len=300.0)
    a = _run(K.BOX_CRYSP, opb, ro, 20_000, 22, (24.0, 24.0, 18.0))
    assert _closed(a, 20_000) and a["nabs_sur"] > 0 and a["ndet"] > 0


def test_multi_position_deterministic():
    """test/runtests.jl:199-206"""
    ro = readout(K.GRID, center=(24, 24), radius=10, sur_R=0.95)
    opb = K.ideal(1.79, wrap_R=0.9, abs_len=300.0)
    ps = [(24.0, 24.0, 5.0), (24.0, 24.0, 30.0)]
    kw = dict(n_photons=10_000, seed=23, positions=ps, tau_ns=0.0)
    a1 = oracle.run_photons(K.BOX_CRYSP, opb, ro, K.TB, **kw)
    a2 = oracle.run_photons(K.BOX_CRYSP, opb, ro, K.TB, threads=1, **kw)
    assert _closed(a1, 10_000)
    assert a1["ndet"] == a2["ndet"] and np.array_equal(a1["counts"], a2["counts"])


def test_two_sided_readout():
    """test/runtests.jl:208-221"""
    ro2 = readout(K.GRID, two_sided=True)
    op = K.ideal(1.95)
    up = oracle.propagate_one(K.BOX_CRYSP, op, ro2, 31, 1, (24, 24, 18, 0, 0, 1, 0))
    dn = oracle.propagate_one(K.BOX_CRYSP, op, ro2, 31, 2, (24, 24, 18, 0, 0, -1, 0))
    assert up[0] == STATUS_DETECTED and up[1] > 0
    assert dn[0] == STATUS_DETECTED and dn[1] < 0
    opb = K.ideal(1.79, wrap_R=0.9, abs_len=300.0)
    a = _run(K.BOX_CRYSP, opb, ro2, 20_000, 32, K.POS_CRYSP)
    assert _closed(a, 20_000)
    assert abs(a["ndet_front"] / a["ndet"] - 0.5) < 0.02


def test_detection_record():
    """test/runtests.jl:341-357: straight down +z at (9, 39) lands in SiPM (2, 7),
    bin 1, at t = 37.2 * 1.95 / c; emission at 1e6 ns lands in the overflow bin."""
    op = K.ideal(1.95)
    st, idx, t, b = oracle.propagate_one(K.BOX_CRYSP, op, K.GRID, 11, 1, (9, 39, 0, 0, 0, 1, 0))
    assert st == STATUS_DETECTED and idx == (7 - 1) * 8 + 2
    texp = np.float32(37.2) * np.float32(1.95) / np.float32(299.792458)
    assert math.isclose(t, texp, rel_tol=1e-4)
    rec = {"status": np.array([st], np.uint8), "idx": np.array([idx], np.int16),
           "t": np.array([t], np.float32), "bounces": np.array([b], np.int32)}
    acc = oracle.reduce_records(rec, K.GRID, K.TB)
    assert acc["counts"][1, 6, 0] == 1
    st2, idx2, t2, _ = oracle.propagate_one(K.BOX_CRYSP, op, K.GRID, 11, 2,
                                            (9, 39, 0, 0, 0, 1, 1e6))
    rec2 = {"status": np.array([st2], np.uint8), "idx": np.array([idx2], np.int16),
            "t": np.array([t2], np.float32), "bounces": np.array([0], np.int32)}
    assert oracle.reduce_records(rec2, K.GRID, K.TB)["counts"][1, 6, -1] == 1


def test_layout_matches_julia_isbits():
    """Byte sizes of the isbits structs (OpticalParams bools at 20/21, Readout 36 B)."""
    s = oracle.layout_sizes()
    assert (s["Box"], s["OpticalParams"], s["SipmGrid"], s["Readout"], s["TimeBinning"]) == \
        (12, 32, 12, 36, 8)


def test_event_reduction_matches_records():
    """src/events.jl:73-87: per-event maps are the detections of each event's contiguous
    photon range, both faces counted, tmin the earliest; npe their sum."""
    rng = np.random.default_rng(1)
    ndep = 40
    dep = np.column_stack([rng.uniform(1, 47, ndep), rng.uniform(1, 47, ndep),
                           rng.uniform(1, 36, ndep), rng.uniform(0, 1, ndep)]).astype(np.float32)
    nph = rng.integers(0, 300, ndep).astype(np.int32)
    ev = np.sort(rng.integers(0, 15, ndep)).astype(np.int32)
    ro2 = readout(K.GRID, two_sided=True)
    r = oracle.run_events(K.BOX_CRYSP, K.OP_ALL, ro2, dep, nph, ev, 15, seed=3, pid0=17,
                          tau_ns=800.0, return_records=True, count_ops=True)
    # recompute with the records mode of the same photons, one deposit at a time
    g0 = 0
    maps = np.zeros((15, 64), np.int64)
    ops = dict.fromkeys(oracle.OPCOUNT_NAMES, 0)
    for d in range(ndep):
        rec = oracle.transport_records(K.BOX_CRYSP, K.OP_ALL, ro2, n_photons=int(nph[d]),
                                       seed=3, pos=tuple(dep[d, :3]), t0s=[dep[d, 3]],
                                       pid_offset=17 + g0, tau_ns=800.0, count_ops=True)
        det = rec["status"] == STATUS_DETECTED
        np.add.at(maps[ev[d]], np.abs(rec["idx"][det].astype(np.int64)) - 1, 1)
        assert np.array_equal(rec["status"], r["status"][g0:g0 + nph[d]])
        for k in ops:
            ops[k] += rec["ops"][k]
        g0 += int(nph[d])
    assert np.array_equal(maps, r["maps"])
    # the event mode's operation counters cover exactly the same photons
    assert r["ops"] == ops
    assert r["ops"]["n_photon"] == int(nph.sum())
    assert r["ops"]["n_detect"] == int(r["npe"].sum())
    assert np.array_equal(r["npe"], maps.sum(1))
    assert np.array_equal(np.isfinite(r["tmin"]), r["maps"] > 0)
