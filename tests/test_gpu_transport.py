"""GPU parity of the HIP transport against the CPU oracle, through the C-ABI.

Bar: integer outputs bit-exact; per-photon fates identical on >= 99.9% of photons
(in practice every photon: the transcendentals are correctly rounded on both sides)."""
import numpy as np
import pytest

import oracle
from crysplight_amd import transport as T
from crysplight_amd.params import Box, SipmGrid, TimeBinning, readout, deg2rad
from tests import cases as K

pytestmark = pytest.mark.gpu


def _fate_match(gpu, ref):
    same = ((gpu.status == ref["status"]) & (gpu.idx == ref["idx"])
            & (gpu.bounces == ref["bounces"]))
    return same.mean(), same


def test_philox_streams_bit_exact(ctx):
    for seed, pid0 in [(1234, 0), (42, 6), (2**40 + 7, 2**32 + 3)]:
        g = T.debug_draws(seed, 300, 64, pid0=pid0, ctx=ctx)
        r = oracle.debug_draws(seed, 300, 64, pid0=pid0)
        assert np.array_equal(g, r), (seed, pid0)


def test_log_exhaustive_on_uniform_domain(ctx):
    """log is only ever evaluated at randu outputs k * 2^-24, k = 1..2^24: all of them,
    against the correctly rounded value and against the oracle's log."""
    u = oracle.uniform_domain()
    g = T.debug_math("log", u, ctx=ctx)
    bad = np.flatnonzero(g != oracle.correctly_rounded("log", u))
    assert bad.size == 0, (bad[:10], u[bad[:10]])
    assert np.array_equal(g, oracle.math("log", u))


def test_absorption_estimate_within_band_on_uniform_domain(ctx):
    """The bulk-absorption decision d_abs < min(d, d_ray) is taken on the estimate
    est = f32(-abs_len ln2) * log2_est(u) unless est lies within a relative band of 2^-12
    of min(d, d_ray); inside it the exact d_abs = -abs_len * log(u) decides.  The outcome is
    the exact one's when |est - d_abs| stays far inside the band: checked here on every
    draw value k * 2^-24 for several absorption lengths, with the device's own log2
    estimate and log, and log2_est(1) = 0 exactly."""
    u = oracle.uniform_domain()
    h = T.debug_math("log2_est", u, ctx=ctx)
    lg = T.debug_math("log", u, ctx=ctx)
    assert u[-1] == 1.0 and h[-1] == 0.0
    ex = np.log2(u.astype(np.float64))
    rel_h = np.abs(h[:-1].astype(np.float64) - ex[:-1]) / np.abs(ex[:-1])
    assert rel_h.max() < 2.0 ** -21, rel_h.max()
    for abs_len in (300.0, 400.0, 500.0, 0.37, 1.0e6):
        k = np.float32(-np.float32(abs_len) * np.float32(0.6931472))
        est = (k * h).astype(np.float64)
        d_abs = (np.float32(-abs_len) * lg).astype(np.float64)
        rel = np.abs(est[:-1] - d_abs[:-1]) / d_abs[:-1]
        assert rel.max() < 2.0 ** -19, (abs_len, rel.max())
        assert est[-1] == 0.0 and d_abs[-1] == 0.0


def test_fresnel_estimate_gap(ctx):
    """The Fresnel reflection decision u < R is taken on an estimate of R (hardware sqrt
    and reciprocals) unless u lies within 2^-14 of it, where the exact R decides.  The
    largest gap between the two, over the crystal/coupling and crystal/air interfaces of
    the workloads, equal indices and n1 < n2, stays below 2^-19 on every 4th float of
    [2^-20, 1] and every float of [1/2, 1]."""
    b = lambda v: int(np.float32(v).view(np.uint32))
    c = np.concatenate([np.arange(b(2.0 ** -20), b(0.5), 4, dtype=np.uint32),
                        np.arange(b(0.5), b(1.0) + 1, dtype=np.uint32)]).view(np.float32)
    g = T.debug_math("fresnel_gap", c, ctx=ctx)
    assert g.max() < 2.0 ** -19, (g.max(), c[g.argmax()])


@pytest.mark.parametrize("fn", ["sin", "cos", "cos1"])
def test_sincos_exhaustive_on_phase_domain(ctx, fn):
    """sin/cos of phi = 2f0*Float32(pi)*u for every randu value u."""
    phi = oracle.phase_domain()
    g = T.debug_math(fn, phi, ctx=ctx)
    bad = np.flatnonzero(g != oracle.correctly_rounded(fn[:3], phi))
    assert bad.size == 0, (bad[:10], phi[bad[:10]])
    assert np.array_equal(g, oracle.math(fn[:3], phi))


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sincos_on_tilt_domain(ctx, fn):
    """alpha = sigma*|g| < 1.2: every 7th binary32 value in [0, 1.2)."""
    x = oracle.tilt_domain(7)
    g = T.debug_math(fn, x, ctx=ctx)
    assert np.count_nonzero(g != oracle.correctly_rounded(fn, x)) == 0


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sincos_tilt_path(ctx, fn):
    """The facet-tilt sin/cos: every 7th binary32 value in [0, 1.2) (series below 1/8,
    the general routine above), and every value in [1/16, 1/8], where the series'
    truncation is largest, plus the switch point."""
    lo, hi = np.float32(0.0625).view(np.uint32), np.float32(0.125).view(np.uint32)
    x = np.concatenate([oracle.tilt_domain(7),
                        np.arange(lo, hi + 2, dtype=np.uint32).view(np.float32)])
    g = T.debug_math(fn + "_tilt", x, ctx=ctx)
    want = oracle.correctly_rounded(fn, x)
    bad = np.flatnonzero(g != want)
    assert bad.size == 0, (x[bad[:10]], g[bad[:10]], want[bad[:10]])


def test_sqrt_correctly_rounded(ctx):
    """sqrt is evaluated at 0 and on [2^-25, 34) (the Box-Muller radius reaches
    sqrt(-2 log 2^-24) = sqrt(33.3)): every float of [0.25, 4) (two binades: the
    correction does not depend on the exponent), every 5th float from 2^-25 to 34, zero."""
    as_f = lambda a: a.astype(np.uint32).view(np.float32)
    b = lambda v: int(np.float32(v).view(np.uint32))
    x = np.concatenate([as_f(np.arange(b(0.25), b(4.0))),
                        as_f(np.arange(b(2.0 ** -25), b(34.0), 5)),
                        np.zeros(1, np.float32)])
    g = T.debug_math("sqrt", x, ctx=ctx)
    assert np.array_equal(g, np.sqrt(x))


def _records_case(ctx, box, op, ro, n, seed, pos, tau, maxb=100_000, positions=None,
                  t0s=None, pid0=0, tau_rise=0.0):
    kw = dict(n_photons=n, seed=seed, pos=pos, positions=positions, t0s=t0s,
              pid_offset=pid0, tau_ns=tau, tau_rise_ns=tau_rise, max_bounces=maxb)
    g = T.transport_records(box, op, ro, ctx=ctx, **kw)
    r = oracle.transport_records(box, op, ro, **kw)
    frac, same = _fate_match(g, r)
    tdiff = np.flatnonzero(same & (g.t != r["t"]))
    print(f"fate parity {frac:.6f} ({(~same).sum()} differ), time mismatches {tdiff.size}")
    assert frac >= 0.999, frac
    assert tdiff.size == 0, tdiff[:10]
    return g, r, frac


def test_records_c1_point_source(ctx):
    g, r, frac = _records_case(ctx, K.C1_BOX, K.C1_OP, K.C1_RO, 100_000, K.C1_SEED,
                               K.C1_POS, K.C1_TAU)
    assert frac == 1.0


def test_records_all_processes(ctx):
    g, r, frac = _records_case(ctx, K.BOX_CRYSP, K.OP_ALL, K.RO_ALL, 100_000, 8,
                               K.POS_CRYSP, 800.0)
    assert frac == 1.0


def test_records_two_sided_rise_multi_position(ctx):
    ro2 = readout(K.GRID, two_sided=True)
    pos = [(24.0, 24.0, 5.0), (24.0, 24.0, 30.0), (3.0, 40.0, 20.0)]
    g, r, frac = _records_case(ctx, K.BOX_CRYSP, K.OP_ALL, ro2, 60_000, 23, None, 800.0,
                               positions=pos, t0s=[0.0, 5.0, 1e6], pid0=1_000_003,
                               tau_rise=20.0)
    assert frac == 1.0
    assert (g.idx < 0).any() and (g.idx > 0).any()


def test_records_photon_ids_cross_2_32_with_64bit_seed(ctx):
    """Photon ids pid_offset + k run across 2^32 (the Philox counter's high word goes from 0
    to 1 mid-launch, src/philox.jl:34-40), under a seed with key bits above 32."""
    g, r, frac = _records_case(ctx, K.C2_BOX, K.C2_OP, K.C1_RO, 20_000, 2 ** 40 + 7,
                               (12.8, 12.8, 10.0), K.C2_TAU, pid0=2 ** 32 - 10_000)
    assert frac == 1.0


def test_accumulate_equals_reduced_records(ctx):
    kw = dict(n_photons=200_000, seed=8, pos=K.POS_CRYSP, tau_ns=800.0)
    acc = T.transport_accumulate(K.BOX_CRYSP, K.OP_ALL, K.RO_ALL, K.TB, ctx=ctx, **kw)
    r = oracle.run_photons(K.BOX_CRYSP, K.OP_ALL, K.RO_ALL, K.TB, **kw)
    assert np.array_equal(acc.counts, r["counts"])
    assert np.array_equal(acc.first_ns, r["first_ns"])
    for name in ("ndet", "ndet_front", "nabs_bulk", "nabs_wall", "nabs_sipm", "ncap",
                 "nabs_sur", "nscat", "sum_bounces_det"):
        assert getattr(acc, name) == r[name], name
    assert acc.total_terminated() == 200_000


@pytest.mark.parametrize("spec,ray", [(True, float("inf")), (False, float("inf")), (False, 333.3)])
def test_accumulate_specialized_optics_equal_oracle(ctx, spec, ray):
    """Accumulate mode on the optics feature sets that run specialized kernels (specular or
    Lambertian wrap over an air gap, rough; Lambertian with Rayleigh scattering, the
    reference's benchmark optics of scripts/bench_metal.jl:7-13): histogram, first times and
    counters identical to the oracle's reduction of its records."""
    from crysplight_amd.params import OpticalParams
    op = OpticalParams(1.95, 1.45, 3000.0, ray, 0.99, spec, True, deg2rad(1.3), 1.0)
    ro = readout(SipmGrid(6.0, 8, 8))
    kw = dict(n_photons=100_000, seed=1, pos=(24.0, 24.0, 18.6), tau_ns=0.0)
    acc = T.transport_accumulate(K.BOX_CRYSP, op, ro, K.TB, ctx=ctx, **kw)
    r = oracle.run_photons(K.BOX_CRYSP, op, ro, K.TB, **kw)
    assert np.array_equal(acc.counts, r["counts"])
    assert np.array_equal(acc.first_ns, r["first_ns"])
    for name in ("ndet", "nabs_bulk", "nabs_wall", "nabs_sipm", "ncap", "nscat"):
        assert getattr(acc, name) == r[name], name


def test_op_counts_match_oracle(ctx):
    kw = dict(n_photons=50_000, seed=1234, pos=K.C1_POS, tau_ns=K.C1_TAU)
    ctx.set_op_counting(True)
    try:
        ctx.op_counts(reset=True)
        T.transport_records(K.C1_BOX, K.C1_OP, K.C1_RO, ctx=ctx, **kw)
        ops = ctx.op_counts(reset=True)
    finally:
        ctx.set_op_counting(False)
    r = oracle.transport_records(K.C1_BOX, K.C1_OP, K.C1_RO, count_ops=True, **kw)
    assert list(ops) == [r["ops"][n] for n in oracle.OPCOUNT_NAMES]


def test_zero_photons_and_bounce_cap(ctx):
    """An empty call writes nothing and leaves the accumulator empty; a small max_bounces
    caps photons (CAP status, src/transport.jl:228-232) exactly where the oracle does."""
    rec = T.transport_records(K.C1_BOX, K.C1_OP, K.C1_RO, n_photons=0, seed=1, pos=K.C1_POS,
                              tau_ns=K.C1_TAU, ctx=ctx)
    assert rec.status.size == 0
    acc = T.transport_accumulate(K.C1_BOX, K.C1_OP, K.C1_RO, K.C1_TB, n_photons=0, seed=1,
                                 pos=K.C1_POS, tau_ns=K.C1_TAU, ctx=ctx)
    assert acc.total_terminated() == 0 and not acc.counts.any() and np.all(np.isinf(acc.first_ns))
    g, r, frac = _records_case(ctx, K.C1_BOX, K.C1_OP, K.C1_RO, 20_000, 77, K.C1_POS,
                               K.C1_TAU, maxb=3)
    assert frac == 1.0
    assert np.count_nonzero(g.status == 5) > 1000
