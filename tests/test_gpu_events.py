"""GPU parity of the event mode (work queue, device maps, hipGraph replay, autotuner)
against the oracle's per-event reduction (src/events.jl:73-88)."""
import numpy as np
import pytest

import oracle
from crysplight_amd import events as E
from crysplight_amd._lib import CryspError
from crysplight_amd.params import Box, SipmGrid, deg2rad, readout
from tests import cases as K

pytestmark = pytest.mark.gpu

RO8 = readout(SipmGrid(3.2, 8, 8))


def _oracle_maps(box, op, ro, dep, seed, tau, tau_rise=0.0, pid0=0):
    return oracle.run_events(box, op, ro, dep.xyzt, dep.nph, dep.event, dep.n_events,
                             seed=seed, pid0=pid0, tau_ns=tau, tau_rise_ns=tau_rise)


def _check(res, ref, nev):
    assert np.array_equal(res.maps.reshape(nev, -1), ref["maps"])
    assert np.array_equal(res.npe, ref["npe"])
    assert np.array_equal(res.tmin.reshape(nev, -1), ref["tmin"])


@pytest.fixture(scope="module")
def bgo_events():
    bgo = E.GammaMaterial("xcom_BGO.csv", 7.13)
    dep, truth = E.gamma_events(K.C2_BOX, bgo, 8500.0, n_events=400, seed=7)
    return dep, truth


def test_events_bgo_host_arrays(ctx, bgo_events):
    dep, _ = bgo_events
    res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
    assert res.n_photons == dep.n_photons
    _check(res, _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU), dep.n_events)


def test_events_device_arrays_graph_replays(ctx, bgo_events):
    """Device-resident inputs go through the captured graph; replays with changed
    contents (same pointers) must track the contents, with eager work in between."""
    import torch
    dep, _ = bgo_events
    dev = torch.device("cuda", 0)
    xyzt = torch.from_numpy(dep.xyzt).to(dev)
    nph = torch.from_numpy(dep.nph).to(dev)
    ev = torch.from_numpy(dep.event).to(dev)
    nev = dep.n_events
    out = (torch.zeros((nev, 8, 8), dtype=torch.int32, device=dev),
           torch.zeros((nev, 8, 8), dtype=torch.float32, device=dev),
           torch.zeros(nev, dtype=torch.int32, device=dev))
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    for rep in range(3):
        r = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                               device_arrays=(xyzt, nph, ev), out=out)
        assert r.n_photons == dep.n_photons
        assert np.array_equal(out[0].cpu().numpy().reshape(nev, -1), ref["maps"]), rep
        assert np.array_equal(out[1].cpu().numpy().reshape(nev, -1), ref["tmin"]), rep
        assert np.array_equal(out[2].cpu().numpy(), ref["npe"]), rep
        _ = (xyzt * 2).sum().item()                 # eager work between replays
    # same pointers, new contents: half the photons
    half = E.Deposits(dep.xyzt, dep.nph // 2, dep.event, dep.e_kev, nev)
    nph.copy_(torch.from_numpy(half.nph))
    r = E.transport_events(K.C2_BOX, K.C2_OP, RO8, half, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=(xyzt, nph, ev), out=out)
    ref2 = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, half, 7, K.C2_TAU)
    assert r.n_photons == half.n_photons
    assert np.array_equal(out[0].cpu().numpy().reshape(nev, -1), ref2["maps"])


def test_events_graph_follows_new_output_arrays(ctx, bgo_events):
    """A replay with the same inputs and maps array but new first-time and npe arrays must
    write the new arrays (each call's outputs are part of the captured graph's key)."""
    import torch
    dep, _ = bgo_events
    dev = torch.device("cuda", 0)
    darr = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, dep.event))
    nev = dep.n_events
    maps = torch.zeros((nev, 8, 8), dtype=torch.int32, device=dev)
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    for rep in range(3):
        tmin = torch.zeros((nev, 8, 8), dtype=torch.float32, device=dev)
        npe = torch.full((nev,), -1, dtype=torch.int32, device=dev)
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=darr, out=(maps, tmin, npe))
        assert np.array_equal(tmin.cpu().numpy().reshape(nev, -1), ref["tmin"]), rep
        assert np.array_equal(npe.cpu().numpy(), ref["npe"]), rep
        assert np.array_equal(maps.cpu().numpy().reshape(nev, -1), ref["maps"]), rep


def test_events_op_counts_match_oracle(ctx, bgo_events):
    """The counting kernel variant in event mode, eager (host arrays) and through the
    captured graph (device arrays, twice), against the oracle's counters of the same
    photons."""
    import torch
    dep, _ = bgo_events
    ref = oracle.run_events(K.C2_BOX, K.C2_OP, RO8, dep.xyzt, dep.nph, dep.event, dep.n_events,
                            seed=7, tau_ns=K.C2_TAU, count_ops=True)
    want = [ref["ops"][n] for n in oracle.OPCOUNT_NAMES]
    assert want[0] == dep.n_photons and want[8] == int(ref["npe"].sum())
    dev = torch.device("cuda", 0)
    darr = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, dep.event))
    ctx.set_op_counting(True)
    try:
        for rep, kw in enumerate([{}, {"device_arrays": darr}, {"device_arrays": darr}]):
            ctx.op_counts(reset=True)
            E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx, **kw)
            assert list(ctx.op_counts(reset=True)) == want, rep
    finally:
        ctx.set_op_counting(False)


def test_events_rise_time_uneven(ctx):
    """Batched uneven single-deposit events with rise + decay emission."""
    dep = E.synthetic_deposits(K.C1_BOX, n_events=300, seed=11, e_kev=(10.0, 511.0),
                               yield_per_mev=2000.0)
    assert dep.nph.min() < 100 < 900 < dep.nph.max()
    res = E.transport_events(K.C1_BOX, K.C1_OP, RO8, dep, seed=5, tau_ns=1000.0,
                             tau_rise_ns=20.0, pid0=12345, ctx=ctx)
    _check(res, _oracle_maps(K.C1_BOX, K.C1_OP, RO8, dep, 5, 1000.0, 20.0, 12345), 300)


def test_events_photon_ids_cross_2_32(ctx, bgo_events):
    """Event-mode stream numbers pid0 + g + 1 run across 2^32 within the launch, under a
    seed with key bits above 32."""
    dep, _ = bgo_events
    sub = E.Deposits(dep.xyzt[:40], dep.nph[:40], dep.event[:40], dep.e_kev[:40],
                     int(dep.event[39]) + 1)
    pid0 = 2 ** 32 - int(sub.nph[:20].sum())
    seed = 2 ** 40 + 7
    res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, sub, seed=seed, tau_ns=K.C2_TAU,
                             pid0=pid0, ctx=ctx)
    _check(res, _oracle_maps(K.C2_BOX, K.C2_OP, RO8, sub, seed, K.C2_TAU, pid0=pid0),
           sub.n_events)


def test_events_generic_and_specialized_optics(ctx):
    """Event mode runs a specialized kernel when the optics match one of its feature sets
    (ESR: specular wrap over an air gap, rough; Teflon: Lambertian wrap over an air gap,
    rough; Teflon with Rayleigh scattering) and the generic kernel otherwise.  Each flavour
    against the oracle on the same deposits: ESR, Teflon, Teflon with Rayleigh, a smooth ESR
    wrap (sigma_alpha 0), a front-painted wrap (no air gap) and the all-processes optics with
    a disc and surround readout (generic)."""
    from crysplight_amd.params import OpticalParams
    rng = np.random.default_rng(21)
    nd = 30
    L = np.array(K.C2_BOX.L, np.float32)
    xyz = (rng.uniform(0.05, 0.95, (nd, 3)) * L).astype(np.float32)
    xyzt = np.column_stack([xyz, rng.uniform(0, 3, nd)]).astype(np.float32)
    nph = rng.integers(50, 400, nd).astype(np.int32)
    ev = np.sort(rng.integers(0, 10, nd)).astype(np.int32)
    dep = E.Deposits(xyzt, nph, ev, np.zeros(nd, np.float32), 10)
    s13 = deg2rad(1.3)
    ro_all = readout(SipmGrid(3.2, 8, 8), center=(12.8, 12.8), radius=11.0, sur_R=0.9)
    cases = [
        (K.C2_OP, RO8),                                                              # ESR
        (OpticalParams(1.79, 1.5, 400.0, float("inf"), 0.98, False, True, s13, 0.3), RO8),  # Teflon
        (OpticalParams(1.95, 1.45, 3000.0, 333.3, 0.99, False, True, s13, 1.0), RO8),    # + Rayleigh
        (OpticalParams(2.15, 1.5, 300.0, float("inf"), 0.98, True, True, 0.0, 0.3), RO8),   # smooth
        (OpticalParams(1.79, 1.5, 400.0, float("inf"), 0.95, False, False, s13, 0.3), RO8),  # painted
        (OpticalParams(1.95, 1.45, 500.0, 400.0, 0.98, False, True, deg2rad(6), 0.4), ro_all),
    ]
    for k, (op, ro) in enumerate(cases):
        res = E.transport_events(K.C2_BOX, op, ro, dep, seed=31, tau_ns=300.0, ctx=ctx)
        ref = oracle.run_events(K.C2_BOX, op, ro, dep.xyzt, dep.nph, dep.event, dep.n_events,
                                seed=31, tau_ns=300.0)
        assert np.array_equal(res.maps.reshape(10, -1), ref["maps"]), k
        assert np.array_equal(res.tmin.reshape(10, -1), ref["tmin"]), k
        assert np.array_equal(res.npe, ref["npe"]), k


def test_events_beyond_2_31_photons_in_one_launch(ctx):
    """One launch past 2^31 photons (64-bit photon numbers, offsets and queue cursor), on the
    CsI(Tl)/Teflon optics, checked by size-independent properties: the photon total, npe
    equal to every map's sum, two launches identical, and the detected fraction equal within
    5 sigma to the oracle's on the same deposits at 200 photons each."""
    rng = np.random.default_rng(5)
    nd = 1000
    L = np.array(K.C1_BOX.L, np.float32)
    xyz = (rng.uniform(0.05, 0.95, (nd, 3)) * L).astype(np.float32)
    xyzt = np.column_stack([xyz, np.zeros(nd)]).astype(np.float32)
    nph = np.full(nd, 2_150_000, np.int32)
    ev = np.arange(nd, dtype=np.int32) // 4
    dep = E.Deposits(xyzt, nph, ev, np.zeros(nd, np.float32), nd // 4)
    total = int(nph.astype(np.int64).sum())
    assert total > 2 ** 31
    a = E.transport_events(K.C1_BOX, K.C1_OP, RO8, dep, seed=3, tau_ns=1000.0, ctx=ctx)
    assert a.n_photons == total
    maps = a.maps.reshape(dep.n_events, -1).astype(np.int64)
    assert np.array_equal(a.npe.astype(np.int64), maps.sum(1))
    b = E.transport_events(K.C1_BOX, K.C1_OP, RO8, dep, seed=3, tau_ns=1000.0, ctx=ctx)
    assert np.array_equal(a.maps, b.maps) and np.array_equal(a.tmin, b.tmin)
    small = E.Deposits(xyzt, np.full(nd, 200, np.int32), ev, np.zeros(nd, np.float32), nd // 4)
    ref = oracle.run_events(K.C1_BOX, K.C1_OP, RO8, small.xyzt, small.nph, small.event,
                            small.n_events, seed=3, tau_ns=1000.0)
    n_small = 200 * nd
    p_ref = ref["npe"].sum() / n_small
    p_big = maps.sum() / total
    sigma = np.sqrt(p_ref * (1 - p_ref) / n_small)
    assert abs(p_big - p_ref) < 5 * sigma, (p_big, p_ref, sigma)


def test_events_schedule_independent_and_autotune(ctx, bgo_events):
    dep, _ = bgo_events
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    old = ctx.get_schedule()
    try:
        for wg, ppt in [(64, 1), (1024, 8), (192, 3)]:
            ctx.set_schedule(wg, ppt)
            res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
            _check(res, ref, dep.n_events)
        best, times = E.autotune(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                                 wg_sizes=(64, 256, 1024), ppts=(1, 4), reps=1)
        assert best in [(w, p) for w in (64, 256, 1024) for p in (1, 4)]
        assert np.all(times > 0)
        assert ctx.get_schedule() == best
    finally:
        ctx.set_schedule(*old)


def test_full_knob_autotuners(ctx, bgo_events):
    """The tuners over workgroup size, photons per thread, queue chunk and refill batch (event
    and accumulate mode) time every combination, keep the fastest in the context, and the
    maps computed with the tuned knobs are still the oracle's."""
    from crysplight_amd import transport as T
    dep, _ = bgo_events
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    old = ctx.get_schedule()
    try:
        best, times = E.autotune_queue(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                                       wg_sizes=(128, 256), ppts=(4,), chunks=(32, 128),
                                       refill_mins=(6, 12), reps=1)
        assert times.shape == (2, 1, 2, 2) and np.all(np.isfinite(times)) and np.all(times > 0)
        assert best[0] in (128, 256) and best[1] == 4 and best[2] in (32, 128) and best[3] in (6, 12)
        assert ctx.get_schedule() == best[:2]
        res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
        _check(res, ref, dep.n_events)
        best_a, times_a = T.autotune_accumulate(K.C1_BOX, K.C1_OP, K.C1_RO, K.C1_TB, n_photons=100_000,
                                                seed=5, pos=K.C1_POS, tau_ns=K.C1_TAU, ctx=ctx,
                                                wg_sizes=(256,), ppts=(1, 4), chunks=(64, 128),
                                                refill_mins=(12,), reps=1)
        assert times_a.shape == (1, 2, 2, 1) and np.all(times_a > 0)
        acc = T.transport_accumulate(K.C1_BOX, K.C1_OP, K.C1_RO, K.C1_TB, n_photons=100_000, seed=5,
                                     pos=K.C1_POS, tau_ns=K.C1_TAU, ctx=ctx)
        r = oracle.run_photons(K.C1_BOX, K.C1_OP, K.C1_RO, K.C1_TB, n_photons=100_000,
                               seed=5, pos=K.C1_POS, tau_ns=K.C1_TAU)
        assert np.array_equal(acc.counts, r["counts"]) and acc.ndet == r["ndet"]
    finally:
        ctx.set_schedule(*old)
        ctx.set_queue(256, 4)


def test_events_queue_parameters_do_not_change_results(ctx, bgo_events):
    """The work-queue grab size and the lane-refill batch only reorder which lane runs which
    photon: every (chunk, refill_min) gives the oracle's maps, including a chunk of one
    photon, chunks smaller and larger than a wave, and refills of one lane or of 64."""
    dep, _ = bgo_events
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    try:
        for chunk, rmin in [(1, 1), (5, 3), (64, 64), (4096, 12), (1 << 20, 7)]:
            ctx.set_queue(chunk, rmin)
            res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
            _check(res, ref, dep.n_events)
        for bad in [(0, 12), (1 << 21, 12), (128, 0), (128, 65)]:
            with pytest.raises(CryspError):
                ctx.set_queue(*bad)
    finally:
        ctx.set_queue(256, 4)


def test_events_bad_event_index_is_an_error(ctx, bgo_events):
    dep, _ = bgo_events
    bad = E.Deposits(dep.xyzt[:5], dep.nph[:5], np.array([0, 1, 2, 3, 99], np.int32),
                     dep.e_kev[:5], 4)
    with pytest.raises(Exception):
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, bad, seed=7, tau_ns=K.C2_TAU, ctx=ctx)


def test_events_empty_tables_and_zero_photon_deposits(ctx):
    """No deposits at all; and deposits with zero photons mixed with events that have no
    deposit (src/events.jl:59-91 keeps such events with empty maps)."""
    empty = E.Deposits(np.zeros((0, 4), np.float32), np.zeros(0, np.int32),
                       np.zeros(0, np.int32), np.zeros(0, np.float32), 5)
    r = E.transport_events(K.C2_BOX, K.C2_OP, RO8, empty, seed=3, tau_ns=K.C2_TAU, ctx=ctx)
    assert r.n_photons == 0 and not r.maps.any() and not r.npe.any()
    assert np.all(np.isinf(r.tmin))
    rng = np.random.default_rng(4)
    nd = 40
    xyzt = np.column_stack([rng.uniform(1, 24, nd), rng.uniform(1, 24, nd), rng.uniform(1, 19, nd),
                            rng.uniform(0, 5, nd)]).astype(np.float32)
    nph = rng.integers(0, 400, nd).astype(np.int32)
    nph[::3] = 0
    ev = np.sort(rng.choice(np.arange(0, 30, 2), nd)).astype(np.int32)   # odd events empty
    dep = E.Deposits(xyzt, nph, ev, np.zeros(nd, np.float32), 31)
    res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=3, tau_ns=K.C2_TAU, ctx=ctx)
    assert res.n_photons == int(nph.sum())
    _check(res, _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 3, K.C2_TAU), 31)
    assert not res.npe[1::2].any()


def test_events_multibox_matches_per_deposit_oracle(ctx):
    """Deposits of three crystal thicknesses in one queue: photon g of the table keeps
    stream pid0 + g + 1 and is transported in its deposit's crystal."""
    boxes = [Box((25.6, 25.6, t)) for t in (10.0, 20.0, 30.0)]
    rng = np.random.default_rng(8)
    nd = 45
    dep_box = rng.integers(0, 3, nd).astype(np.int32)
    L = np.array([b.L for b in boxes], np.float32)[dep_box]
    xyz = rng.uniform(0.05, 0.95, (nd, 3)).astype(np.float32) * L
    xyzt = np.column_stack([xyz, rng.uniform(0, 3, nd)]).astype(np.float32)
    nph = rng.integers(0, 300, nd).astype(np.int32)
    ev = np.sort(rng.integers(0, 12, nd)).astype(np.int32)
    dep = E.Deposits(xyzt, nph, ev, np.zeros(nd, np.float32), 12)
    pid0 = 1000
    res = E.transport_events_multibox(boxes, dep_box, K.C2_OP, RO8, dep, seed=9, tau_ns=K.C2_TAU,
                                      pid0=pid0, ctx=ctx)
    maps = np.zeros((12, 64), np.int64)
    tmin = np.full((12, 64), np.inf, np.float32)
    g0 = 0
    for d in range(nd):
        n = int(nph[d])
        if n:
            rec = oracle.transport_records(boxes[dep_box[d]], K.C2_OP, RO8, n_photons=n, seed=9,
                                           pos=tuple(xyzt[d, :3]), t0s=[xyzt[d, 3]],
                                           pid_offset=pid0 + g0, tau_ns=K.C2_TAU)
            det = rec["status"] == 1
            cell = np.abs(rec["idx"][det].astype(np.int64)) - 1
            np.add.at(maps[ev[d]], cell, 1)
            np.minimum.at(tmin[ev[d]], cell, rec["t"][det])
        g0 += n
    assert np.array_equal(res.maps.reshape(12, -1), maps)
    assert np.array_equal(res.npe, maps.sum(1))
    assert np.array_equal(res.tmin.reshape(12, -1), tmin)
    # the same queue from device-resident tables into device-resident maps
    import torch
    dev = torch.device("cuda", 0)
    darr = tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                 for a in (dep.xyzt, dep.nph, dep.event, dep_box))
    out = (torch.zeros((12, 8, 8), dtype=torch.int32, device=dev),
           torch.zeros((12, 8, 8), dtype=torch.float32, device=dev),
           torch.zeros(12, dtype=torch.int32, device=dev))
    for rep in range(2):
        r2 = E.transport_events_multibox(boxes, None, K.C2_OP, RO8, dep, seed=9, tau_ns=K.C2_TAU,
                                         pid0=pid0, ctx=ctx, device_arrays=darr, out=out)
        assert r2.n_photons == int(nph.sum())
        assert np.array_equal(out[0].cpu().numpy().reshape(12, -1), maps), rep
        assert np.array_equal(out[1].cpu().numpy().reshape(12, -1), tmin), rep
        assert np.array_equal(out[2].cpu().numpy(), maps.sum(1)), rep


def test_device_tables_with_bad_indices_raise(ctx, bgo_events):
    """Device-resident event and crystal indices are checked on the device: an index out
    of range makes the call fail (CRYSP_EINVAL) instead of reading outside the tables, and
    the context stays usable."""
    import torch
    dep, _ = bgo_events
    dev = torch.device("cuda", 0)
    nev = dep.n_events
    out = (torch.zeros((nev, 8, 8), dtype=torch.int32, device=dev),
           torch.zeros((nev, 8, 8), dtype=torch.float32, device=dev),
           torch.zeros(nev, dtype=torch.int32, device=dev))
    ev_bad = dep.event.copy()
    ev_bad[5] = nev                                   # one past the last event
    darr = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, ev_bad))
    with pytest.raises(CryspError, match="dep_event"):
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=darr, out=out)
    # host counts with device event indices take the same check
    darr_h = (torch.from_numpy(dep.xyzt).to(dev), dep.nph.copy(), torch.from_numpy(ev_bad).to(dev))
    with pytest.raises(CryspError, match="dep_event"):
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=darr_h, out=out)
    boxes = [K.C2_BOX, Box((25.6, 25.6, 10.0))]
    box_bad = np.zeros(dep.xyzt.shape[0], np.int32)
    box_bad[3] = 2                                    # only crystals 0 and 1 exist
    darr4 = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, dep.event, box_bad))
    with pytest.raises(CryspError, match="dep_box"):
        E.transport_events_multibox(boxes, None, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU,
                                    ctx=ctx, device_arrays=darr4, out=out)
    res = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
    _check(res, _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU), nev)


def test_device_tables_longer_than_one_scan_pass(ctx):
    """The launch set-up scans device-resident photon counts in passes of 16384 deposits
    with a carry between passes, and checks the index tables in the same passes.  40000
    deposits (three passes, the last partial; counts 0 to about 12, so zeros sit inside the
    runs) of three deposits per event: maps, first times, npe and the photon total against
    the oracle; a negative count in the third pass is an error; events out of order only in
    the second pass make the owned-event path stand aside (the atomic path's maps)."""
    import torch
    n = 40000
    single = E.synthetic_deposits(K.C2_BOX, n_events=n, seed=23, e_kev=(0.2, 1.5),
                                  yield_per_mev=8500.0)
    ev = (np.arange(n) // 3).astype(np.int32)
    nev = int(ev[-1]) + 1
    dep = E.Deposits(single.xyzt, single.nph, ev, single.e_kev, nev)
    assert (dep.nph == 0).any() and dep.nph.max() > 5
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    dev = torch.device("cuda", 0)
    out = (torch.zeros((nev, 8, 8), dtype=torch.int32, device=dev),
           torch.zeros((nev, 8, 8), dtype=torch.float32, device=dev),
           torch.zeros(nev, dtype=torch.int32, device=dev))
    darr = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, dep.event))
    print(f"{n} deposits, {nev} events, {int(dep.nph.sum())} photons", flush=True)
    for rep in range(2):
        r = E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                               device_arrays=darr, out=out)
        assert r.n_photons == int(dep.nph.sum()), rep
        assert np.array_equal(out[0].cpu().numpy().reshape(nev, -1), ref["maps"]), rep
        assert np.array_equal(out[1].cpu().numpy().reshape(nev, -1), ref["tmin"]), rep
        assert np.array_equal(out[2].cpu().numpy(), ref["npe"]), rep
    nph_bad = dep.nph.copy()
    nph_bad[2 * 16384 + 5] = -1
    darr_bad = (darr[0], torch.from_numpy(nph_bad).to(dev), darr[2])
    with pytest.raises(CryspError, match="negative"):
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=darr_bad, out=out)
    # swap two deposits' event indices (and with them nothing else: the oracle follows the
    # same table) inside the second pass
    ev_uns = dep.event.copy()
    i = 16384 + 1000
    ev_uns[i], ev_uns[i + 3] = ev_uns[i + 3], ev_uns[i]
    dep_uns = E.Deposits(dep.xyzt, dep.nph, ev_uns, dep.e_kev, nev)
    ref_uns = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep_uns, 7, K.C2_TAU)
    darr_uns = (darr[0], darr[1], torch.from_numpy(ev_uns).to(dev))
    ctx.set_event_sessions(True, 1024)
    try:
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep_uns, seed=7, tau_ns=K.C2_TAU, ctx=ctx,
                           device_arrays=darr_uns, out=out)
    finally:
        ctx.set_event_sessions(False, 1024)
    assert np.array_equal(out[0].cpu().numpy().reshape(nev, -1), ref_uns["maps"])
    assert np.array_equal(out[2].cpu().numpy(), ref_uns["npe"])


# ---- events owned by workgroups ---------------------------------------------------
# With deposits ordered by event, workgroups own whole events and write each map row once
# (helpers' partial maps merged afterwards); otherwise, or when switched off, every
# detection is a global atomic.  Both must give the oracle's maps on the same deposits.

def _with_sessions(ctx, enable, piece, fn):
    ctx.set_event_sessions(enable, piece)
    try:
        return fn()
    finally:
        ctx.set_event_sessions(False, 1024)


@pytest.mark.parametrize("piece", [1, 37, 512, 1 << 20])
def test_owned_events_equal_oracle_any_piece(ctx, bgo_events, piece):
    """Pieces of one photon up to whole events per claim, host and device tables."""
    import torch
    dep, _ = bgo_events
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 7, K.C2_TAU)
    print(f"piece {piece}, {dep.n_events} events, {dep.n_photons} photons", flush=True)
    res = _with_sessions(ctx, True, piece, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx))
    _check(res, ref, dep.n_events)
    dev = torch.device("cuda", 0)
    darr = tuple(torch.from_numpy(a).to(dev) for a in (dep.xyzt, dep.nph, dep.event))
    res = _with_sessions(ctx, True, piece, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx, device_arrays=darr))
    assert np.array_equal(np.asarray(res.maps).reshape(dep.n_events, -1), ref["maps"])
    assert np.array_equal(np.asarray(res.tmin).reshape(dep.n_events, -1), ref["tmin"])
    assert np.array_equal(np.asarray(res.npe), ref["npe"])


def test_owned_events_off_equals_on(ctx, bgo_events):
    dep, _ = bgo_events
    on = _with_sessions(ctx, True, 512, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx))
    off = _with_sessions(ctx, False, 512, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx))
    for a in ("maps", "tmin", "npe"):
        assert np.array_equal(np.asarray(getattr(on, a)), np.asarray(getattr(off, a))), a


def test_few_large_events_are_helped(ctx):
    """Three events of 1.5e5 photons each, far more workgroups than events: the owners'
    events are split with helpers, whose partial maps are merged afterwards."""
    rng = np.random.default_rng(21)
    nd = 9
    xyzt = np.column_stack([rng.uniform(1, 24, nd), rng.uniform(1, 24, nd), rng.uniform(1, 19, nd),
                            rng.uniform(0, 5, nd)]).astype(np.float32)
    nph = np.full(nd, 50_000, np.int32)
    ev = np.repeat(np.arange(3, dtype=np.int32), 3)
    dep = E.Deposits(xyzt, nph, ev, np.zeros(nd, np.float32), 3)
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 9, K.C2_TAU)
    print(f"{dep.n_photons} photons in 3 events", flush=True)
    res = _with_sessions(ctx, True, 512, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=9, tau_ns=K.C2_TAU, ctx=ctx))
    _check(res, ref, 3)


def test_owned_events_many_small_and_empty(ctx):
    """6000 events of 0-30 photons, some events with no deposit at all, zero-photon
    deposits, the last events empty."""
    rng = np.random.default_rng(22)
    nev = 6000
    evs = np.sort(rng.choice(nev - 50, 9000)).astype(np.int32)
    nd = evs.size
    xyzt = np.column_stack([rng.uniform(0, 25.6, nd), rng.uniform(0, 25.6, nd), rng.uniform(0, 20, nd),
                            rng.uniform(0, 50, nd)]).astype(np.float32)
    nph = rng.integers(0, 31, nd).astype(np.int32)
    dep = E.Deposits(xyzt, nph, evs, np.zeros(nd, np.float32), nev)
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, dep, 3, K.C2_TAU)
    res = _with_sessions(ctx, True, 512, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, dep, seed=3, tau_ns=K.C2_TAU, ctx=ctx))
    _check(res, ref, nev)
    assert (np.asarray(res.npe)[nev - 50:] == 0).all()
    assert np.isinf(np.asarray(res.tmin).reshape(nev, -1)[nev - 50:]).all()


def test_unordered_tables_take_the_atomic_path(ctx, bgo_events):
    """Deposits out of event order (host and device tables): the device check sends the
    call to per-detection atomics, with the same maps."""
    import torch
    dep, _ = bgo_events
    perm = np.random.default_rng(4).permutation(dep.nph.size)
    # photons are numbered in table order, so the permuted table is a different photon
    # set: compare with the oracle on the permuted table itself
    shuf = E.Deposits(np.ascontiguousarray(dep.xyzt[perm]), np.ascontiguousarray(dep.nph[perm]),
                      np.ascontiguousarray(dep.event[perm]), dep.e_kev[perm], dep.n_events)
    ref = _oracle_maps(K.C2_BOX, K.C2_OP, RO8, shuf, 7, K.C2_TAU)
    res = _with_sessions(ctx, True, 512, lambda: E.transport_events(
        K.C2_BOX, K.C2_OP, RO8, shuf, seed=7, tau_ns=K.C2_TAU, ctx=ctx))
    _check(res, ref, dep.n_events)
    dev = torch.device("cuda", 0)
    darr = tuple(torch.from_numpy(a).to(dev) for a in (shuf.xyzt, shuf.nph, shuf.event))
    for rep in range(2):
        res = _with_sessions(ctx, True, 512, lambda: E.transport_events(
            K.C2_BOX, K.C2_OP, RO8, shuf, seed=7, tau_ns=K.C2_TAU, ctx=ctx, device_arrays=darr))
        assert np.array_equal(np.asarray(res.maps).reshape(dep.n_events, -1), ref["maps"]), rep
        assert np.array_equal(np.asarray(res.npe), ref["npe"]), rep


def test_wave_timeline(ctx, bgo_events):
    """The diagnostics timeline: every wave of the last launch reports its end; the spread
    of the ends lies within the launch."""
    dep, _ = bgo_events
    ctx.set_wave_timeline(True)
    ctx.set_timing(True)
    ctx.kernel_time(reset=True)
    try:
        E.transport_events(K.C2_BOX, K.C2_OP, RO8, dep, seed=7, tau_ns=K.C2_TAU, ctx=ctx)
        ms, n = ctx.kernel_time(reset=True)
        before_end = ctx.wave_timeline()
    finally:
        ctx.set_wave_timeline(False)
        ctx.set_timing(False)
    assert before_end.size > 0 and before_end.min() == 0.0
    assert before_end.max() <= ms / n / 1e3
