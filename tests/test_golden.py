"""The committed golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py
from the pinned oracle): the oracle must still reproduce them on the CPU, and the HIP path
must reproduce them bit for bit on the GPU (records, accumulate-mode maps and counters,
event maps, Philox words)."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.golden import make_fixtures as F

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / name) as z:
        return {k: z[k] for k in z.files}


def _same_records(got, want):
    for k in ("status", "idx", "t", "bounces"):
        assert np.array_equal(np.asarray(got[k]), want[k]), k


# ---------------------------------------------------------------- the oracle (CPU) ----

def test_oracle_philox_draws_fixture():
    assert np.array_equal(F.draws(), load("philox_draws.npz")["draws"])


@pytest.mark.parametrize("name,kw", [("records_c1.npz", F.records_kw_c1),
                                     ("records_all_processes.npz", F.records_kw_all)])
def test_oracle_records_fixture(name, kw):
    _same_records(F.records(kw()), load(name))


def test_oracle_accumulate_fixture():
    got, want = F.accumulate(F.accumulate_kw_c1()), load("accumulate_c1_1e6.npz")
    for k in ("counts", "first_ns", "counters"):
        assert np.array_equal(got[k], want[k]), k
    assert want["counters"][0] > 0 and want["counts"].sum() == want["counters"][0]


def test_oracle_events_fixture():
    z = load("events_bgo.npz")
    box, op, ro, tau = F.events_setup()
    r = oracle.run_events(box, op, ro, z["xyzt"], z["nph"], z["event"], int(z["n_events"]),
                          seed=F.EV_SEED, tau_ns=tau)
    for k in ("maps", "tmin", "npe"):
        assert np.array_equal(r[k], z[k]), k
    assert z["npe"].sum() > 0


# --------------------------------------------------------------- the HIP path (GPU) ----

@pytest.mark.gpu
def test_gpu_philox_draws_fixture(ctx):
    from crysplight_amd import transport as T
    want = load("philox_draws.npz")
    for p, row in zip(want["pids"], want["draws"]):
        assert np.array_equal(T.debug_draws(1234, 1, 64, pid0=int(p) - 1, ctx=ctx)[0], row), p


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw", [("records_c1.npz", F.records_kw_c1),
                                     ("records_all_processes.npz", F.records_kw_all)])
def test_gpu_records_fixture(ctx, name, kw):
    from crysplight_amd import transport as T
    kw = dict(kw())
    box, op, ro = kw.pop("box"), kw.pop("op"), kw.pop("ro")
    g = T.transport_records(box, op, ro, ctx=ctx, **kw)
    _same_records({"status": g.status, "idx": g.idx, "t": g.t, "bounces": g.bounces}, load(name))


@pytest.mark.gpu
def test_gpu_accumulate_fixture(ctx):
    from crysplight_amd import transport as T
    kw = dict(F.accumulate_kw_c1())
    box, op, ro, tb = kw.pop("box"), kw.pop("op"), kw.pop("ro"), kw.pop("tb")
    acc = T.transport_accumulate(box, op, ro, tb, ctx=ctx, **kw)
    want = load("accumulate_c1_1e6.npz")
    assert np.array_equal(acc.counts, want["counts"])
    assert np.array_equal(acc.first_ns, want["first_ns"])
    got = np.array([getattr(acc, n) for n in oracle.COUNTER_NAMES], np.int64)
    assert np.array_equal(got, want["counters"])


@pytest.mark.gpu
def test_gpu_events_fixture(ctx):
    from crysplight_amd import events as E
    z = load("events_bgo.npz")
    nev = int(z["n_events"])
    dep = E.Deposits(z["xyzt"], z["nph"], z["event"], np.zeros(z["nph"].size, np.float32), nev)
    box, op, ro, tau = F.events_setup()
    res = E.transport_events(box, op, ro, dep, seed=F.EV_SEED, tau_ns=tau, ctx=ctx)
    assert np.array_equal(res.maps.reshape(nev, -1), z["maps"])
    assert np.array_equal(res.tmin.reshape(nev, -1), z["tmin"])
    assert np.array_equal(res.npe, z["npe"])
