"""Write the golden fixtures of the transport from the oracle (oracle/crysp_oracle.c).

The oracle is trusted for this only because it passes its own pins first
(tests/test_oracle.py): the Random123 known-answer vectors and the rocRAND cross-check of
Philox4x32-10, the exhaustive check of randu, the closed forms and exact cases of the
reference's tests (test/runtests.jl:35-174), and its statistical anchors.  The fixtures then
hold those results fixed: a later change to the oracle or to the device path that moves a
single draw, record or map cell fails tests/test_golden.py.

Run from the repository root:   python tests/golden/make_fixtures.py

Files (numpy .npz, no object arrays):
* philox_draws.npz -- the first 64 u32 words of the streams of pids 1, 2 and 2^32 + 5 under
  seed 1234 (src/philox.jl:34-62).
* records_c1.npz -- per-photon records (status, idx, t, bounces; src/kernel.jl:159-305) of
  4096 photons of the CsI(Tl)/Teflon point source (seed 1234, decay 1000 ns).
* records_all_processes.npz -- the same for the all-processes configuration of
  test/runtests.jl:229-232 (seed 8, decay 800 ns, Rayleigh, disc readout with surround).
* accumulate_c1_1e6.npz -- the CsI(Tl)/Teflon point source at 1e6 photons in accumulate mode
  (src/kernel.jl:337-365): time-binned counts, first times and the nine counters.
* events_bgo.npz -- 64 gamma events of 511 keV in BGO (ESR over an air gap): the deposit
  table they start from and the per-event maps, first times and npe (src/events.jl:65-91).
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from crysplight_amd import events as E  # noqa: E402
from crysplight_amd.params import SipmGrid, readout  # noqa: E402
from tests import cases as K  # noqa: E402

PIDS = (1, 2, 2 ** 32 + 5)
N_RECORDS = 4096
EV_SEED = 7
EV_N = 64


def records_kw_c1():
    return dict(box=K.C1_BOX, op=K.C1_OP, ro=K.C1_RO, n_photons=N_RECORDS, seed=K.C1_SEED,
                pos=K.C1_POS, tau_ns=K.C1_TAU)


def records_kw_all():
    return dict(box=K.BOX_CRYSP, op=K.OP_ALL, ro=K.RO_ALL, n_photons=N_RECORDS, seed=8,
                pos=K.POS_CRYSP, tau_ns=800.0)


def accumulate_kw_c1():
    return dict(box=K.C1_BOX, op=K.C1_OP, ro=K.C1_RO, tb=K.C1_TB, n_photons=1_000_000,
                seed=K.C1_SEED, pos=K.C1_POS, tau_ns=K.C1_TAU)


def events_setup():
    ro = readout(SipmGrid(3.2, 8, 8))
    return K.C2_BOX, K.C2_OP, ro, K.C2_TAU


def draws():
    return np.stack([oracle.debug_draws(1234, 1, 64, pid0=p - 1)[0] for p in PIDS])


def records(kw):
    kw = dict(kw)
    box, op, ro = kw.pop("box"), kw.pop("op"), kw.pop("ro")
    r = oracle.transport_records(box, op, ro, **kw)
    return {k: r[k] for k in ("status", "idx", "t", "bounces")}


def accumulate(kw):
    kw = dict(kw)
    box, op, ro, tb = kw.pop("box"), kw.pop("op"), kw.pop("ro"), kw.pop("tb")
    r = oracle.run_photons(box, op, ro, tb, **kw)
    counters = np.array([r[n] for n in oracle.COUNTER_NAMES], np.int64)
    return {"counts": np.asarray(r["counts"]), "first_ns": np.asarray(r["first_ns"]),
            "counters": counters}


def main():
    np.savez_compressed(HERE / "philox_draws.npz", pids=np.array(PIDS, np.uint64),
                        draws=draws())
    np.savez_compressed(HERE / "records_c1.npz", **records(records_kw_c1()))
    np.savez_compressed(HERE / "records_all_processes.npz", **records(records_kw_all()))
    np.savez_compressed(HERE / "accumulate_c1_1e6.npz", **accumulate(accumulate_kw_c1()))
    box, op, ro, tau = events_setup()
    bgo = E.GammaMaterial("xcom_BGO.csv", 7.13, "BGO")
    dep, _ = E.gamma_events(box, bgo, 8500.0, n_events=EV_N, seed=EV_SEED)
    r = oracle.run_events(box, op, ro, dep.xyzt, dep.nph, dep.event, dep.n_events,
                          seed=EV_SEED, tau_ns=tau)
    np.savez_compressed(HERE / "events_bgo.npz", xyzt=dep.xyzt, nph=dep.nph, event=dep.event,
                        n_events=np.int64(dep.n_events), maps=r["maps"], tmin=r["tmin"],
                        npe=r["npe"])
    for p in sorted(HERE.glob("*.npz")):
        print(f"{p.name}: {p.stat().st_size} bytes")


if __name__ == "__main__":
    main()
