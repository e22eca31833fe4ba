"""The gamma driver end to end on the GPU (scripts/shoot_gammas.jl): a small campaign
through `python -m crysplight_amd.shoot_gammas`'s entry point, output/<tag>/events.h5
read back by the HDF5 C library (h5py) and compared with the same campaign's arrays
recomputed in memory; the event datasets chunked at 4096 events along the event axis
(scripts/shoot_gammas.jl:11-17), maps as UInt16 (src/events.jl:49,81)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.test_h5 import READER

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(READER is None, reason="no h5py/HDF5 library in this image")]

HERE = Path(__file__).resolve().parent.parent / "crysplight_amd"

READ = r"""
import json, sys
import h5py, numpy as np
with h5py.File(sys.argv[1], "r") as f:
    np.savez(sys.argv[2], **{k: f[k][()] for k in f})
    meta = {"chunks": {k: list(f[k].chunks or []) for k in f},
            "attrs": {k: (v.decode() if isinstance(v, bytes) else np.asarray(v).item())
                      for k, v in f.attrs.items()}}
print(json.dumps(meta))
"""


def test_shoot_gammas_writes_events_h5(tmp_path, ctx):
    from crysplight_amd import shoot_gammas
    from crysplight_amd.config import load_gamma_campaign
    from crysplight_amd.events import GammaMaterial, run_events
    from crysplight_amd.params import readout

    n = 4500                                    # two chunks, the second partial
    src = (HERE / "runs" / "gammas_bgo_c2.toml").read_text()
    src = src.replace('tag = "gammas_bgo_c2"', 'tag = "h5_small"').replace("n_events = 10000", f"n_events = {n}")
    runs = tmp_path / "runs"
    runs.mkdir()
    cfg = runs / "h5_small.toml"
    cfg.write_text(src)
    print(f"campaign {cfg} n_events={n}", flush=True)
    assert shoot_gammas.main([str(cfg)]) == 0
    out = tmp_path / "output" / "h5_small"
    assert (out / "h5_small.toml").read_text() == src

    r = subprocess.run([str(READER), "-c", READ, str(out / "events.h5"), str(tmp_path / "back.npz")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    meta = json.loads(r.stdout)
    back = dict(np.load(tmp_path / "back.npz"))

    c = load_gamma_campaign(cfg)
    mat = GammaMaterial(HERE / "data" / "xcom_BGO.csv", c.density, c.crystal_name)
    want = run_events(c.box, c.op, readout(c.grid), c.tb, mat, c.yield_per_mev, n_events=n,
                      seed=c.seed, tau_ns=c.tau_ns, tau_rise_ns=c.tau_rise_ns, ctx=ctx)
    dt = {"edep_kev": ("edep", np.float32), "npe": ("npe", np.int32), "tmin_ns": ("tmin", np.float32),
          "xyz1_mm": ("xyz1", np.float32), "e1_kev": ("e1", np.float32), "xyz2_mm": ("xyz2", np.float32),
          "e2_kev": ("e2", np.float32), "er_kev": ("er", np.float32), "int_type": ("int_type", np.int8),
          "n_int": ("n_int", np.int16)}
    assert sorted(back) == sorted(list(dt) + ["maps"])
    for k, (src_key, t) in dt.items():
        a = np.asarray(want[src_key]).astype(t)
        assert back[k].dtype == t and np.array_equal(back[k], a), k
    maps = np.asarray(want["maps"])
    assert back["maps"].dtype == np.uint16 and back["maps"].shape == (n, c.grid.ny, c.grid.nx)
    assert np.array_equal(back["maps"], (maps % 65536).astype(np.uint16))
    # the pe counts of one event add up to its npe (no event here comes near 2^16 per SiPM)
    assert maps.max() < 65536
    assert np.array_equal(back["maps"].reshape(n, -1).sum(1, dtype=np.int64), back["npe"])
    for k, a in back.items():
        assert meta["chunks"][k] == [min(4096, n)] + list(a.shape[1:]), k
    attrs = meta["attrs"]
    assert attrs["n_events"] == n and attrs["seed"] == c.seed and attrs["tag"] == "h5_small"
    assert attrs["crystal"] == "bgo_c2" and attrs["tmin_nohit"] == "Inf"
