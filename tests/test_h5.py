"""HDF5 outputs (crysplight_amd.h5) read back by an independent HDF5 library.

The files follow src/output.jl:10-32 (light.h5) and scripts/shoot_gammas.jl:60-78
(events.h5).  The reader is h5py over the HDF5 C library, run in a separate
interpreter that has them (the image's Anaconda Python); h5dump checks the same files.
Without such a reader the tests skip: nothing here can check the bytes on its own."""
from __future__ import annotations

import hashlib
import json
import subprocess
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from crysplight_amd import h5
from crysplight_amd.output import write_events, write_output

READERS = [Path("/opt/conda/bin/python3.9"), Path("/opt/conda/bin/python3")]
H5DUMP = Path("/opt/conda/bin/h5dump")

DUMP = r"""
import hashlib, json, sys
import h5py, numpy as np
out = {"datasets": {}, "attrs": {}}
with h5py.File(sys.argv[1], "r") as f:
    out["keys"] = list(f.keys())
    for k in f:
        a = f[k][()]
        out["datasets"][k] = [a.dtype.str, list(a.shape), hashlib.sha256(a.tobytes()).hexdigest()]
        out.setdefault("chunks", {})[k] = list(f[k].chunks) if f[k].chunks else None
        if f[k].chunks and a.shape[0] > 1:     # partial reads through the chunk index
            lo, hi = a.shape[0] // 3, a.shape[0] - 1
            part = f[k][lo:hi]
            out.setdefault("slices", {})[k] = bool(np.array_equal(part, a[lo:hi]))
    for k, v in f.attrs.items():
        if isinstance(v, bytes):
            out["attrs"][k] = ["bytes", v.decode()]
        else:
            out["attrs"][k] = [np.asarray(v).dtype.str, v.item()]
print(json.dumps(out))
"""


def _reader():
    for p in READERS:
        if p.exists():
            r = subprocess.run([str(p), "-c", "import h5py"], capture_output=True)
            if r.returncode == 0:
                return p
    return None


READER = _reader()
needs_reader = pytest.mark.skipif(READER is None, reason="no h5py/HDF5 library in this image")


def _dump(path: Path) -> dict:
    r = subprocess.run([str(READER), "-c", DUMP, str(path)], capture_output=True, text=True,
                       cwd=path.parent, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _check(path: Path, datasets: dict, attrs: dict, chunks: dict | None = None):
    d = _dump(path)
    assert d["keys"] == sorted(datasets)
    if chunks is not None:
        assert d.get("chunks", {}) == {k: chunks.get(k) for k in datasets}
        assert all(d.get("slices", {}).values())
    for k, a in datasets.items():
        a = np.asarray(a)
        assert d["datasets"][k] == [a.dtype.str, list(a.shape), _sha(a)], k
    assert set(d["attrs"]) == set(attrs)
    for k, v in attrs.items():
        kind, got = d["attrs"][k]
        if isinstance(v, str):
            assert (kind, got) == ("bytes", v), k
        elif isinstance(v, np.generic):
            assert kind == v.dtype.str and got == v.item(), k
        elif isinstance(v, int):
            assert kind == "<i8" and got == v, k
        else:
            assert kind == "<f8" and got == v, k
    if H5DUMP.exists():
        r = subprocess.run([str(H5DUMP), str(path)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and "error" not in r.stderr.lower(), r.stderr


@needs_reader
def test_every_dtype_shape_and_attribute(tmp_path):
    rng = np.random.default_rng(3)
    ds = {
        "maps": rng.integers(0, 65536, (7, 8, 8)).astype(np.uint16),
        "tmin_ns": np.where(rng.random((7, 8, 8)) < 0.3, np.inf, rng.random((7, 8, 8))).astype(np.float32),
        "counts": rng.integers(0, 2**32, (81, 8, 8), dtype=np.uint64).astype(np.uint32),
        "int_type": rng.integers(-1, 5, 7).astype(np.int8),
        "n_int": rng.integers(0, 9, 7).astype(np.int16),
        "npe": rng.integers(-5, 30000, 7).astype(np.int32),
        "big": rng.integers(-2**62, 2**62, 5).astype(np.int64),
        "f64": rng.standard_normal((3, 4)),
        "u8": rng.integers(0, 256, 9).astype(np.uint8),
        "xyz1_mm": rng.random((7, 3)).astype(np.float32),
        "scalar_like": np.array([1.5], np.float32),
        "empty": np.zeros((0, 8, 8), np.uint16),
    }
    attrs = {"tag": "csitl_crysp", "n_events": 7, "seed": 2**40 + 3, "pde": np.float32(0.3),
             "bin_ns": np.float32(100.0), "nbins": np.int32(80), "sigma_alpha_deg": 1.3,
             "int_type_key": "-1 = no interaction, 0 = direct photoelectric, X>=1 = X Compton scatters",
             "tmin_nohit": "Inf", "negative": -17, "empty_string": ""}
    p = tmp_path / "all.h5"
    h5.write(p, ds, attrs)
    _check(p, ds, attrs)


@needs_reader
def test_many_names_in_one_symbol_table_node(tmp_path):
    # 37 links exceed the default leaf-node capacity of 8: the writer raises K.
    names = [f"d{i:02d}" for i in range(36)] + ["Z_upper_sorts_first"]
    ds = {n: np.full(i + 1, i, np.int32) for i, n in enumerate(names)}
    p = tmp_path / "many.h5"
    h5.write(p, ds)
    _check(p, ds, {})


@needs_reader
def test_no_datasets(tmp_path):
    p = tmp_path / "none.h5"
    h5.write(p, {}, {"tag": "x"})
    _check(p, {}, {"tag": "x"})


def test_rejects_what_it_cannot_write(tmp_path):
    with pytest.raises(ValueError):
        h5.write(tmp_path / "a.h5", {"a/b": np.zeros(2)})
    with pytest.raises(ValueError):
        h5.write(tmp_path / "a.h5", {"c": np.zeros(2, np.complex64)})
    with pytest.raises(ValueError):
        h5.write(tmp_path / "a.h5", {"c": np.zeros(2)}, {"flag": True})


def _cfg(tmp_path, **extra):
    p = tmp_path / "run.toml"
    p.write_text("[run]\n")
    base = dict(cfg_path=p, tag="t", seed=1234,
                cfg={"crystal": {"material": "csitl"}, "wrap": {"material": "teflon"},
                     "surface": {"finish": "backpainted", "sigma_alpha_deg": 1.3}},
                op=SimpleNamespace(pde=np.float32(0.3)))
    base.update(extra)
    return SimpleNamespace(**base)


@needs_reader
def test_events_h5_layout(tmp_path):
    """scripts/shoot_gammas.jl:60-78: dataset names, dtypes, h5py axis order, attributes."""
    rng = np.random.default_rng(5)
    n = 11
    r = {"edep": rng.random(n) * 511, "npe": rng.integers(0, 3000, n), "maps": rng.integers(0, 70000, (n, 8, 8)),
         "tmin": rng.random((n, 8, 8)), "xyz1": rng.random((n, 3)), "e1": rng.random(n),
         "xyz2": rng.random((n, 3)), "e2": rng.random(n), "er": rng.random(n),
         "int_type": rng.integers(-1, 4, n), "n_int": rng.integers(0, 5, n)}
    c = _cfg(tmp_path, n_events=n)
    write_events(c, r, 0.1, tmp_path / "out", verbose=False)
    p = tmp_path / "out" / "events.h5"
    want = {"edep_kev": r["edep"].astype(np.float32), "npe": r["npe"].astype(np.int32),
            "maps": (r["maps"] % 65536).astype(np.uint16), "tmin_ns": r["tmin"].astype(np.float32),
            "xyz1_mm": r["xyz1"].astype(np.float32), "e1_kev": r["e1"].astype(np.float32),
            "xyz2_mm": r["xyz2"].astype(np.float32), "e2_kev": r["e2"].astype(np.float32),
            "er_kev": r["er"].astype(np.float32), "int_type": r["int_type"].astype(np.int8),
            "n_int": r["n_int"].astype(np.int16)}
    attrs = {"tag": "t", "n_events": n, "seed": 1234, "crystal": "csitl", "pde": np.float32(0.3),
             "backend": "hip-gfx950",
             "int_type_key": "-1 = no interaction, 0 = direct photoelectric, X>=1 = X Compton scatters",
             "tmin_nohit": "Inf"}
    _check(p, want, attrs, chunks={k: [n] + list(a.shape[1:]) for k, a in want.items()})


@needs_reader
def test_chunked_datasets(tmp_path):
    """Chunks of 4096 events along the first h5py axis with the full other dimensions,
    the edge chunk partial, more chunks than the library's default B-tree node holds."""
    rng = np.random.default_rng(8)
    ds = {"maps": rng.integers(0, 65536, (4096 * 3 + 17, 8, 8)).astype(np.uint16),
          "tmin_ns": rng.random((4096, 8, 8)).astype(np.float32),
          "xyz1_mm": rng.random((4095, 3)).astype(np.float32),
          "npe": rng.integers(0, 9000, 70 * 64 + 1).astype(np.int32),
          "one": np.array([7], np.int16)}
    p = tmp_path / "chunked.h5"
    h5.write(p, ds, {"n_events": 1}, chunk_events=4096)
    _check(p, ds, {"n_events": 1}, chunks={"maps": [4096, 8, 8], "tmin_ns": [4096, 8, 8],
                                            "xyz1_mm": [4095, 3], "npe": [4096], "one": [1]})
    q = tmp_path / "many_chunks.h5"
    h5.write(q, {"npe": ds["npe"]}, chunk_events=64)       # 71 chunks: K raised above 32
    _check(q, {"npe": ds["npe"]}, {}, chunks={"npe": [64]})


def test_uint16_maps_wrap():
    """src/events.jl:49,81: `maps[ii, jj, ev] += UInt16(1)` wraps at 2^16."""
    from crysplight_amd.output import uint16_maps
    m = np.array([0, 1, 65535, 65536, 65537, 200000], np.int64)
    assert uint16_maps(m).tolist() == [0, 1, 65535, 0, 1, 200000 % 65536]
    # counting one by one in UInt16 gives the same
    acc = np.zeros(1, np.uint16)
    for _ in range(65536 + 5):
        acc += np.uint16(1)
    assert acc[0] == uint16_maps(np.array([65536 + 5]))[0]
    with pytest.raises(ValueError):
        uint16_maps(np.array([-1]))


@needs_reader
def test_light_h5_layout(tmp_path):
    """src/output.jl:10-32: counts (nbins+1, ny, nx) UInt32, first_time_ns (ny, nx), attributes."""
    rng = np.random.default_rng(6)
    nx, ny, nb = 8, 8, 81
    acc = SimpleNamespace(counts=rng.integers(0, 1000, (nx, ny, nb)).astype(np.uint32),
                          first_ns=rng.random((nx, ny)).astype(np.float32), ndet=100, nabs_bulk=5,
                          nabs_wall=6, nabs_sipm=7, ncap=0, nabs_sur=1, nscat=2, sum_bounces_det=900)
    st = _cfg(tmp_path, n_photons=121, outdir=tmp_path / "out",
              tb=SimpleNamespace(bin_ns=np.float32(100.0), nbins=np.int32(nb - 1)))
    write_output(st, acc, 0.5, verbose=False)
    want = {"counts": acc.counts.transpose(2, 1, 0), "first_time_ns": acc.first_ns.T}
    attrs = {"tag": "t", "n_photons": 121, "seed": 1234, "bin_ns": np.float32(100.0),
             "nbins": np.int32(nb - 1), "crystal": "csitl", "wrap": "teflon", "pde": np.float32(0.3),
             "ndet": 100, "nabs_bulk": 5, "nabs_wall": 6, "nabs_sipm": 7, "ncap": 0,
             "nabs_surround": 1, "nscat": 2, "finish": "backpainted", "sigma_alpha_deg": 1.3}
    _check(tmp_path / "out" / "light.h5", want, attrs)
    # h5py's (nbins+1, ny, nx) element [b, j, i] is the reference's counts[i, j, b].
    d = _dump(tmp_path / "out" / "light.h5")
    assert d["datasets"]["counts"][1] == [nb, ny, nx]
