"""Run configuration: TOML files name the materials and the run, as in the reference.

load_material      <- load_material (src/materials.jl:8-12)
surface_has_gap    <- surface_has_gap (src/config.jl:147-150)
load_setup         <- load_setup (src/config.jl:152-194), plus an optional tau_rise_ns
run_from_config    <- run_from_config (src/config.jl:197-204), transport on the MI355X
load_gamma_campaign / run_gamma_campaign <- scripts/shoot_gammas.jl:19-91

Material files are looked up in data_dir (default: crysplight_amd/data/materials, then the
directory of the run file's ../data, which is where the reference keeps them).
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

try:
    import tomllib as _toml  # Python >= 3.11
except ImportError:  # pragma: no cover - this image has 3.10 with tomli
    import tomli as _toml

from .params import Box, OpticalParams, SipmGrid, TimeBinning, readout

HERE = Path(__file__).resolve().parent
DATA_DIR = HERE / "data" / "materials"


def _candidates(name: str, data_dir=None, run_path=None):
    dirs = []
    if data_dir:
        dirs.append(Path(data_dir))
    dirs.append(DATA_DIR)
    if run_path is not None:
        dirs.append(Path(run_path).resolve().parent.parent / "data")
    return [d / f"{name}.toml" for d in dirs]


def load_material(name: str, data_dir=None, run_path=None) -> dict:
    """data/<name>.toml as a dict; raises if no candidate file exists."""
    for p in _candidates(name, data_dir, run_path):
        if p.is_file():
            with open(p, "rb") as f:
                return _toml.load(f)
    raise FileNotFoundError(f"material file not found: {name}.toml (looked in "
                            + ", ".join(str(p.parent) for p in _candidates(name, data_dir, run_path)) + ")")


def surface_has_gap(finish: str) -> bool:
    """True for an air-gap back-painted wrap; only backpainted/frontpainted are valid."""
    if finish not in ("backpainted", "frontpainted"):
        raise ValueError(f"unknown finish: {finish}")
    return finish == "backpainted"


def deg2rad32(deg) -> float:
    return float(np.float32(float(deg) * (math.pi / 180.0)))


@dataclass
class RunSetup:
    """src/config.jl:130-144"""
    tag: str
    cfg: dict
    box: Box
    op: OpticalParams
    grid: SipmGrid
    tb: TimeBinning
    pos: tuple
    tau_ns: float
    n_photons: int
    seed: int
    max_bounces: int
    outdir: Path
    cfg_path: Path
    tau_rise_ns: float = 0.0


def _optics(cfg, crystal, wrap, sipm):
    surf = cfg.get("surface", {})
    finish = surf.get("finish", "backpainted")
    pde = float(cfg["sipm"].get("pde_override", sipm["pde"]))
    op = OpticalParams(float(crystal["n"]), float(cfg["sipm"]["coupling_n"]),
                       float(crystal["abs_length_mm"]), float(crystal.get("rayleigh_mm", math.inf)),
                       float(wrap["reflectivity"]), wrap["model"] == "specular",
                       surface_has_gap(finish), deg2rad32(surf.get("sigma_alpha_deg", 0.0)), pde)
    grid = SipmGrid(float(sipm["pitch_mm"]), int(sipm["nx"]), int(sipm["ny"]))
    b = cfg["binning"]
    bin_ns = float(np.float32(b["bin_ns"]))
    tb = TimeBinning(bin_ns, int(round(float(b["window_ns"]) / bin_ns)))
    return op, grid, tb


def load_setup(cfg_path, data_dir=None) -> RunSetup:
    """src/config.jl:152-194: TOML -> Box, OpticalParams, SipmGrid, TimeBinning, source."""
    cfg_path = Path(cfg_path)
    with open(cfg_path, "rb") as f:
        cfg = _toml.load(f)
    run = cfg["run"]
    tag = run.get("tag", cfg_path.stem)
    mat = lambda name: load_material(name, data_dir, cfg_path)
    crystal = mat(cfg["crystal"]["material"])
    wrap = mat(cfg["wrap"]["material"])
    sipm = mat(cfg["sipm"]["material"])
    sz = [float(v) for v in cfg["crystal"]["size_mm"]]
    op, grid, tb = _optics(cfg, crystal, wrap, sipm)
    src = cfg["source"]
    if src["type"] != "point":
        raise ValueError("only the point source is supported")
    p = tuple(float(np.float32(v)) for v in src["position_mm"])
    delta = src["time_profile"] == "delta"
    tau = 0.0 if delta else float(crystal["tau_ns"])
    tau_rise = 0.0 if delta else float(crystal.get("tau_rise_ns", 0.0))
    outdir = cfg_path.resolve().parent.parent / cfg.get("output", {}).get("dir", "output") / tag
    return RunSetup(tag, cfg, Box(sz), op, grid, tb, p, tau, int(run["n_photons"]),
                    int(run["seed"]), int(run.get("max_bounces", 100_000)), outdir,
                    cfg_path.resolve(), tau_rise)


def run_from_config(cfg_path, data_dir=None, ctx=None, write=True, verbose=True) -> dict:
    """src/config.jl:197-204: transport, write output/<tag>/, print and return the summary."""
    from .transport import run_photons
    from .output import write_output
    st = load_setup(cfg_path, data_dir)
    t0 = time.perf_counter()
    acc = run_photons(st.box, st.op, readout(st.grid), st.tb, n_photons=st.n_photons,
                      seed=st.seed, pos=st.pos, tau_ns=st.tau_ns, tau_rise_ns=st.tau_rise_ns,
                      max_bounces=st.max_bounces, ctx=ctx)
    elapsed = time.perf_counter() - t0
    return write_output(st, acc, elapsed, write=write, verbose=verbose)


@dataclass
class GammaCampaign:
    """The fields scripts/shoot_gammas.jl:20-47 reads from a gamma run file."""
    tag: str
    cfg: dict
    box: Box
    op: OpticalParams
    grid: SipmGrid
    tb: TimeBinning
    n_events: int
    seed: int
    batch_photons: int
    yield_per_mev: float
    tau_ns: float
    tau_rise_ns: float
    density: float
    xcom: str
    crystal_name: str
    cfg_path: Path


def load_gamma_campaign(cfg_path, data_dir=None) -> GammaCampaign:
    cfg_path = Path(cfg_path)
    with open(cfg_path, "rb") as f:
        cfg = _toml.load(f)
    camp = cfg["campaign"]
    mat = lambda name: load_material(name, data_dir, cfg_path)
    crystal = mat(cfg["crystal"]["material"])
    wrap = mat(cfg["wrap"]["material"])
    sipm = mat(cfg["sipm"]["material"])
    op, grid, tb = _optics(cfg, crystal, wrap, sipm)
    sz = [float(v) for v in cfg["crystal"]["size_mm"]]
    return GammaCampaign(camp["tag"], cfg, Box(sz), op, grid, tb, int(camp["n_events"]),
                         int(camp["seed"]), int(camp.get("batch_photons", 2_000_000)),
                         float(crystal["yield_per_mev"]), float(crystal["tau_ns"]),
                         float(crystal.get("tau_rise_ns", 0.0)), float(cfg["gamma"]["density"]),
                         cfg["gamma"]["xcom"], crystal.get("name", cfg["crystal"]["material"]),
                         cfg_path.resolve())


def run_gamma_campaign(cfg_path, data_dir=None, ctx=None, write=True, verbose=True) -> dict:
    """scripts/shoot_gammas.jl:19-91: gamma events -> optical transport on the MI355X ->
    output/<tag>/events (the events.h5 datasets) and a photopeak summary."""
    from .events import GammaMaterial, run_events
    from .output import write_events
    c = load_gamma_campaign(cfg_path, data_dir)
    xcom = Path(c.xcom)
    if not xcom.is_absolute():
        for cand in (HERE / "data" / xcom, c.cfg_path.parent.parent / "data" / xcom):
            if cand.exists():
                xcom = cand
                break
    mat = GammaMaterial(xcom, c.density, c.crystal_name)
    t0 = time.perf_counter()
    r = run_events(c.box, c.op, readout(c.grid), c.tb, mat, c.yield_per_mev,
                   n_events=c.n_events, seed=c.seed, tau_ns=c.tau_ns, tau_rise_ns=c.tau_rise_ns,
                   batch_photons=c.batch_photons, ctx=ctx)
    elapsed = time.perf_counter() - t0
    outdir = c.cfg_path.parent.parent / "output" / c.tag
    return write_events(c, r, elapsed, outdir, write=write, verbose=verbose)
