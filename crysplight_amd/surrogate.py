"""Learned light-response surrogate: an MLP from (deposit position, energy) to the mean
SiPM count and arrival-time maps, trained on Monte Carlo targets of this library's own
transport, and Poisson sampling of event maps from it.  No reference counterpart.

Model.  For a deposit of energy E at x, the photon count is Poisson(Y E / 1000) and each
photon lands in SiPM c with probability p_c(x), so the SiPM counts are independent
Poisson(Y E / 1000 * p_c(x)) (Poisson thinning): the mean count map determines the
count distribution exactly.  The network (4 -> 256 -> 256 -> 256 -> 128, GELU) predicts
the per-photon map as log-rates log p_c (outputs 0..63, trained with a Poisson negative
log-likelihood of the observed counts, which is unbiased and weighs relative errors
evenly) and the mean arrival time of a detected photon after the deposit (outputs
64..127, standardized, squared error weighted by the counts); the mean count map is
Y E / 1000 * p.  Energy enters as the fourth input.

Computed light.  Near the readout face the maps are sharp: light that reaches the face
straight from the deposit lands within d tan(theta_c) of the foot point, so a cell's share
jumps across cell boundaries on the scale of the depth d.  That part is computed, not
learned: the direct light (Fresnel transmission into the coupling, bulk attenuation,
PDE, integrated over each cell) and, behind an air gap, its mirror images in the side
walls (total internal reflection), weighted by a smoothstep that is 1 within 2 mm of the
readout face and 0 beyond 3 mm.  The per-photon rate of cell c is e^{y_c} + D_c; the
network learns e^{y_c}, which is smooth (training uses D as a rate offset in the Poisson
loss).

Inference.  `predict` and `sample` run the whole chain on the device: input features and
symmetry folding, the computed light, the network (one fused BF16 MFMA kernel) and, for
events, the per-event Poisson sampling.  `bake` and `sample_events` keep the older path
that interpolates a baked grid of the network.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .events import Deposits, transport_events
from .params import Box, Emission, OpticalParams, SipmGrid, readout


def _bind():
    lib = _lib.load()
    if getattr(lib, "_mlp_bound", False):
        return lib
    vp = C.c_void_p
    lib.crysp_mlp_create.restype = vp
    lib.crysp_mlp_create.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]
    lib.crysp_mlp_destroy.argtypes = [vp]
    lib.crysp_mlp_last_error.restype = C.c_char_p
    lib.crysp_mlp_num_params.restype = C.c_int64
    lib.crysp_mlp_num_params.argtypes = [vp]
    lib.crysp_mlp_flops_per_row.restype = C.c_double
    lib.crysp_mlp_flops_per_row.argtypes = [vp]
    lib.crysp_mlp_get_params.argtypes = [vp, vp]
    lib.crysp_mlp_set_params.argtypes = [vp, vp]
    lib.crysp_mlp_set_input_norm.argtypes = [vp, vp, vp]
    lib.crysp_mlp_set_poisson_outputs.argtypes = [vp, C.c_int32]
    lib.crysp_mlp_forward.argtypes = [vp, vp, C.c_int64, vp]
    lib.crysp_mlp_gradient.argtypes = [vp, vp, vp, vp, vp, C.c_int32, vp, vp]
    lib.crysp_mlp_train_batches.argtypes = [vp, vp, vp, vp, vp, C.c_int64, C.c_int32, C.c_float,
                                            C.c_float, C.c_float, C.c_float, C.c_float,
                                            C.c_float, vp]
    lib.crysp_surrogate_create.restype = vp
    lib.crysp_surrogate_create.argtypes = [vp, C.POINTER(Box), C.POINTER(OpticalParams), vp,
                                           C.POINTER(SurrogateConfig)]
    lib.crysp_surrogate_destroy.argtypes = [vp]
    lib.crysp_surrogate_direct_light.argtypes = [vp, vp, C.c_int64, vp]
    lib.crysp_surrogate_predict.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
    lib.crysp_surrogate_sample.argtypes = [vp, vp, vp, vp, vp, C.c_int64, C.c_int64, C.c_double,
                                           C.c_uint64, vp, vp, vp]
    lib.crysp_surrogate_sample_events.argtypes = [C.c_int, vp, C.c_int32, C.c_int32, C.c_int32,
                                                  C.c_int32, C.POINTER(Box), vp, vp, vp, vp,
                                                  C.c_int64, C.c_int64, C.c_double, C.c_uint64,
                                                  vp, vp, vp]
    lib.crysp_debug_gemm.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_int32, vp, C.c_int32, vp,
                                     C.c_int32, vp]
    lib.crysp_light_maps.argtypes = [vp, C.POINTER(Box), C.POINTER(OpticalParams), vp, vp, vp,
                                     C.c_int64, C.c_uint64, C.c_uint64, C.c_int32,
                                     C.POINTER(Emission), vp, vp]
    lib._mlp_bound = True
    return lib


class SurrogateConfig(C.Structure):
    """crysp_surrogate_config (include/crysplight.h)."""
    _fields_ = [("fold", C.c_uint8), ("direct", C.c_uint8), ("images", C.c_uint8), ("pad_", C.c_uint8),
                ("log_depth_mm", C.c_float), ("t_shift", C.c_float), ("t_scale", C.c_float)]


def _mcheck(rc, what):
    if rc != 0:
        raise _lib.CryspError(f"{what} failed ({rc}): " + _bind().crysp_mlp_last_error().decode())


def _ptr(a):
    return None if a is None else C.c_void_p(_lib.ptr(a))


class MLP:
    """The network on the device: BF16 MFMA forward/backward, FP32 master weights, Adam."""

    def __init__(self, n_in=4, hidden=256, n_hidden_layers=3, n_out=128, seed=1, device=0):
        lib = _bind()
        h = lib.crysp_mlp_create(device, n_in, hidden, n_hidden_layers, n_out, seed)
        if not h:
            raise _lib.CryspError("crysp_mlp_create: " + lib.crysp_mlp_last_error().decode())
        self._h = h
        self.n_in, self.hidden, self.n_hidden_layers, self.n_out = n_in, hidden, n_hidden_layers, n_out
        self.shapes = [(b, a) for a, b in zip([n_in] + [hidden] * n_hidden_layers,
                                              [hidden] * n_hidden_layers + [n_out])]
        self.shift = np.zeros(n_in, np.float32)
        self.scale = np.ones(n_in, np.float32)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _bind().crysp_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return int(_bind().crysp_mlp_num_params(self._h))

    @property
    def flops_per_row(self) -> float:
        return float(_bind().crysp_mlp_flops_per_row(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _mcheck(_bind().crysp_mlp_get_params(self._h, p.ctypes.data), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _mcheck(_bind().crysp_mlp_set_params(self._h, p.ctypes.data), "set_params")

    def set_input_norm(self, shift, scale):
        self.shift = np.ascontiguousarray(shift, np.float32)
        self.scale = np.ascontiguousarray(scale, np.float32)
        _mcheck(_bind().crysp_mlp_set_input_norm(self._h, self.shift.ctypes.data,
                                                 self.scale.ctypes.data), "set_input_norm")

    def set_poisson_outputs(self, n_pois: int):
        _mcheck(_bind().crysp_mlp_set_poisson_outputs(self._h, int(n_pois)), "set_poisson_outputs")
        self.n_pois = int(n_pois)

    def forward(self, X, out=None):
        n = int(X.shape[0])
        if out is None:
            Xh = np.ascontiguousarray(X, np.float32)
            out = np.zeros((n, self.n_out), np.float32)
            _mcheck(_bind().crysp_mlp_forward(self._h, Xh.ctypes.data, n, out.ctypes.data), "forward")
            return out
        _lib.check_array(X, "X", np.float32, (n, self.n_in))
        _lib.check_array(out, "out", np.float32, (n, self.n_out))
        _lib.ready(X, out)
        _mcheck(_bind().crysp_mlp_forward(self._h, _ptr(X), n, _ptr(out)), "forward")
        return out

    def gradient(self, X, T, W=None, R0=None):
        """Loss and gradient of one batch (host arrays); R0: optional rate offsets
        [rows x n_pois] (the Poisson rate is e^y + R0)."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        W = None if W is None else np.ascontiguousarray(W, np.float32)
        R0 = None if R0 is None else np.ascontiguousarray(R0, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        _mcheck(_bind().crysp_mlp_gradient(self._h, X.ctypes.data, T.ctypes.data,
                                           None if W is None else W.ctypes.data,
                                           None if R0 is None else R0.ctypes.data, X.shape[0],
                                           g.ctypes.data, loss.ctypes.data), "gradient")
        return float(loss[0]), g

    def train_batches(self, X, T, W, n_batches, batch, lr=1e-3, betas=(0.9, 0.999), eps=1e-8,
                      weight_decay=0.0, mean_row_weight=1.0, R0=None):
        """Adam steps over consecutive rows of device tensors X, T, W (and the rate offsets
        R0, rows of n_pois, when given)."""
        losses = np.zeros(max(1, n_batches), np.float32)
        rows = int(n_batches) * int(batch)
        n_pois = getattr(self, "n_pois", 0)
        for a, name, cols in ((X, "X", self.n_in), (T, "T", self.n_out), (W, "W", self.n_out),
                              (R0, "R0", n_pois)):
            if a is None and name in ("W", "R0"):
                continue
            _lib.check_array(a, name, np.float32)
            if a.ndim != 2 or a.shape[1] != cols or a.shape[0] < rows:
                raise ValueError(f"{name}: shape {tuple(a.shape)}, expected at least "
                                 f"({rows}, {cols})")
        _lib.ready(*(a for a in (X, T, W, R0) if a is not None))
        _mcheck(_bind().crysp_mlp_train_batches(self._h, _ptr(X), _ptr(T), _ptr(W), _ptr(R0),
                                                int(n_batches),
                                                int(batch), lr, betas[0], betas[1], eps,
                                                weight_decay, mean_row_weight,
                                                losses.ctypes.data), "train_batches")
        return losses[:n_batches]


def debug_gemm(A, B, a_t=False, b_t=False, device=0):
    """C = op(A) op(B) through the device GEMM; A, B given as float arrays (rounded to
    bf16).  Shapes: op(A) [M x K], op(B) [K x N]."""
    import torch
    a = torch.from_numpy(np.ascontiguousarray(A, np.float32)).to(torch.bfloat16).view(torch.int16).numpy()
    b = torch.from_numpy(np.ascontiguousarray(B, np.float32)).to(torch.bfloat16).view(torch.int16).numpy()
    M = A.shape[1] if a_t else A.shape[0]
    K = A.shape[0] if a_t else A.shape[1]
    N = B.shape[0] if b_t else B.shape[1]
    Cm = np.zeros((M, N), np.float32)
    _mcheck(_bind().crysp_debug_gemm(device, M, N, K, a.ctypes.data, int(a_t), b.ctypes.data,
                                     int(b_t), Cm.ctypes.data), "debug_gemm")
    return Cm


def mc_light_maps(box: Box, op: OpticalParams, ro, xyz, nph, *, seed: int, tau_ns: float,
                  tau_rise_ns: float = 0.0, pid0: int = 0, max_bounces: int = 100_000, ctx=None):
    """Monte Carlo light-response maps of point sources: counts [n, nx*ny] (int32) and
    the exact sums of arrival times [n, nx*ny] in ns (binary64 from integer sums)."""
    ctx = ctx or _lib.default_context()
    ro = readout(ro) if isinstance(ro, SipmGrid) else ro
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    nph = np.ascontiguousarray(np.broadcast_to(np.asarray(nph, np.int32), (n,)))
    nxy = ro.grid.nx * ro.grid.ny
    counts = np.zeros((n, nxy), np.int32)
    tsum = np.zeros((n, nxy), np.uint64)
    em = Emission(tau_ns, tau_rise_ns)
    rc = _bind().crysp_light_maps(ctx.handle, C.byref(box), C.byref(op), C.c_void_p(C.addressof(ro)),
                                  xyz.ctypes.data, nph.ctypes.data, n, int(seed), int(pid0),
                                  int(max_bounces), C.byref(em), counts.ctypes.data,
                                  tsum.ctypes.data)
    _lib.check(rc, "crysp_light_maps")
    return counts, tsum.astype(np.float64) / 1024.0


# ---------------------------------------------------------------- symmetry ----
# A square crystal face read by a square SiPM grid that spans it exactly is symmetric
# under x -> L-x, y -> L-y and x <-> y (8 operations).  SiPM (i, j) (0-based, i along x)
# is cell i + n*j.  A position folds into the fundamental triangle 0 <= x <= y <= L/2
# (octant code: bit 0 x mirrored, bit 1 y mirrored, bit 2 swapped); the network learns
# maps of folded positions only, in the folded cell order.

def fold_xy(x, y, L):
    """(x, y) -> folded (xc, yc) and the octant code of each point."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    h = 0.5 * L
    mx, my = x > h, y > h
    xf = np.where(mx, L - x, x)
    yf = np.where(my, L - y, y)
    sw = xf > yf
    xc = np.where(sw, yf, xf)
    yc = np.where(sw, xf, yf)
    code = mx.astype(np.int64) | (my.astype(np.int64) << 1) | (sw.astype(np.int64) << 2)
    return xc, yc, code


def fold_perms(n: int) -> np.ndarray:
    """[8, n*n]: for octant code o, perms[o][c'] is the actual cell of folded cell c'."""
    a1, b1 = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")   # a1 along x
    a1, b1 = a1.ravel(), b1.ravel()                                   # c' = a1 + n*b1
    out = np.empty((8, n * n), np.int64)
    for o in range(8):
        mx, my, sw = o & 1, (o >> 1) & 1, (o >> 2) & 1
        a, b = (b1, a1) if sw else (a1, b1)
        i = n - 1 - a if mx else a
        j = n - 1 - b if my else b
        out[o] = i + n * j
    return out


def to_folded(maps, code, perms):
    """Actual-order maps [k, n*n] -> folded order."""
    return np.take_along_axis(np.asarray(maps), perms[code], axis=1)


def from_folded(maps_f, code, perms):
    """Folded-order maps [k, n*n] -> actual order."""
    out = np.empty_like(maps_f)
    np.put_along_axis(out, perms[code], maps_f, axis=1)
    return out


def folds(box: Box, ro) -> bool:
    """True when the optics have the 8-fold symmetry the folding uses."""
    g = ro.grid
    L0, L1 = float(box.L[0]), float(box.L[1])
    return (abs(L0 - L1) <= 1e-6 * L0 and g.nx == g.ny and not ro.disc
            and abs(g.pitch_mm * g.nx - L0) <= 1e-5 * L0)


@dataclass
class TrainingSet:
    X: np.ndarray        # (n, 4) x, y, z [mm], E [keV]
    counts: np.ndarray   # (n, 64) int32
    tsum: np.ndarray     # (n, 64) float64 ns
    nph: int

    @property
    def p(self):
        return self.counts / float(self.nph)

    @property
    def tmean(self):
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(self.counts > 0, self.tsum / np.maximum(self.counts, 1), np.nan)


class LightSurrogate:
    """Surrogate of one crystal/optics configuration."""

    def __init__(self, box: Box, op: OpticalParams, ro, *, tau_ns: float,
                 tau_rise_ns: float = 0.0, e_range=(10.0, 511.0), hidden=256, n_hidden_layers=3,
                 seed=1, device=0, log_depth_mm: float = 0.0, fold: bool = True,
                 direct: bool = True, images: bool = True):
        self.box, self.op = box, op
        self.ro = readout(ro) if isinstance(ro, SipmGrid) else ro
        self.ncell = self.ro.grid.nx * self.ro.grid.ny
        # fold positions into 1/8 of the face (and permute maps) when the optics allow it
        self.fold = bool(fold) and folds(box, self.ro)
        self.perms = fold_perms(self.ro.grid.nx) if self.fold else None
        self.tau_ns, self.tau_rise_ns = float(tau_ns), float(tau_rise_ns)
        self.e_range = tuple(float(v) for v in e_range)
        self.net = MLP(4, hidden, n_hidden_layers, 2 * self.ncell, seed=seed, device=device)
        # log_depth_mm > 0: the depth input is -log(Lz - z + log_depth_mm), which stretches
        # the region next to the readout face where the maps sharpen like 1 / distance^2
        self.log_depth_mm = float(log_depth_mm)
        L = np.asarray(box.L, np.float32)
        lo = np.array([0, 0, 0, self.e_range[0]], np.float32)
        hi = np.array([L[0], L[1], L[2], self.e_range[1]], np.float32)
        if self.fold:
            hi[0] = hi[1] = 0.5 * L[0]
        if self.log_depth_mm > 0:
            lo[2], hi[2] = self._depth(np.float32(0.0)), self._depth(L[2])
        self.net.set_input_norm((lo + hi) / 2, 2.0 / (hi - lo))
        self.net.set_poisson_outputs(self.ncell)
        self.pois_w = 1.0
        self.t_shift = self.tau_ns + self.tau_rise_ns
        self.t_scale = max(self.tau_ns + self.tau_rise_ns, 1.0)
        self.device = device
        # the light that reaches the readout face straight from the deposit is computed, not
        # learned (sharp near the face: it lands within d tan(theta_c) of the foot point)
        self.direct = bool(direct) and not self.ro.disc and not self.ro.two_sided
        # and, behind an air gap, its mirror images in the side walls near the readout face
        self.images = self.direct and bool(images) and bool(op.air_gap)
        cfg = SurrogateConfig(int(self.fold), int(self.direct), int(self.images), 0,
                              self.log_depth_mm, self.t_shift, self.t_scale)
        h = _bind().crysp_surrogate_create(self.net._h, C.byref(box), C.byref(op),
                                           C.c_void_p(C.addressof(self.ro)), C.byref(cfg))
        if not h:
            raise _lib.CryspError("crysp_surrogate_create: " + _bind().crysp_mlp_last_error().decode())
        self._sh = h

    def __del__(self):
        try:
            if getattr(self, "_sh", None):
                _bind().crysp_surrogate_destroy(self._sh)
                self._sh = None
        except Exception:
            pass

    def direct_light(self, xyz, out=None):
        """The computed direct-light part D [n, ncell] (cell order i + nx j) at positions
        xyz [n, 3] as given (no folding); a device tensor `out` receives it in place."""
        xyz = xyz if hasattr(xyz, "data_ptr") else np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = int(xyz.shape[0])
        _lib.check_array(xyz, "xyz", np.float32, (n, 3), device=self.device)
        D = np.zeros((n, self.ncell), np.float32) if out is None else out
        _lib.check_array(D, "D", np.float32, (n, self.ncell), device=self.device)
        _lib.ready(xyz, D)
        _mcheck(_bind().crysp_surrogate_direct_light(self._sh, _ptr(xyz), n, _ptr(D)),
                "crysp_surrogate_direct_light")
        return D

    def _depth(self, z):
        return -np.log(np.float32(self.box.L[2]) - z + np.float32(self.log_depth_mm))

    def features(self, xyz, e_kev):
        """Network inputs (x, y, depth, E) for positions [n, 3] and energies [n]."""
        X = np.column_stack([np.asarray(xyz, np.float32).reshape(-1, 3),
                             np.asarray(e_kev, np.float32).reshape(-1)]).astype(np.float32)
        if self.fold:
            xc, yc, _ = fold_xy(X[:, 0], X[:, 1], float(self.box.L[0]))
            X[:, 0], X[:, 1] = xc, yc
        if self.log_depth_mm > 0:
            X[:, 2] = self._depth(X[:, 2])
        return X

    # ------------------------------------------------------------------ data ----
    def generate(self, n_samples: int, photons_per_sample: int, *, seed: int, pid0: int = 0,
                 near_readout_frac: float = 0.0, near_readout_mm: float = 2.0,
                 ctx=None) -> TrainingSet:
        """Positions uniform in the crystal and energies uniform in e_range (Philox
        domain 2), each traced at photons_per_sample photons.  near_readout_frac of the
        samples are drawn uniformly in the slab within near_readout_mm of the readout face
        instead, where the maps change fastest."""
        from .events import synthetic_deposits
        n_near = int(round(near_readout_frac * n_samples))
        d = synthetic_deposits(self.box, n_events=n_samples - n_near, seed=seed,
                               e_kev=self.e_range, n_fixed=photons_per_sample)
        X = np.column_stack([d.xyzt[:, :3], d.e_kev]).astype(np.float32)
        if n_near:
            L = self.box.L
            dn = synthetic_deposits(self.box, n_events=n_near, seed=seed + 1, e_kev=self.e_range,
                                    n_fixed=photons_per_sample,
                                    lo=(0.0, 0.0, max(0.0, L[2] - near_readout_mm)),
                                    hi=(L[0], L[1], L[2]))
            X = np.concatenate([X, np.column_stack([dn.xyzt[:, :3], dn.e_kev]).astype(np.float32)])
        counts, tsum = mc_light_maps(self.box, self.op, self.ro, X[:, :3], photons_per_sample,
                                     seed=seed, tau_ns=self.tau_ns, tau_rise_ns=self.tau_rise_ns,
                                     pid0=pid0, ctx=ctx)
        return TrainingSet(X, counts, tsum, int(photons_per_sample))

    def targets(self, ts: TrainingSet, device=None):
        """Observed per-photon rates (Poisson outputs, weight 1 / mean rate so gradients
        read as relative errors) and standardized mean times (weight c / (c + 20)), as
        float32 [n, 2 * ncell] tensors on the device (numpy arrays with device=None), and
        the direct-light rates R0 [n, ncell] at the same (folded) positions (None without
        the direct term): the network learns the rest of each rate."""
        import torch
        dev = torch.device("cpu") if device is None else device
        cnt = torch.from_numpy(np.ascontiguousarray(ts.counts)).to(dev)
        tsum = torch.from_numpy(np.ascontiguousarray(ts.tsum)).to(dev)
        xyz_f = np.ascontiguousarray(ts.X[:, :3], np.float32).copy()
        if self.fold:
            xc, yc, code = fold_xy(ts.X[:, 0], ts.X[:, 1], float(self.box.L[0]))
            xyz_f[:, 0], xyz_f[:, 1] = xc, yc
            idx = torch.from_numpy(self.perms).to(dev)[torch.from_numpy(code.astype(np.int64)).to(dev)]
            cnt, tsum = torch.gather(cnt, 1, idx), torch.gather(tsum, 1, idx)
        p = cnt.double() / float(ts.nph)
        tm = tsum / cnt.double()                     # 0 / 0 where a cell saw nothing
        n = ts.X.shape[0]
        T = torch.empty((n, 2 * self.ncell), dtype=torch.float32, device=dev)
        T[:, :self.ncell] = p
        T[:, self.ncell:] = torch.nan_to_num((tm - self.t_shift) / self.t_scale, nan=0.0)
        W = torch.empty_like(T)
        self.pois_w = float(1.0 / max(float(p.mean()), 1e-12))
        W[:, :self.ncell] = self.pois_w
        c = cnt.float()
        W[:, self.ncell:] = 0.25 * c / (c + 20.0)
        R0 = None
        if self.direct:
            if device is None:
                R0 = self.direct_light(xyz_f)
            else:
                R0 = torch.empty((n, self.ncell), dtype=torch.float32, device=dev)
                self.direct_light(torch.from_numpy(xyz_f).to(dev), out=R0)
        if device is None:
            return T.numpy(), W.numpy(), R0
        return T, W, R0

    def init_output_bias(self, T, R0=None):
        """Start the log-rate outputs at the log of each cell's mean learned rate (T:
        targets, R0: the direct-light rates or None; numpy or tensors)."""
        par = self.net.get_params()
        n_out = 2 * self.ncell
        b = par.size - n_out
        rest = T[:, :self.ncell] if R0 is None else T[:, :self.ncell] - R0
        mean = rest.mean(0)
        mean = mean.cpu().numpy() if hasattr(mean, "cpu") else np.asarray(mean)
        par[b:b + self.ncell] = np.log(np.maximum(mean, 1e-9))
        par[b + self.ncell:] = 0.0
        self.net.set_params(par)

    # ------------------------------------------------------------------ train ---
    def fit(self, ts: TrainingSet, *, epochs: int = 10, batch: int = 4096, lr: float = 2e-3,
            lr_final: float = 1e-4, seed: int = 0, verbose: bool = False):
        """Adam over shuffled mini-batches, learning rate decayed geometrically per epoch.
        Returns the per-step losses."""
        import torch
        n = (ts.X.shape[0] // batch) * batch
        if n == 0:
            raise ValueError("training set smaller than one batch")
        dev = torch.device("cuda", self.device)
        T_d, W_d, R_d = self.targets(ts, device=dev)  # prepared on the device
        if self.net.get_params()[-2 * self.ncell:].any() == 0:
            self.init_output_bias(T_d, R_d)
        X_d = torch.from_numpy(self.features(ts.X[:, :3], ts.X[:, 3])).to(dev)
        mrw = float(W_d.sum(1).mean() / W_d.shape[1])
        g = torch.Generator(device=dev).manual_seed(seed)     # permutations drawn on the device
        # shuffled copies in the same buffers every epoch: the library keeps its captured
        # training step while the arrays stay the same
        Xb = torch.empty((n, X_d.shape[1]), dtype=X_d.dtype, device=dev)
        Tb = torch.empty((n, T_d.shape[1]), dtype=T_d.dtype, device=dev)
        Wb = torch.empty((n, W_d.shape[1]), dtype=W_d.dtype, device=dev)
        Rb = None if R_d is None else torch.empty((n, R_d.shape[1]), dtype=R_d.dtype, device=dev)
        losses = []
        self.fit_step_s = 0.0                           # time in the training steps alone
        for ep in range(epochs):
            lr_ep = lr * (lr_final / lr) ** (ep / max(1, epochs - 1))
            perm = torch.randperm(ts.X.shape[0], generator=g, device=dev)[:n]
            torch.index_select(X_d, 0, perm, out=Xb)
            torch.index_select(T_d, 0, perm, out=Tb)
            torch.index_select(W_d, 0, perm, out=Wb)
            if Rb is not None:
                torch.index_select(R_d, 0, perm, out=Rb)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            l = self.net.train_batches(Xb, Tb, Wb, n // batch, batch, lr=lr_ep,
                                       mean_row_weight=mrw, R0=Rb)      # synchronizes
            self.fit_step_s += time.perf_counter() - t0
            losses.append(l)
            if verbose:
                print(f"epoch {ep}: lr {lr_ep:.2e} loss {l.mean():.5f}", flush=True)
        return np.concatenate(losses)

    # -------------------------------------------------------------- inference ----
    def predict(self, xyz, e_kev):
        """Per-photon detection maps p [n, ncell] and mean arrival times [n, ncell] (ns
        after the deposit) in actual cell order: features, direct light, network and
        unfolding on the device."""
        xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0]
        dep = np.zeros((n, 4), np.float32)
        dep[:, :3] = xyz
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(e_kev, np.float32), (n,)))
        p = np.zeros((n, self.ncell), np.float32)
        t = np.zeros((n, self.ncell), np.float32)
        _mcheck(_bind().crysp_surrogate_predict(self._sh, dep.ctypes.data, e.ctypes.data, n,
                                                p.ctypes.data, t.ctypes.data),
                "crysp_surrogate_predict")
        return p, t

    def sample(self, dep: Deposits, ev_first, use, yield_per_mev: float, *, seed: int,
               out=None):
        """Poisson event maps from the network for the deposits with use != 0 (all on the
        device): returns (maps [nev, ncell] int32, mean times [nev, ncell])."""
        nev = dep.n_events
        ndep = int(dep.xyzt.shape[0])
        if out is None:
            maps = np.zeros((nev, self.ncell), np.int32)
            tmean = np.zeros((nev, self.ncell), np.float32)
        else:
            maps, tmean = out
        use8 = np.ascontiguousarray(use, np.uint8) if isinstance(use, np.ndarray) else use
        _lib.check_array(dep.xyzt, "xyzt", np.float32, (ndep, 4), device=self.device)
        _lib.check_array(dep.e_kev, "e_kev", np.float32, (ndep,), device=self.device)
        _lib.check_array(use8, "use", np.uint8, (ndep,), device=self.device)
        _lib.check_array(ev_first, "ev_first", np.int64, (nev + 1,), device=self.device)
        _lib.check_array(maps, "maps", np.int32, numel=nev * self.ncell, device=self.device)
        _lib.check_array(tmean, "tmean", np.float32, numel=nev * self.ncell, device=self.device)
        _lib.ready(dep.xyzt, dep.e_kev, use8, ev_first, maps, tmean)
        _mcheck(_bind().crysp_surrogate_sample(self._sh, _ptr(dep.xyzt), _ptr(dep.e_kev), _ptr(use8),
                                               _ptr(ev_first), ndep, nev, float(yield_per_mev),
                                               int(seed), _ptr(maps), _ptr(tmean), None),
                "crysp_surrogate_sample")
        return maps, tmean

    def bake(self, shape=(48, 48, 40), e_kev=None):
        """The network on a regular (gx, gy, gz) grid spanning the crystal, at one energy
        (default mid-range): float32 [gz, gy, gx, 2, ncell] (p, then mean time)."""
        gx, gy, gz = shape
        L = np.asarray(self.box.L, np.float64)
        xs = np.linspace(0, L[0], gx)
        ys = np.linspace(0, L[1], gy)
        zs = np.linspace(0, L[2], gz)
        Z, Yg, Xg = np.meshgrid(zs, ys, xs, indexing="ij")
        e = np.full(Xg.size, np.mean(self.e_range) if e_kev is None else e_kev, np.float32)
        p, t = self.predict(np.column_stack([Xg.ravel(), Yg.ravel(), Z.ravel()]), e)
        grid = np.empty((gz, gy, gx, 2, self.ncell), np.float32)
        grid[..., 0, :] = p.reshape(gz, gy, gx, self.ncell)
        grid[..., 1, :] = t.reshape(gz, gy, gx, self.ncell)
        self.grid = grid
        return grid


def split_deposits(dep: Deposits, box: Box, fallback_mm: float = 1.0) -> np.ndarray:
    """True where a deposit is at least fallback_mm from every face (surrogate region)."""
    L = np.asarray(box.L, np.float32)
    x = dep.xyzt[:, :3]
    return np.all((x >= fallback_mm) & (x <= L - fallback_mm), axis=1)


def sample_events(grid, box: Box, dep: Deposits, ev_first, use, yield_per_mev: float, *,
                  seed: int, device: int = 0, out=None):
    """Poisson event maps from a baked grid for the deposits with use != 0."""
    gz, gy, gx, two, ncell = grid.shape
    nev = dep.n_events
    lib = _bind()
    if out is None:
        maps = np.zeros((nev, ncell), np.int32)
        tmean = np.zeros((nev, ncell), np.float32)
    else:
        maps, tmean = out
    use8 = np.ascontiguousarray(use, np.uint8) if isinstance(use, np.ndarray) else use
    ndep = int(dep.xyzt.shape[0])
    _lib.check_array(grid, "grid", np.float32, (gz, gy, gx, 2, ncell), device=device)
    _lib.check_array(dep.xyzt, "xyzt", np.float32, (ndep, 4), device=device)
    _lib.check_array(dep.e_kev, "e_kev", np.float32, (ndep,), device=device)
    _lib.check_array(use8, "use", np.uint8, (ndep,), device=device)
    _lib.check_array(ev_first, "ev_first", np.int64, (nev + 1,), device=device)
    _lib.check_array(maps, "maps", np.int32, numel=nev * ncell, device=device)
    _lib.check_array(tmean, "tmean", np.float32, numel=nev * ncell, device=device)
    _lib.ready(grid, dep.xyzt, dep.e_kev, use8, ev_first, maps, tmean)
    rc = lib.crysp_surrogate_sample_events(
        device, _ptr(grid), gx, gy, gz, ncell, C.byref(box), _ptr(dep.xyzt), _ptr(dep.e_kev),
        _ptr(use8), _ptr(ev_first), int(dep.xyzt.shape[0]), nev, float(yield_per_mev),
        int(seed), _ptr(maps), _ptr(tmean), None)
    _mcheck(rc, "crysp_surrogate_sample_events")
    return maps, tmean


def count_error(p_pred, p_true, floor_frac=0.0):
    """Mean per-SiPM relative count error, count-weighted per point:
    sum_c |p_pred - p_true| / sum_c p_true, averaged over points."""
    num = np.abs(p_pred - p_true).sum(1)
    den = p_true.sum(1)
    return float(np.mean(num / np.maximum(den, 1e-30)))
