"""The surrogate's CPU-side pins (FP64 MLP oracle against finite differences) and its
GPU checks: MFMA GEMM layouts against numpy, device gradients against the FP64 oracle,
training that lowers the loss, and a trained surrogate against high-statistics MC."""
import numpy as np
import pytest

from oracle import mlp as M


def test_oracle_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    shapes = M.layer_shapes(4, 16, 2, 8)
    n = sum(a * b + a for a, b in shapes)
    p = rng.normal(0, 0.3, n)
    X = rng.uniform(-1, 1, (7, 4))
    T = rng.normal(0, 1, (7, 8))
    W = rng.uniform(0, 1, (7, 8))
    T[:, :3] = np.abs(T[:, :3])
    L, g = M.loss_and_grad(p, X, T, W, shapes, n_pois=3)
    idx = rng.choice(n, 60, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy(); q[i] += h
        Lp, _ = M.loss_and_grad(q, X, T, W, shapes, n_pois=3)
        q[i] -= 2 * h
        Lm, _ = M.loss_and_grad(q, X, T, W, shapes, n_pois=3)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_gelu_closed_forms():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    from scipy.stats import norm
    np.testing.assert_allclose(M.gelu(z), z * norm.cdf(z), rtol=1e-12, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), rtol=1e-6)


def test_round_bf16():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3], np.float32)
    r = M.round_bf16(x)
    # 8 significant bits, ties to even
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] - 3.140625) < 1e-12


def _iso_map(x, y, n, pitch, width=4.0):
    """A light map that depends only on the distance to each SiPM centre: invariant under
    the square face's mirrors and the x<->y swap.  Cell i + n*j, i along x."""
    c = (np.arange(n) + 0.5) * pitch
    dx = x[:, None, None] - c[None, None, :]          # [k, 1, i]
    dy = y[:, None, None] - c[None, :, None]          # [k, j, 1]
    return np.exp(-(dx ** 2 + dy ** 2) / width).reshape(len(x), n * n)


def test_symmetry_fold_against_closed_form():
    from crysplight_amd import surrogate as S
    n, pitch = 8, 3.2
    L = n * pitch
    rng = np.random.default_rng(5)
    x, y = rng.uniform(0, L, 4000), rng.uniform(0, L, 4000)
    xc, yc, code = S.fold_xy(x, y, L)
    assert np.all((0 <= xc) & (xc <= yc + 1e-12) & (yc <= L / 2 + 1e-12))
    assert set(np.unique(code)) == set(range(8))
    perms = S.fold_perms(n)
    for o in range(8):
        assert sorted(perms[o]) == list(range(n * n))
    # the map at the folded point, in folded cell order, unfolds to the map at the point
    got = S.from_folded(_iso_map(xc, yc, n, pitch), code, perms)
    assert np.allclose(got, _iso_map(x, y, n, pitch), rtol=1e-12, atol=1e-300)
    # and folding the actual map gives the folded point's map
    assert np.allclose(S.to_folded(_iso_map(x, y, n, pitch), code, perms),
                       _iso_map(xc, yc, n, pitch), rtol=1e-12, atol=1e-300)
    # a wrong permutation (mirror without the swap) is caught
    bad = S.from_folded(_iso_map(xc, yc, n, pitch), code & 3, perms)
    assert not np.allclose(bad, _iso_map(x, y, n, pitch))


def test_symmetry_fold_only_when_the_optics_allow_it():
    from crysplight_amd import surrogate as S
    from crysplight_amd.params import Box, SipmGrid, readout
    g = SipmGrid(3.2, 8, 8)
    assert S.folds(Box((25.6, 25.6, 20.0)), readout(g))
    assert not S.folds(Box((25.6, 24.0, 20.0)), readout(g))
    assert not S.folds(Box((30.0, 30.0, 20.0)), readout(g))          # grid short of the face
    assert not S.folds(Box((25.6, 25.6, 20.0)), readout(g, center=(12.8, 12.8), radius=5.0))


gpu = pytest.mark.gpu


@gpu
@pytest.mark.parametrize("a_t,b_t", [(False, True), (True, False), (False, False)])
def test_gemm_layout_against_numpy(a_t, b_t):
    from crysplight_amd.surrogate import debug_gemm
    rng = np.random.default_rng(1)
    Mm, N, K = 128, 192, 96
    A = M.round_bf16(rng.normal(size=(K, Mm) if a_t else (Mm, K))).astype(np.float32)
    B = M.round_bf16(rng.normal(size=(N, K) if b_t else (K, N))).astype(np.float32)
    C = debug_gemm(A, B, a_t, b_t)
    ref = (A.T if a_t else A).astype(np.float64) @ (B.T if b_t else B).astype(np.float64)
    np.testing.assert_allclose(C, ref, rtol=1e-4, atol=1e-3)


@gpu
def test_mlp_gradient_against_fp64_oracle():
    from crysplight_amd.surrogate import MLP
    net = MLP(4, 256, 3, 128, seed=3)
    rng = np.random.default_rng(2)
    shift = np.array([12.8, 12.8, 10.0, 260.0], np.float32)
    scale = np.array([1 / 12.8, 1 / 12.8, 1 / 10.0, 1 / 250.0], np.float32)
    net.set_input_norm(shift, scale)
    X = np.column_stack([rng.uniform(0, 25.6, 256), rng.uniform(0, 25.6, 256),
                         rng.uniform(0, 20, 256), rng.uniform(10, 511, 256)]).astype(np.float32)
    T = rng.normal(0, 1, (256, 128)).astype(np.float32)
    W = rng.uniform(0, 1, (256, 128)).astype(np.float32)
    p = net.get_params()
    y_dev = net.forward(X)
    y_ref, _ = M.forward(p, X, net.shapes, shift, scale, bf16=True)
    assert np.linalg.norm(y_dev - y_ref) / np.linalg.norm(y_ref) < 2e-2
    net.set_poisson_outputs(64)
    T[:, :64] = np.abs(T[:, :64]) * 0.01
    loss, g = net.gradient(X, T, W)
    L, gr = M.loss_and_grad(p, X, T, W, net.shapes, shift, scale, bf16=True, n_pois=64)
    assert abs(loss - L) / L < 1e-2
    o = 0
    for (n_o, n_i) in net.shapes:
        for size in (n_o * n_i, n_o):
            a, b = g[o:o + size], gr[o:o + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            assert rel < 5e-2 and cos > 0.998, (o, size, rel, cos)
            o += size


@gpu
def test_training_reduces_loss_on_a_smooth_target():
    import torch
    from crysplight_amd.surrogate import MLP
    net = MLP(4, 256, 3, 128, seed=4)
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (16384, 4)).astype(np.float32)
    k = np.arange(128)
    T = (np.sin(X[:, :1] * (1 + k % 4) + X[:, 1:2] * (k // 32)) + 0.5 * X[:, 2:3] * X[:, 3:4]).astype(np.float32)
    dev = torch.device("cuda", 0)
    Xd, Td = torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev)
    first = net.train_batches(Xd, Td, None, 4, 4096, lr=2e-3)
    for _ in range(30):
        last = net.train_batches(Xd, Td, None, 4, 4096, lr=2e-3)
    assert last.mean() < 0.1 * first[0], (first, last)


@gpu
def test_cached_training_step_follows_new_rates_and_contents():
    """The library keeps its captured training step across calls with the same arrays. A
    second epoch with a new learning rate and reshuffled contents in the same buffers must
    give what one eager step per batch gives (same rates, same batches), and a new loss
    layout (set_poisson_outputs) must retire the captured step."""
    import torch
    from crysplight_amd.surrogate import MLP
    rng = np.random.default_rng(8)
    B, nb = 2048, 3
    X = rng.uniform(-1, 1, (nb * B, 4)).astype(np.float32)
    T = np.concatenate([rng.uniform(0, 0.05, (nb * B, 64)), rng.normal(0, 1, (nb * B, 64))], 1).astype(np.float32)
    W = rng.uniform(0.5, 2.0, (nb * B, 128)).astype(np.float32)
    dev = torch.device("cuda", 0)
    perm = rng.permutation(nb * B)
    epochs = [(X, T, W, 2e-3), (X[perm], T[perm], W[perm], 5e-4)]
    nets = {}
    for mode in ("graph", "eager"):
        net = MLP(4, 256, 3, 128, seed=6)
        net.set_poisson_outputs(64)
        Xd, Td, Wd = (torch.empty(a.shape, dtype=torch.float32, device=dev) for a in (X, T, W))
        losses = []
        for ex, et, ew, lr in epochs:
            for d, h in ((Xd, ex), (Td, et), (Wd, ew)):
                d.copy_(torch.from_numpy(np.ascontiguousarray(h)))
            if mode == "graph":
                losses.append(net.train_batches(Xd, Td, Wd, nb, B, lr=lr))
            else:
                for b in range(nb):
                    sl = slice(b * B, (b + 1) * B)
                    losses.append(net.train_batches(Xd[sl], Td[sl], Wd[sl], 1, B, lr=lr))
        nets[mode] = (net, np.concatenate(losses), (Xd, Td, Wd))
    (ng, lg, (Xd, Td, Wd)), (ne, le, _) = nets["graph"], nets["eager"]
    np.testing.assert_allclose(lg, le, rtol=1e-5)
    np.testing.assert_allclose(ng.get_params(), ne.get_params(), rtol=1e-4, atol=1e-6)
    # a new loss layout on the same buffers: the cached step must not be reused
    ng.set_poisson_outputs(0)
    ne.set_poisson_outputs(0)
    a = ng.train_batches(Xd, Td, Wd, nb, B, lr=1e-3)
    b = np.concatenate([ne.train_batches(Xd[k * B:(k + 1) * B], Td[k * B:(k + 1) * B],
                                         Wd[k * B:(k + 1) * B], 1, B, lr=1e-3) for k in range(nb)])
    np.testing.assert_allclose(a, b, rtol=1e-4)


@gpu
def test_surrogate_end_to_end_small():
    """A small surrogate trains, beats its untrained start by a wide margin on held-out MC
    means, and its event maps (with MC fallback) keep the energy sum within a few percent
    of the plain MC while running faster."""
    from crysplight_amd import surrogate_eval as SE
    from crysplight_amd.events import GammaMaterial
    from crysplight_amd.params import Box
    from tests import cases as K
    box = Box((25.6, 25.6, 20.0))
    sur0, _ = SE.train_for_box(box, K.C1_OP, K.C1_RO, tau_ns=1000.0, n_samples=8192,
                               photons=2000, epochs=1, batch=8192)
    err0 = SE.validate(sur0, n_points=100, photons=400_000)["count_error_beyond_fallback"]
    sur, info = SE.train_for_box(box, K.C1_OP, K.C1_RO, tau_ns=1000.0, n_samples=100_000,
                                 photons=4000, epochs=15, batch=4096)
    v = SE.validate(sur, n_points=100, photons=400_000)
    assert v["count_error_beyond_fallback"] < 0.12, v
    assert v["count_error_beyond_fallback"] < 0.5 * err0, (err0, v)
    csi = GammaMaterial("xcom_CSI.csv", 4.51)
    dep, part_of, _, boxes = SE.gamma_mix((20.0,), 25.6, csi, 54_000.0, 3000, seed=3)
    assert dep.n_events == 3000
    r = SE.compare_events([sur], boxes, K.C1_OP, K.C1_RO, dep, part_of, yield_per_mev=54_000.0,
                          tau_ns=1000.0)
    assert abs(r["energy_sum_bias"]) < 0.05, r
    # 3000 events take a few ms either way, so launch costs blur the ratio; the gain at
    # full size (1e6 events, about 600x) is the bench's measurement
    assert r["speedup_sampling_only"] > 2, r


def test_training_targets_fold_cells_and_weights():
    """Targets of a tiny training set: folded cell order (the map of a folded point),
    per-photon rates, standardized mean times (0 where a cell saw nothing) and weights."""
    from crysplight_amd import surrogate as S
    from crysplight_amd.params import Box
    sur = S.LightSurrogate.__new__(S.LightSurrogate)
    sur.box, sur.fold, sur.ncell = Box((25.6, 25.6, 20.0)), True, 64
    sur.perms, sur.t_shift, sur.t_scale = S.fold_perms(8), 100.0, 50.0
    rng = np.random.default_rng(3)
    n, nph = 500, 1000
    X = np.column_stack([rng.uniform(0, 25.6, n), rng.uniform(0, 25.6, n), rng.uniform(0, 20, n),
                         rng.uniform(10, 511, n)]).astype(np.float32)
    counts = rng.integers(0, 40, (n, 64)).astype(np.int32)
    tsum = counts * rng.uniform(50, 400, (n, 64))
    ts = S.TrainingSet(X, counts, tsum, nph)
    T, W = sur.targets(ts)
    _, _, code = S.fold_xy(X[:, 0], X[:, 1], 25.6)
    cf = S.to_folded(counts, code, sur.perms)
    tf = S.to_folded(tsum, code, sur.perms)
    assert np.array_equal(T[:, :64], (cf / nph).astype(np.float32))
    with np.errstate(invalid="ignore"):
        want_t = np.nan_to_num((tf / cf - 100.0) / 50.0, nan=0.0).astype(np.float32)
    assert np.array_equal(T[:, 64:], want_t)
    assert np.all(T[:, 64:][cf == 0] == 0.0)
    assert np.allclose(W[:, :64], 1.0 / (counts / nph).mean(), rtol=1e-7)   # float32
    assert np.array_equal(W[:, 64:], (0.25 * cf.astype(np.float32) / (cf.astype(np.float32) + 20.0)))
