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


def test_oracle_mlp_rate_offset_loss_and_gradient():
    """The Poisson loss with rate offsets (rate e^y + R0): finite differences of the
    gradient, and R0 = 0 gives the plain loss."""
    rng = np.random.default_rng(1)
    shapes = M.layer_shapes(4, 16, 2, 8)
    n = sum(a * b + a for a, b in shapes)
    p = rng.normal(0, 0.3, n)
    X = rng.uniform(-1, 1, (7, 4))
    T = np.abs(rng.normal(0, 1, (7, 8)))
    W = rng.uniform(0, 1, (7, 8))
    R0 = rng.uniform(0.0, 0.5, (7, 3))
    L, g = M.loss_and_grad(p, X, T, W, shapes, n_pois=3, R0=R0)
    h = 1e-6
    for i in rng.choice(n, 60, replace=False):
        q = p.copy(); q[i] += h
        Lp, _ = M.loss_and_grad(q, X, T, W, shapes, n_pois=3, R0=R0)
        q[i] -= 2 * h
        Lm, _ = M.loss_and_grad(q, X, T, W, shapes, n_pois=3, R0=R0)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    L0, g0 = M.loss_and_grad(p, X, T, W, shapes, n_pois=3, R0=np.zeros((7, 3)))
    L1, g1 = M.loss_and_grad(p, X, T, W, shapes, n_pois=3)
    assert abs(L0 - L1) < 1e-12 and np.allclose(g0, g1, rtol=1e-12, atol=1e-15)


def test_direct_light_reference_pins():
    """The FP64 direct-light reference: with T = 1 and no attenuation every cell gets
    PDE * (solid angle of the cell) / 4 pi (closed form); with Fresnel transmission and
    attenuation, a cone that stays inside the face sums to PDE / 2 * int T(u) e^{-k/u} du."""
    import math
    from scipy.integrate import quad
    from oracle import direct_light as DL
    L, p = (25.6, 25.6, 20.0), 3.2
    for x0, y0, z in [(5.1, 7.3, 19.0), (3.2, 3.2, 18.0), (1.0, 0.5, 19.9)]:
        D = DL.direct_light(x0, y0, z, L=L, n1=1.5, n2=1.5, pde=0.3, pitch=p, nx=8, ny=8,
                            transmission=lambda u: np.ones_like(u), ns=1200, nphi=2400)
        d = L[2] - z
        ref = np.array([0.3 * DL.rectangle_solid_angle(i * p - x0, (i + 1) * p - x0,
                                                       j * p - y0, (j + 1) * p - y0, d) / (4 * math.pi)
                        for j in range(8) for i in range(8)])
        assert np.abs(D - ref).max() < 1e-4 * ref.max(), (x0, y0, z)
    n1, n2 = 1.79, 1.5
    uc = DL.critical_cos(n1, n2)
    for z, lam in [(19.0, 1 / 400), (19.9, 0.0)]:
        d = L[2] - z
        D = DL.direct_light(12.8, 12.8, z, L=L, n1=n1, n2=n2, pde=0.3, pitch=p, nx=8, ny=8,
                            inv_lambda=lam, ns=800, nphi=1600)
        I = quad(lambda u: DL.fresnel_T(n1, n2, u) * math.exp(-d * lam / u), uc, 1, limit=200)[0]
        assert abs(D.sum() - 0.15 * I) < 1e-6 * I
    # a wrong sign in the attenuation or a missing transmission factor is caught
    D_bad = DL.direct_light(12.8, 12.8, 15.0, L=L, n1=n1, n2=n2, pde=0.3, pitch=p, nx=8, ny=8,
                            inv_lambda=-1 / 400, ns=400, nphi=800)
    I = quad(lambda u: DL.fresnel_T(n1, n2, u) * math.exp(-5.0 / 400 / u), uc, 1)[0]
    assert abs(D_bad.sum() - 0.15 * I) > 1e-3 * I


def test_oracle_gelu_closed_forms():
    """The oracle's activation is the tanh form of GELU: 0 and slope 1/2 at the origin,
    x -> x and -> 0 far out, within 5e-4 of the exact GELU x Phi(x) everywhere (the known
    accuracy of the tanh form), the odd part x/2 (GELU(x) - GELU(-x) = x), and a
    derivative that equals the central differences of the function."""
    from scipy.stats import norm
    z = np.linspace(-8.0, 8.0, 4001)
    g = M.gelu(z)
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    np.testing.assert_allclose(g - M.gelu(-z), z, atol=1e-12)
    assert np.abs(g - z * norm.cdf(z)).max() < 5e-4
    assert abs(M.gelu(np.array([12.0]))[0] - 12.0) < 1e-12 and abs(M.gelu(np.array([-12.0]))[0]) < 1e-12
    h = 1e-6
    zz = np.array([-3.0, -1.0, -0.2, 0.5, 2.0])
    np.testing.assert_allclose(M.gelu_grad(zz), (M.gelu(zz + h) - M.gelu(zz - h)) / (2 * h), rtol=1e-6)


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
@pytest.mark.parametrize("hidden,n_hidden,n_out,rows", [
    (256, 3, 128, 256),        # the surrogate's network; 4 batch slices of one K stage
    (256, 3, 128, 4096),       # 32 slices of two stages
    (192, 2, 64, 1088),        # 64-row weight-gradient tiles; one slice of 17 stages
    (64, 1, 192, 200),         # padded rows (200 of 256), 64-wide hidden layer
    (256, 3, 128, 65600),      # 128-row workgroups (two row blocks per wave), the last padded
])
def test_mlp_gradient_against_fp64_oracle(hidden, n_hidden, n_out, rows):
    """Forward outputs, loss and every layer's weight and bias gradient of the device step
    against the FP64 oracle on the same bf16-rounded weights and inputs, over network shapes
    and batch sizes that take each tile shape of the weight-gradient kernel and several
    batch-slice layouts (the bias gradients are the slices' dZ row sums there), and both
    row-kernel shapes (64-row workgroups, and 128-row ones once the batch gives every compute
    unit two)."""
    from crysplight_amd.surrogate import MLP
    net = MLP(4, hidden, n_hidden, n_out, seed=3)
    rng = np.random.default_rng(2)
    shift = np.array([12.8, 12.8, 10.0, 260.0], np.float32)
    scale = np.array([1 / 12.8, 1 / 12.8, 1 / 10.0, 1 / 250.0], np.float32)
    net.set_input_norm(shift, scale)
    X = np.column_stack([rng.uniform(0, 25.6, rows), rng.uniform(0, 25.6, rows),
                         rng.uniform(0, 20, rows), rng.uniform(10, 511, rows)]).astype(np.float32)
    T = rng.normal(0, 1, (rows, n_out)).astype(np.float32)
    W = rng.uniform(0, 1, (rows, n_out)).astype(np.float32)
    p = net.get_params()
    y_dev = net.forward(X)
    y_ref, _ = M.forward(p, X, net.shapes, shift, scale, bf16=True)
    assert np.linalg.norm(y_dev - y_ref) / np.linalg.norm(y_ref) < 2e-2
    n_pois = n_out // 2
    net.set_poisson_outputs(n_pois)
    T[:, :n_pois] = np.abs(T[:, :n_pois]) * 0.01
    R0 = rng.uniform(0.0, 0.01, (rows, n_pois)).astype(np.float32)
    for r0 in (None, R0):
        loss, g = net.gradient(X, T, W, R0=r0)
        L, gr = M.loss_and_grad(p, X, T, W, net.shapes, shift, scale, bf16=True, n_pois=n_pois, R0=r0)
        assert abs(loss - L) / abs(L) < 1e-2, (r0 is None, loss, L)
        o = 0
        for (n_o, n_i) in net.shapes:
            for size in (n_o * n_i, n_o):
                a, b = g[o:o + size], gr[o:o + size]
                rel = np.linalg.norm(a - b) / np.linalg.norm(b)
                cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
                assert rel < 5e-2 and cos > 0.998, (r0 is None, o, size, rel, cos)
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
    err0 = SE.validate(sur0, n_points=100, photons=400_000, region_points_n=0)["uniform"]["rms_rel_debiased"]
    sur, info = SE.train_for_box(box, K.C1_OP, K.C1_RO, tau_ns=1000.0, n_samples=100_000,
                                 photons=4000, epochs=15, batch=4096)
    v = SE.validate(sur, n_points=100, photons=400_000, region_points_n=0)["uniform"]
    assert v["rms_rel_debiased"] < 0.12, v
    assert v["rms_rel_debiased"] < 0.5 * err0, (err0, v)
    csi = GammaMaterial("xcom_CSI.csv", 4.51)
    dep, part_of, _, boxes = SE.gamma_mix((20.0,), 25.6, csi, 54_000.0, 3000, seed=3)
    assert dep.n_events == 3000
    r = SE.compare_events([sur], boxes, K.C1_OP, K.C1_RO, dep, part_of, yield_per_mev=54_000.0,
                          tau_ns=1000.0, rules={"any_face_1mm": {"any_face_mm": 1.0},
                                                "edges_0.5mm": {"edge_mm": 0.5}})
    for name, line in r["lines"].items():
        assert abs(line["energy_sum_bias"]) < 0.05, (name, line)
    assert r["lines"]["edges_0.5mm"]["fallback_deposit_fraction"] < \
        r["lines"]["any_face_1mm"]["fallback_deposit_fraction"]


@gpu
def test_surrogate_events_three_thickness_mix():
    """The configuration's event mix: gamma events over three crystal thicknesses in one
    several-crystal queue, each crystal's deposits sampled from its own small surrogate with
    MC fallback near the faces, against the plain MC of the same deposits. Every thickness
    gets events, its energy sum stays within a few percent, and the position-binned map error
    is reported for all of them together."""
    from crysplight_amd import surrogate_eval as SE
    from crysplight_amd.events import GammaMaterial
    from crysplight_amd.params import Box
    from tests import cases as K
    sur = []
    for t in (10.0, 20.0, 30.0):
        s_, _ = SE.train_for_box(Box((25.6, 25.6, t)), K.C1_OP, K.C1_RO, tau_ns=1000.0,
                                 n_samples=50_000, photons=4000, epochs=10, batch=4096)
        sur.append(s_)
    csi = GammaMaterial("xcom_CSI.csv", 4.51)
    dep, part_of, _, boxes = SE.gamma_mix((10.0, 20.0, 30.0), 25.6, csi, 54_000.0, 3000, seed=5)
    assert dep.n_events == 3000 and sorted(set(part_of.tolist())) == [0, 1, 2]
    r = SE.compare_events(sur, boxes, K.C1_OP, K.C1_RO, dep, part_of, yield_per_mev=54_000.0,
                          tau_ns=1000.0, rules={"any_face_1mm": {"any_face_mm": 1.0},
                                                "corners_0.5mm": {"corner_mm": 0.5}})
    for name, line in r["lines"].items():
        assert set(line["by_thickness"]) == {"10mm", "20mm", "30mm"}, (name, line)
        for t, v in line["by_thickness"].items():
            assert abs(v["energy_sum_bias"]) < 0.06, (name, t, v)
        assert line["position_binned"]["bins"] > 0, (name, line)
    assert r["lines"]["corners_0.5mm"]["fallback_deposit_fraction"] < \
        r["lines"]["any_face_1mm"]["fallback_deposit_fraction"]


def test_training_targets_fold_cells_and_weights():
    """Targets of a tiny training set: folded cell order (the map of a folded point),
    per-photon rates, standardized mean times (0 where a cell saw nothing) and weights."""
    from crysplight_amd import surrogate as S
    from crysplight_amd.params import Box
    sur = S.LightSurrogate.__new__(S.LightSurrogate)
    sur.box, sur.fold, sur.ncell, sur.direct = Box((25.6, 25.6, 20.0)), True, 64, False
    sur.perms, sur.t_shift, sur.t_scale = S.fold_perms(8), 100.0, 50.0
    rng = np.random.default_rng(3)
    n, nph = 500, 1000
    X = np.column_stack([rng.uniform(0, 25.6, n), rng.uniform(0, 25.6, n), rng.uniform(0, 20, n),
                         rng.uniform(10, 511, n)]).astype(np.float32)
    counts = rng.integers(0, 40, (n, 64)).astype(np.int32)
    tsum = counts * rng.uniform(50, 400, (n, 64))
    ts = S.TrainingSet(X, counts, tsum, nph)
    T, W, _ = sur.targets(ts)
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


def _taper(a):
    """1 below 2 mm, 0 beyond 3 mm, a cubic smoothstep between (the weight of the computed
    light terms)."""
    u = np.clip(3.0 - a, 0.0, 1.0)
    return u * u * (3.0 - 2.0 * u)


def _direct_points(Lz):
    """Sources near the readout face (on a cell, on a cell boundary, on a corner of four
    cells, near the face edges and corners) and deeper ones."""
    return np.array([[5.1, 7.3, Lz - 0.05], [6.4, 7.0, Lz - 0.3], [9.6, 9.6, Lz - 0.2],
                     [0.2, 12.0, Lz - 0.5], [0.3, 0.4, Lz - 0.8], [25.4, 25.5, Lz - 0.1],
                     [12.8, 12.8, Lz - 1.0], [3.1, 14.5, Lz - 2.5], [12.0, 5.0, Lz - 6.0],
                     [20.0, 3.0, 2.0]], np.float32)


@gpu
def test_direct_light_against_fp64_reference():
    """The device's direct-light term (piecewise azimuth integral over a tabulated H)
    against the FP64 direct-direction quadrature, at sources near the readout face, on
    cell boundaries and corners, near the face edges, and deep."""
    from oracle import direct_light as DL
    from crysplight_amd.surrogate import LightSurrogate
    from tests import cases as K
    sur = LightSurrogate(K.C1_BOX, K.C1_OP, K.C1_RO, tau_ns=1000.0, fold=False, images=False)
    pts = _direct_points(20.0)
    D = sur.direct_light(pts)
    for q, got in zip(pts, D):
        ref = DL.direct_light(float(q[0]), float(q[1]), float(q[2]), L=(25.6, 25.6, 20.0),
                              n1=1.79, n2=1.5, pde=0.3, pitch=3.2, nx=8, ny=8,
                              inv_lambda=1 / 400.0) * _taper(20.0 - float(q[2]))
        if _taper(20.0 - float(q[2])) == 0.0:
            assert not got.any(), q                  # deeper than 3 mm: nothing computed
            continue
        # the reference's own cell-assignment error grows with the cone (deep sources);
        # near the readout face it is below 2e-4 of the largest cell
        tol = 4e-4 if 20.0 - q[2] < 3.0 else 8e-3
        assert np.abs(got - ref).max() <= tol * ref.max(), (q, np.abs(got - ref).max(), ref.max())
        assert abs(got.sum() - ref.sum()) <= 2e-4 * ref.sum(), (q, got.sum(), ref.sum())


@gpu
def test_direct_light_equals_monte_carlo_with_black_walls(ctx):
    """With every wall black (front-painted, R = 0) only light that reaches the readout
    face straight is detected (a Fresnel reflection there sends it to a black wall), so the
    transport's maps must equal the direct-light term within Poisson noise."""
    from crysplight_amd.params import OpticalParams
    from crysplight_amd.surrogate import LightSurrogate, mc_light_maps
    from tests import cases as K
    black = OpticalParams(1.79, 1.5, 400.0, float("inf"), 0.0, False, False, 0.0, 0.3)
    sur = LightSurrogate(K.C1_BOX, K.C1_OP, K.C1_RO, tau_ns=1000.0, fold=False, images=False)
    pts = _direct_points(20.0)
    pts = pts[20.0 - pts[:, 2] <= 2.0]               # where the computed term has full weight
    N = 4_000_000
    counts, _ = mc_light_maps(K.C1_BOX, black, K.C1_RO, pts, N, seed=91, tau_ns=0.0, ctx=ctx)
    mu = N * sur.direct_light(pts).astype(np.float64)
    m = mu > 20
    z = (counts - mu)[m] / np.sqrt(mu[m])
    assert np.all(counts[~m] <= mu[~m] + 5 * np.sqrt(mu[~m]) + 5)
    assert np.abs(z).max() < 5.5, np.abs(z).max()
    assert np.mean(z * z) < 1.3, np.mean(z * z)
    # a direct term 10% off would be caught
    assert not np.all(np.abs((counts[m] - mu[m] / 0.9)) < 5.5 * np.sqrt(mu[m]))


def _constant_grid_events(lam_cell, n_ev, dev):
    """Single-deposit events on a constant baked grid: every cell's rate is lam_cell."""
    import torch
    from crysplight_amd.events import Deposits
    grid = np.zeros((2, 2, 2, 2, 64), np.float32)
    grid[:, :, :, 0, :] = 1.0                        # p = 1 everywhere
    xyzt = np.tile(np.array([[12.8, 12.8, 10.0, 0.0]], np.float32), (n_ev, 1))
    e = np.full(n_ev, 1000.0, np.float32)            # yield * E / 1000 = yield
    dep = Deposits(torch.from_numpy(xyzt).to(dev), None, None, torch.from_numpy(e).to(dev), n_ev)
    evf = torch.arange(n_ev + 1, dtype=torch.int64, device=dev)
    use = torch.ones(n_ev, dtype=torch.uint8, device=dev)
    return torch.from_numpy(grid).to(dev), dep, evf, use


@gpu
@pytest.mark.parametrize("lam", [3.5, 20.0, 50.0, 51.0, 4000.0])
def test_device_poisson_sampler_moments(lam):
    """The device Poisson sampler (src/events.jl:11-22: Knuth for lambda <= 50, rounded
    Gaussian above) through crysp_surrogate_sample_events on a constant grid: mean lambda
    and variance lambda (lambda + 1/12 on the rounded-Gaussian branch) within 5 sigma over
    1.28e6 draws; on the Knuth branch also the probabilities of 0..3 counts."""
    import math
    import torch
    from crysplight_amd import surrogate as S
    from crysplight_amd.params import Box
    dev = torch.device("cuda", 0)
    n_ev = 20_000
    grid, dep, evf, use = _constant_grid_events(lam, n_ev, dev)
    maps = torch.zeros((n_ev, 64), dtype=torch.int32, device=dev)
    tm = torch.zeros((n_ev, 64), dtype=torch.float32, device=dev)
    S.sample_events(grid, Box((25.6, 25.6, 20.0)), dep, evf, use, lam, seed=17, out=(maps, tm))
    x = maps.cpu().numpy().astype(np.float64).ravel()
    n = x.size
    var_want = lam + (1.0 / 12.0 if lam > 50 else 0.0)
    assert abs(x.mean() - lam) < 5 * math.sqrt(lam / n), (x.mean(), lam)
    sd_var = math.sqrt((lam + 2 * lam * lam) / n)
    assert abs(x.var() - var_want) < 5 * sd_var, (x.var(), var_want)
    if lam <= 50:
        for k in range(4):
            pk = math.exp(-lam) * lam ** k / math.factorial(k)
            assert abs(np.mean(x == k) - pk) < 5 * math.sqrt(pk * (1 - pk) / n) + 1e-6, (k, pk)
    # different events and cells draw from different streams
    assert len(np.unique(maps.cpu().numpy()[:50, :8].ravel())) > 5 or lam < 5


@gpu
def test_grid_lookup_at_nodes_and_midpoints():
    """The baked-grid lookup: at a grid node the rate is the node's value, midway between
    nodes the trilinear mean of the 8 neighbours (numpy, float32)."""
    import torch
    from crysplight_amd import _lib
    from crysplight_amd import surrogate as S
    from crysplight_amd.params import Box
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(9)
    gz, gy, gx = 5, 6, 7
    L = np.array([25.6, 25.6, 20.0])
    grid = rng.uniform(0.001, 0.01, (gz, gy, gx, 2, 64)).astype(np.float32)
    ix, iy, iz = rng.integers(0, gx - 1, 40), rng.integers(0, gy - 1, 40), rng.integers(0, gz - 1, 40)
    nodes = np.column_stack([ix / (gx - 1), iy / (gy - 1), iz / (gz - 1)]) * L
    mids = np.column_stack([(ix + 0.5) / (gx - 1), (iy + 0.5) / (gy - 1), (iz + 0.5) / (gz - 1)]) * L
    xyz = np.concatenate([nodes, mids]).astype(np.float32)
    n = xyz.shape[0]
    xyzt = np.zeros((n, 4), np.float32)
    xyzt[:, :3] = xyz
    lam = np.zeros((n, 64), np.float32)
    maps = np.zeros((n, 64), np.int32)
    lib = S._bind()
    rc = lib.crysp_surrogate_sample_events(0, grid.ctypes.data, gx, gy, gz, 64,
                                           C_box(Box(tuple(L))), xyzt.ctypes.data,
                                           np.ones(n, np.float32).ctypes.data,
                                           np.ones(n, np.uint8).ctypes.data,
                                           np.arange(n + 1, dtype=np.int64).ctypes.data, n, n,
                                           1000.0, 3, maps.ctypes.data, None, lam.ctypes.data)
    assert rc == 0
    np.testing.assert_allclose(lam[:40], grid[iz, iy, ix, 0, :], rtol=2e-6)
    mean8 = np.zeros((40, 64))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                mean8 += grid[iz + dz, iy + dy, ix + dx, 0, :] / 8.0
    np.testing.assert_allclose(lam[40:], mean8, rtol=1e-5)


def C_box(b):
    import ctypes as C
    return C.byref(b)


def test_binned_map_errors_statistics():
    """The position-binned error of event maps: two independent Poisson samples of the same
    means give a debiased RMS near 0 (raw mean at the noise floor); a 3% scale error on
    one side is recovered."""
    from crysplight_amd import surrogate_eval as SE
    from crysplight_amd.events import Deposits
    from crysplight_amd.params import Box
    rng = np.random.default_rng(4)
    nev = 40_000
    xyz = rng.uniform(0, 1, (nev, 3)) * np.array([25.6, 25.6, 20.0])
    dep = Deposits(np.column_stack([xyz, np.zeros(nev)]).astype(np.float32),
                   np.full(nev, 1000, np.int32), np.arange(nev, dtype=np.int32),
                   np.ones(nev, np.float32), nev)
    part = np.zeros(nev, np.int32)
    lam = rng.uniform(5, 40, (1, 64)) * (1 + xyz[:, :1] / 25.6)
    a, b = rng.poisson(lam), rng.poisson(lam)
    r0 = SE.binned_map_errors(a, b, dep, part, [Box((25.6, 25.6, 20.0))])
    assert r0["rms_rel_debiased"] < 0.004, r0
    assert abs(r0["mean_abs_rel"] - r0["noise_floor"]) < 0.15 * r0["noise_floor"], r0
    r3 = SE.binned_map_errors(a, rng.poisson(lam * 1.03), dep, part, [Box((25.6, 25.6, 20.0))])
    assert abs(r3["rms_rel_debiased"] - 0.03) < 0.004, r3


def test_near_face_mask_rules():
    from crysplight_amd import surrogate_eval as SE
    L = np.array([[25.6, 25.6, 20.0]] * 5)
    xyz = np.array([[10, 10, 10], [0.5, 10, 10], [0.5, 0.5, 10], [10, 10, 19.8], [0.3, 10, 19.7]])
    assert SE.near_face_mask(xyz, L, any_face_mm=1.0).tolist() == [False, True, True, True, True]
    assert SE.near_face_mask(xyz, L, edge_mm=1.0).tolist() == [False, False, True, False, True]
    assert SE.near_face_mask(xyz, L, readout_mm=0.35).tolist() == [False, False, False, True, True]
    corner = np.array([[0.4, 0.3, 0.2], [0.4, 0.3, 5.0]])
    assert SE.near_face_mask(corner, L[:2], corner_mm=0.5).tolist() == [True, False]


def test_fallback_rule_from_region_errors():
    """The measured fallback rule: the narrowest distances whose complement keeps every
    region's error within the bar (readout bands, two-face edges, three-face corners)."""
    from crysplight_amd import surrogate_eval as SE

    def val(errs):
        return {"regions": {"by_region": {k: {"rms_rel_debiased": v} for k, v in errs.items()}}}
    a = val({"readout_0-0.05mm": 0.01, "readout_0.05-0.1mm": 0.01,
             "edge_side_side_0-0.25mm": 0.025, "edge_side_side_0.25-0.5mm": 0.015,
             "edge_side_front_0-0.25mm": 0.01, "corner_front_0-0.25mm": 0.03,
             "corner_front_0.25-0.5mm": 0.021, "corner_front_0.5-1.0mm": 0.01})
    b = val({"readout_0-0.05mm": 0.03, "readout_0.05-0.1mm": 0.01, "edge_side_side_0-0.25mm": 0.01})
    r = SE.fallback_from_regions([a, b], bar=0.02)
    assert r == {"readout_mm": 0.05, "edge_mm": 0.25, "corner_mm": 0.5}, r


@gpu
def test_direct_light_mirror_images_near_side_walls():
    """Near a side wall and the readout face (air-gap wrap), the computed term adds the
    direct light of the source's mirror images in the side walls (total internal
    reflection), each weighted by a smoothstep of its walls' distances and of the depth:
    restated here with the FP64 reference."""
    from oracle import direct_light as DL
    from crysplight_amd.surrogate import LightSurrogate
    from tests import cases as K

    taper = _taper

    sur = LightSurrogate(K.C1_BOX, K.C1_OP, K.C1_RO, tau_ns=1000.0, fold=False, images=True)
    pts = np.array([[0.2, 12.0, 19.5], [0.3, 0.4, 19.2], [25.1, 2.4, 18.5], [1.5, 24.9, 19.9],
                    [12.8, 12.8, 19.0], [0.5, 12.0, 16.0]], np.float32)
    got = sur.direct_light(pts)
    kw = dict(L=(25.6, 25.6, 20.0), n1=1.79, n2=1.5, pde=0.3, pitch=3.2, nx=8, ny=8,
              inv_lambda=1 / 400.0)
    for q, g in zip(pts, got):
        x, y, z = (float(v) for v in q)
        want = DL.direct_light(x, y, z, **kw)
        wd = taper(20.0 - z)
        xs, wx = (x, -x, 51.2 - x), (1.0, taper(x), taper(25.6 - x))
        ys, wy = (y, -y, 51.2 - y), (1.0, taper(y), taper(25.6 - y))
        for a in range(3):
            for b in range(3):
                w = wx[a] * wy[b]
                if (a or b) and w > 0:
                    want = want + w * DL.direct_light(xs[a], ys[b], z, **kw)
        want = want * wd
        assert np.abs(g - want).max() <= 4e-4 * want.max() + 1e-7, (q, np.abs(g - want).max(), want.max())
    # a source away from the walls gets no image light, a deep one nothing computed
    assert np.allclose(got[4], DL.direct_light(12.8, 12.8, 19.0, **kw), rtol=0, atol=4e-4 * got[4].max())
    assert not got[5].any()
