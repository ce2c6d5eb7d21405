"""Pins for the CPU restatement of the path-guiding network (oracle/guide_ref.py).

Each check is independent of the restated formula it covers:
  * hash-grid encoding: trilinear interpolation reproduces a linear function
    of the lattice exactly, on dense levels and on collision-free regions of
    hashed levels;
  * vMF mixture: normalisation by quadrature, the closed-form mean cosine
    coth(k) - 1/k of the sampler, and the importance identity
    E_q[1/q] = 4 pi for the mixture sampler/pdf pair;
  * loss and back-propagation: float64 torch autograd of the same model;
  * Adam: the closed form of constant-gradient steps (each moves lr * sign(g)
    once bias correction is applied);
  * normal encoding (real spherical harmonics): scipy's complex Y_l^m
    converted to the real basis, orthonormality by exact quadrature, and
    the closed-form values at the poles;
  * MLP forward: positive homogeneity of a bias-free ReLU network (bit for
    bit, since doubling is exact in FP16 and FP32), identity and
    permutation weights whose output is known without a matrix product, and
    a one-hot input through non-negative weights, which reads back a row of
    the product W1 W2 W3 W4;
  * hash slots: the published primes evaluated with Python integers.
"""
import math

import numpy as np
import pytest
import torch

from oracle import guide_ref as G


def _params(seed, t_log2, grid_scale=0.5, w_scale=1.0):
    rng = np.random.default_rng(seed)
    P = G.GuideParams(seed=seed, t_log2=t_log2)
    P.grid = rng.uniform(-grid_scale, grid_scale, P.grid.shape).astype(np.float32)
    P.W = [(w * np.float32(w_scale)).astype(np.float32) for w in P.W]
    return P


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


# ---- hash grid -----------------------------------------------------------------

def test_level_resolutions_geometric():
    r = G.level_resolutions()
    assert r[0] == 16 and r[-1] in (2047, 2048)
    assert np.all(np.diff(r) > 0)
    ratio = r[1:] / r[:-1]
    assert np.all(np.abs(ratio - (2048 / 16) ** (1 / 15)) < 0.07)


def _linear_grid(P, level, coef, corners):
    """Writes coef . (ix, iy, iz) + coef[3] into the level's slots of the
    given integer corners; returns False if two corners share a slot."""
    res = int(P.res[level])
    s = G.corner_index(corners[:, 0], corners[:, 1], corners[:, 2], res, P.t_log2)
    if len(np.unique(s)) != len(s):
        return False
    vals = corners @ np.asarray(coef[:3], np.float64) + coef[3]
    P.grid[level, s, 0] = vals
    P.grid[level, s, 1] = -vals
    return True


def test_trilinear_reproduces_linear_function_dense_and_hashed():
    rng = np.random.default_rng(5)
    P = _params(0, 14)
    coef = (1.0, 2.0, 3.0, -5.0)
    checked_dense = checked_hash = 0
    for level in range(G.L_LEVELS):
        res = int(P.res[level])
        dense = (res + 1) ** 3 <= (1 << P.t_log2)
        # a 3x3x3-cell sub-box of the lattice (whole lattice when dense)
        lo = rng.integers(0, res - 3, 3)
        span = res + 1 if dense else 4
        base = np.zeros(3, np.int64) if dense else lo
        ax = [np.arange(span) + base[k] for k in range(3)]
        corners = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
        if not _linear_grid(P, level, coef, corners):
            continue
        # integer values up to ~(3*2048)*6: keep them exact in FP16 only where small
        if np.abs(P.grid[level]).max() > 2048:
            continue
        hi = base + span - 1
        x = rng.uniform(base + 0.01, hi - 0.01, size=(200, 3)) / res
        enc, _, _ = G.encode(P, x.astype(np.float32), _unit(rng, 200))
        want = (x.astype(np.float32) * np.float32(res)) @ np.array(coef[:3]) + coef[3]
        tol = 2e-3 * np.abs(want).max() + 1e-3
        assert np.abs(enc[:, 2 * level] - want).max() <= tol, level
        assert np.abs(enc[:, 2 * level + 1] + want).max() <= tol, level
        checked_dense += dense
        checked_hash += not dense
    assert checked_dense >= 2 and checked_hash >= 3


def test_hash_slots_in_range():
    rng = np.random.default_rng(1)
    for t in (8, 14, 19):
        for res in (16, 300, 2048):
            c = rng.integers(0, res + 1, size=(1000, 3))
            s = G.corner_index(c[:, 0], c[:, 1], c[:, 2], res, t)
            assert s.min() >= 0 and s.max() < (1 << t)


# ---- vMF mixture -------------------------------------------------------------------

@pytest.mark.parametrize("kappa", [1e-3, 0.02, 1.0, 30.0, 2980.0])
def test_vmf_normalisation_quadrature(kappa):
    from scipy.integrate import quad
    c = float(G.vmf_norm(np.float32(kappa)))
    f = lambda t: 2 * math.pi * c * math.exp(kappa * (t - 1.0))  # noqa: E731
    cut = max(-1.0, 1.0 - 60.0 / kappa)
    total = quad(f, cut, 1.0, limit=400)[0] + (quad(f, -1.0, cut, limit=400)[0] if cut > -1 else 0.0)
    assert abs(total - 1.0) < 2e-6


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 500.0])
def test_vmf_sampler_mean_cosine(kappa):
    rng = np.random.default_rng(int(kappa * 10))
    n = 200_000
    mu = _unit(rng, 1)[0]
    mus = np.tile(mu, (n, G.N_LOBES, 1)).astype(np.float32)
    kap = np.full((n, G.N_LOBES), kappa, np.float32)
    pi = np.zeros((n, G.N_LOBES), np.float32)
    pi[:, 0] = 1.0
    u = rng.random((3, n)).astype(np.float32)
    w = G.sample_mixture(mus, kap, pi, u[0], u[1], u[2])
    assert np.abs(np.linalg.norm(w, axis=1) - 1).max() < 1e-5
    c = w @ mu
    want = 1.0 / math.tanh(kappa) - 1.0 / kappa
    assert abs(c.mean() - want) < 4 * c.std() / math.sqrt(n) + 1e-6


def _random_mixture(rng, n, kmax=3.0):
    raw = np.zeros((n, G.N_OUT), np.float32)
    r = raw[:, :40].reshape(n, 8, 5)
    base = rng.normal(size=(8, 5))
    r[:] = base[None]
    r[:, :, 3] = rng.uniform(-2, math.log(kmax), 8)[None]
    return G.decode(raw)


def test_mixture_importance_identity_and_pdf_mass():
    rng = np.random.default_rng(3)
    n = 400_000
    mu, kappa, pi, _ = _random_mixture(rng, n, kmax=20.0)
    u = rng.random((3, n)).astype(np.float32)
    w = G.sample_mixture(mu, kappa, pi, u[0], u[1], u[2])
    q, _ = G.mixture_pdf(mu, kappa, pi, w)
    x = 1.0 / q.astype(np.float64)
    assert abs(x.mean() - 4 * math.pi) < 4 * x.std() / math.sqrt(n)
    # pdf mass over uniform sphere directions
    d = _unit(rng, n)
    qu, _ = G.mixture_pdf(mu, kappa, pi, d)
    m = 4 * math.pi * qu.astype(np.float64)
    assert abs(m.mean() - 1.0) < 4 * m.std() / math.sqrt(n)


# ---- loss and back-propagation vs float64 autograd -----------------------------------

class _Half(torch.autograd.Function):
    """Rounds activations to FP16 where the device stores them; the gradient
    passes straight through (the device back-propagates through the stored
    FP16 values the same way)."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.float16).to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g


def _torch_loss(P, slots, weights, nrm, w, lam):
    """Float64 torch model: grid gather with the encoder's slots and
    trilinear weights, SH, bias-free ReLU MLP, vMF mixture, -lam log q;
    activations rounded to FP16 as stored."""
    grid = torch.tensor(G.f16(P.grid), dtype=torch.float64, requires_grad=True)
    Ws = [torch.tensor(G.f16(W), dtype=torch.float64, requires_grad=True) for W in P.W]
    feats = []
    for lv in range(G.L_LEVELS):
        s = torch.tensor(slots[lv])
        wt = torch.tensor(weights[lv], dtype=torch.float64)
        feats.append((grid[lv][s] * wt[..., None]).sum(1))
    enc = _Half.apply(torch.cat(feats + [torch.tensor(G.sh4(nrm), dtype=torch.float64)], 1))
    h = enc
    for k, W in enumerate(Ws):
        h = h @ W
        if k < 3:
            h = _Half.apply(torch.relu(h))
    raw = h[:, :40].reshape(-1, 8, 5)
    m = raw[..., 0:3]
    mu = m / torch.sqrt((m * m).sum(-1, keepdim=True) + 1e-12)
    kap = torch.exp(torch.clamp(raw[..., 3], G.KAPPA_LOG_MIN, G.KAPPA_LOG_MAX))
    pi = torch.softmax(raw[..., 4], -1)
    wt = torch.tensor(w, dtype=torch.float64)
    dots = (mu * wt[:, None, :]).sum(-1)
    lobes = kap / (2 * math.pi * -torch.expm1(-2 * kap)) * torch.exp(kap * (dots - 1))
    q = (pi * lobes).sum(-1)
    loss = (-torch.tensor(lam, dtype=torch.float64) * torch.log(q)).mean()
    loss.backward()
    return float(loss.detach()), grid.grad.numpy(), [W.grad.numpy() for W in Ws]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_backprop_matches_autograd():
    rng = np.random.default_rng(11)
    P = _params(2, 12, grid_scale=0.3)
    n = 512
    p01 = rng.random((n, 3)).astype(np.float32)
    nrm = _unit(rng, n)
    w = _unit(rng, n)
    lam = rng.uniform(0.5, 1.5, n).astype(np.float32)
    loss, gg, gW = G.train_step_grads(P, p01, nrm, w, lam)
    enc, slots, weights = G.encode(P, p01, nrm)
    out, _ = G.mlp_forward(P, enc)
    _, g_out = G.loss_and_grad_out(out, w, lam)
    assert np.abs(g_out).max() < G.GRAD_CLIP   # the clip is inactive on this batch
    tl, tg, tW = _torch_loss(P, slots, weights, nrm, w, lam)
    assert abs(loss - tl) <= 1e-3 * abs(tl) + 1e-4
    for k in range(4):
        assert _rel(gW[k], tW[k]) < 1e-3, k
    assert _rel(gg, tg) < 1e-3


def test_loss_gradient_of_outputs_matches_autograd():
    """Output-gradient formulas, including the kappa clamp (zero gradient
    outside [-4, 8]) and near-isotropic lobes."""
    rng = np.random.default_rng(12)
    n = 256
    out = rng.normal(size=(n, G.N_OUT)).astype(np.float32)
    r = out[:, :40].reshape(n, 8, 5)
    r[:, :, 3] = rng.uniform(-6, 10, (n, 8))
    w = _unit(rng, n)
    lam = rng.uniform(0.1, 3.0, n).astype(np.float32)
    loss, g = G.loss_and_grad_out(out, w, lam)
    t = torch.tensor(out, dtype=torch.float64, requires_grad=True)
    raw = t[:, :40].reshape(-1, 8, 5)
    m = raw[..., 0:3]
    mu = m / torch.sqrt((m * m).sum(-1, keepdim=True) + 1e-12)
    kap = torch.exp(torch.clamp(raw[..., 3], G.KAPPA_LOG_MIN, G.KAPPA_LOG_MAX))
    pi = torch.softmax(raw[..., 4], -1)
    wt = torch.tensor(w, dtype=torch.float64)
    lobes = kap / (2 * math.pi * -torch.expm1(-2 * kap)) * torch.exp(kap * ((mu * wt[:, None]).sum(-1) - 1))
    q = torch.clamp((pi * lobes).sum(-1), min=1e-20)
    tl = -torch.tensor(lam, dtype=torch.float64) * torch.log(q)
    tl.sum().backward()
    ref = np.clip(t.grad.numpy(), -G.GRAD_CLIP, G.GRAD_CLIP)
    assert np.allclose(loss, tl.detach().numpy(), rtol=1e-4, atol=1e-4)
    assert np.abs(g - ref).max() <= 1e-3 * np.abs(ref).max() + 1e-5


# ---- Adam -------------------------------------------------------------------------------

def test_adam_constant_gradient_closed_form():
    rng = np.random.default_rng(4)
    p = rng.normal(size=1000).astype(np.float32)
    g = (rng.choice([-1.0, 1.0], 1000) * 10.0 ** rng.uniform(-3, 1, 1000)).astype(np.float32)
    opt = G.Adam(lr=5e-3)
    p0 = p.copy()
    for t in range(1, 6):
        opt.t = t
        opt.step("p", p, g)
        # bias-corrected moments of a constant gradient are g and g^2
        assert np.allclose(p, p0 - t * 5e-3 * np.sign(g), atol=1e-5)
    # (without bias correction step 2 would already move 0.19 / sqrt(0.0199) = 1.35 lr)


# ---- normal encoding: real spherical harmonics -----------------------------------------

def _real_sh_scipy(nrm):
    """Real SH of degree <= 3 from scipy's complex Y_l^m (Condon-Shortley
    phase included): index l*l + l + m; m > 0: sqrt(2) Re Y_l^m, m < 0:
    sqrt(2) Im Y_l^|m|, m = 0: Y_l^0."""
    from scipy.special import sph_harm_y

    theta = np.arccos(np.clip(nrm[:, 2].astype(np.float64), -1.0, 1.0))
    phi = np.arctan2(nrm[:, 1].astype(np.float64), nrm[:, 0].astype(np.float64))
    out = np.zeros((len(nrm), 16))
    for l in range(4):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            out[:, l * l + l + m] = y.real if m == 0 else math.sqrt(2.0) * (y.real if m > 0 else y.imag)
    return out


def test_sh4_matches_scipy_real_basis():
    rng = np.random.default_rng(21)
    nrm = _unit(rng, 5000).astype(np.float64)
    got = G.sh4(nrm.astype(np.float32)).astype(np.float64)
    want = _real_sh_scipy(nrm)
    assert np.abs(got - want).max() < 2e-6


def test_sh4_orthonormal_by_quadrature():
    # Gauss-Legendre in z (16 nodes) x 32 equispaced azimuths integrates
    # every polynomial of degree <= 6 in (x, y, z) on the sphere exactly
    z, wz = np.polynomial.legendre.leggauss(16)
    phi = (np.arange(32) + 0.5) * (2 * math.pi / 32)
    Z, P = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - Z * Z)
    nrm = np.stack([r * np.cos(P), r * np.sin(P), Z], -1).reshape(-1, 3)
    w = (wz[:, None] * np.full(32, 2 * math.pi / 32)[None, :]).reshape(-1)
    Y = G.sh4(nrm.astype(np.float32)).astype(np.float64)
    gram = (Y * w[:, None]).T @ Y
    assert np.abs(gram - np.eye(16)).max() < 1e-5


def test_sh4_pole_values():
    for s in (1.0, -1.0):
        y = G.sh4(np.array([[0.0, 0.0, s]], np.float32))[0].astype(np.float64)
        for l in range(4):
            for m in range(-l, l + 1):
                want = s ** l * math.sqrt((2 * l + 1) / (4 * math.pi)) if m == 0 else 0.0
                assert abs(y[l * l + l + m] - want) < 1e-6, (s, l, m)


# ---- MLP forward --------------------------------------------------------------------------

def test_mlp_forward_positive_homogeneity_bitwise():
    rng = np.random.default_rng(31)
    P = _params(3, 12)
    enc = G.f16(rng.normal(0, 0.5, (300, G.N_IN)))
    out1, acts1 = G.mlp_forward(P, enc)
    out2, acts2 = G.mlp_forward(P, (enc * np.float32(2)).astype(np.float32))
    # rounding to FP16 commutes with doubling except for FP16 subnormals
    # (below 2^-14, whose spacing is fixed): rows holding one are compared
    # by tolerance
    sub = np.zeros(len(enc), bool)
    for a in acts1[1:4]:
        sub |= ((a > 0) & (a < 2.0 ** -14)).any(1)
    assert (~sub).sum() >= 250
    assert np.array_equal(out2[~sub], out1[~sub] * np.float32(2))
    for a1, a2 in zip(acts1[1:4], acts2[1:4]):
        assert np.array_equal(a2[~sub], a1[~sub] * np.float32(2))
    assert np.allclose(out2[sub], out1[sub] * 2, rtol=1e-3, atol=1e-5)
    # a negative scale is not a symmetry of ReLU: the map is not linear
    out3, _ = G.mlp_forward(P, (enc * np.float32(-1)).astype(np.float32))
    assert not np.allclose(out3, -out1)


def test_mlp_forward_identity_and_permutation_weights():
    rng = np.random.default_rng(32)
    P = _params(4, 12)
    sigma = rng.permutation(G.N_HID)
    W1 = np.zeros((G.N_IN, G.N_HID), np.float32)
    W1[np.arange(G.N_IN), np.arange(G.N_IN)] = 1.0
    W2 = np.zeros((G.N_HID, G.N_HID), np.float32)
    W2[np.arange(G.N_HID), sigma] = 1.0            # hidden unit i moves to sigma(i)
    W3 = np.eye(G.N_HID, dtype=np.float32)
    W4 = np.zeros((G.N_HID, G.N_OUT), np.float32)
    W4[np.arange(G.N_OUT), np.arange(G.N_OUT)] = 1.0
    P.W = [W1, W2, W3, W4]
    enc = G.f16(rng.normal(0, 1.0, (50, G.N_IN)))
    out, _ = G.mlp_forward(P, enc)
    want = np.zeros((50, G.N_OUT), np.float32)
    for i in range(G.N_IN):                      # input i -> hidden i -> hidden sigma(i) -> output sigma(i)
        if sigma[i] < G.N_OUT:
            want[:, sigma[i]] = np.maximum(enc[:, i], 0.0)
    assert np.array_equal(out, want)


def test_mlp_forward_one_hot_reads_row_of_weight_product():
    rng = np.random.default_rng(33)
    P = _params(5, 12)
    P.W = [np.abs(W).astype(np.float32) for W in P.W]   # ReLUs stay inactive
    prod = G.f16(P.W[0]).astype(np.float64)
    for W in P.W[1:]:
        prod = prod @ G.f16(W).astype(np.float64)
    for i in (0, 7, 31, 47):
        e = np.zeros((1, G.N_IN), np.float32)
        e[0, i] = 1.0
        out, _ = G.mlp_forward(P, e)
        # three FP16 roundings of the hidden activations: relative 3 * 2^-11
        assert np.allclose(out[0], prod[i], rtol=2e-3, atol=1e-6), i


# ---- hash slots and the unit-cube map ------------------------------------------------------

def test_hashed_corner_slots_python_integers():
    rng = np.random.default_rng(41)
    t = 14
    res = 2048
    c = rng.integers(0, res + 1, size=(500, 3))
    got = G.corner_index(c[:, 0], c[:, 1], c[:, 2], res, t)
    for k in range(len(c)):
        x, y, z = (int(v) for v in c[k])
        h = ((x * 1) ^ (y * 2654435761) ^ (z * 805459861)) % (1 << 32)
        assert got[k] == h % (1 << t)
    # a dense level is row-major x + n (y + n z)
    d = G.corner_index(np.array([3]), np.array([5]), np.array([7]), 16, t)
    assert d[0] == 3 + 17 * (5 + 17 * 7)


def test_unit_pos_maps_box_to_unit_cube():
    lo, hi = (-1.0, 2.0, 0.0), (3.0, 4.0, 0.0)   # a flat axis maps through a unit extent
    p = np.array([[-1.0, 2.0, 0.0], [3.0, 4.0, 0.5], [1.0, 3.0, -1.0], [5.0, 1.0, 2.0]], np.float32)
    q = G.unit_pos(p, lo, hi)
    assert np.array_equal(q[0], [0.0, 0.0, 0.0])
    assert np.array_equal(q[1], [G.POS_MAX, G.POS_MAX, 0.5])
    assert np.array_equal(q[2], [0.5, 0.5, 0.0])
    assert np.array_equal(q[3], [G.POS_MAX, 0.0, G.POS_MAX])
