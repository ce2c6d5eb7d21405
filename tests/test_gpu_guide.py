"""Path-guiding kernels against the CPU FP32 restatement (oracle/guide_ref.py):
network evaluation (hash-grid gather + FP16 MFMA MLP), the fused loss and
back-propagation (summed gradients), two Adam steps and the refreshed FP16
copies, mixture sampling and pdf queries, and an end-to-end fit of a
position-dependent vMF target."""
import math

import numpy as np
import pytest

from oracle import guide_ref as G

pytestmark = pytest.mark.gpu

T_LOG2 = 14   # 2 dense + 14 hashed levels


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _params(seed, grid_scale=0.3, w_scale=1.0):
    rng = np.random.default_rng(seed)
    P = G.GuideParams(seed=seed, t_log2=T_LOG2)
    P.grid = rng.uniform(-grid_scale, grid_scale, P.grid.shape).astype(np.float32)
    P.W = [(w * np.float32(w_scale)).astype(np.float32) for w in P.W]
    return P


def _flat_w(P):
    return np.concatenate([W.ravel() for W in P.W]).astype(np.float32)


def _split_w(w):
    out, o = [], 0
    for a, b in [(48, 64), (64, 64), (64, 64), (64, 48)]:
        out.append(w[o:o + a * b].reshape(a, b))
        o += a * b
    return out


@pytest.fixture
def guided():
    from xenon_amd import HipRenderer
    with HipRenderer(8, 8, 4, 2) as r:
        r.guide_enable((0, 0, 0), (1, 1, 1), t_log2=T_LOG2, learning_rate=5e-3, seed=3)
        yield r


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_guide_requires_enable():
    from xenon_amd import HipRenderer, XnhError
    with HipRenderer(8, 8) as r:
        with pytest.raises(XnhError):
            r.guide_eval(np.zeros((1, 3), np.float32), np.array([[0, 0, 1]], np.float32))


def test_guide_params_roundtrip(guided):
    P = _params(1)
    guided.guide_set_params(P.grid, _flat_w(P))
    g, w = guided.guide_params()
    assert np.array_equal(g, P.grid) and np.array_equal(w, _flat_w(P))


def test_guide_eval_matches_oracle(guided):
    rng = np.random.default_rng(7)
    P = _params(2)
    guided.guide_set_params(P.grid, _flat_w(P))
    n = 1000
    p = rng.random((n, 3)).astype(np.float32)
    p[:4] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [-3, 2, 0.25]]   # corners, centre, clamped
    nrm = _unit(rng, n)
    print(f"[guide] eval n={n}", flush=True)
    dev = guided.guide_eval(p, nrm)
    enc, _, _ = G.encode(P, p, nrm)
    ref, _ = G.mlp_forward(P, enc)
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.abs(dev - ref).max() <= 1e-2 * scale
    assert _rel(dev, ref) < 3e-3


def test_guide_gradients_match_oracle(guided):
    rng = np.random.default_rng(8)
    P = _params(3)
    guided.guide_set_params(P.grid, _flat_w(P))
    n = 2048
    p = rng.random((n, 3)).astype(np.float32)
    nrm, w = _unit(rng, n), _unit(rng, n)
    lam = rng.uniform(0.5, 1.5, n).astype(np.float32)
    print(f"[guide] gradients n={n}", flush=True)
    loss, gg, gw = guided.guide_train(p, nrm, w, lam, apply=False, want_grads=True)
    lam_s = (lam * np.float32(n / lam.astype(np.float64).sum())).astype(np.float32)
    rl, rg, rW = G.train_step_grads(P, p, nrm, w, lam_s)
    assert abs(loss - rl) <= 2e-3 * abs(rl) + 1e-3
    dW = _split_w(gw)
    for k in range(4):
        assert _rel(dW[k], rW[k] * n) < 2e-2, k
    assert _rel(gg, rg * n) < 2e-2
    # a second call starts from cleared gradients
    loss2, gg2, gw2 = guided.guide_train(p, nrm, w, lam, apply=False, want_grads=True)
    assert _rel(gw2, gw) < 1e-4 and _rel(gg2, gg) < 1e-4


def test_guide_adam_two_steps(guided):
    rng = np.random.default_rng(9)
    P = _params(4)
    guided.guide_set_params(P.grid, _flat_w(P))
    n = 4096
    p = rng.random((n, 3)).astype(np.float32)
    nrm, w = _unit(rng, n), _unit(rng, n)
    lam = rng.uniform(0.5, 1.5, n).astype(np.float32)
    print("[guide] adam two steps", flush=True)
    grads = []
    for _ in range(2):
        _, gg, gw = guided.guide_train(p, nrm, w, lam, apply=False, want_grads=True)
        grads.append((gg / n, gw / n))
        guided.guide_train(p, nrm, w, lam, apply=True)
    g2, w2 = guided.guide_params()
    opt = G.Adam(lr=5e-3, b1=0.9, b2=0.99, eps=1e-8)
    pg, pw = P.grid.copy(), _flat_w(P)
    for t, (gg, gw) in enumerate(grads, 1):
        opt.t = t
        opt.step("grid", pg, gg.astype(np.float32))
        opt.step("w", pw, gw.astype(np.float32))
    # The device recomputes each gradient inside the apply call; FP32 atomic
    # summation order makes it differ from the returned copy in the last
    # bits only, which moves an Adam step by far less than 2e-6 unless a
    # gradient sits at the noise floor (a sign flip: at most 2 lr per step).
    for dev, ref in ((g2, pg), (w2, pw)):
        err = np.abs(dev.ravel() - ref.ravel())
        assert np.mean(err <= 2e-6) > 0.9999
        assert err.max() <= 4 * 5e-3 + 1e-6
    # the FP16 copies used by the kernels were refreshed from the masters
    Pn = _params(4)
    Pn.grid, Pn.W = g2, _split_w(w2)
    q = rng.random((256, 3)).astype(np.float32)
    qn = _unit(rng, 256)
    ref, _ = G.mlp_forward(Pn, G.encode(Pn, q, qn)[0])
    assert _rel(guided.guide_eval(q, qn), ref) < 3e-3


def test_guide_query_matches_oracle(guided):
    rng = np.random.default_rng(10)
    P = _params(5, w_scale=0.5)
    guided.guide_set_params(P.grid, _flat_w(P))
    n = 20000
    p = rng.random((n, 3)).astype(np.float32)
    nrm = _unit(rng, n)
    print(f"[guide] query n={n}", flush=True)
    raw = guided.guide_eval(p, nrm)
    mu, kappa, pi, _ = G.decode(raw)
    d = _unit(rng, n)
    pdf = guided.guide_pdf(p, nrm, d)
    ref, _ = G.mixture_pdf(mu, kappa, pi, d)
    assert np.allclose(pdf, ref, rtol=2e-4, atol=1e-8)
    u = rng.random((n, 3)).astype(np.float32)
    ds, ps = guided.guide_sample(p, nrm, u)
    assert np.abs(np.linalg.norm(ds, axis=1) - 1).max() < 1e-4
    rs = G.sample_mixture(mu, kappa, pi, u[:, 0], u[:, 1], u[:, 2])
    cdf = np.cumsum(pi, 1)
    clear = np.abs(cdf - u[:, :1]).min(1) > 1e-5
    assert clear.mean() > 0.99
    assert np.abs(ds[clear] - rs[clear]).max() < 2e-3
    rp, _ = G.mixture_pdf(mu, kappa, pi, ds)
    assert np.allclose(ps, rp, rtol=2e-4, atol=1e-8)
    # importance identity of the device sampler/pdf pair: E_q[1/q] = 4 pi
    x = 1.0 / ps.astype(np.float64)
    assert abs(x.mean() - 4 * math.pi) < 4 * x.std() / math.sqrt(n)


def test_guide_fits_position_dependent_vmf(guided):
    """Samples drawn from vMF(mu(p), 20) with mu depending on position: the
    loss falls from the near-uniform value log(4 pi) toward the target's
    entropy (-0.16)."""
    rng = np.random.default_rng(11)
    guided.guide_enable((0, 0, 0), (1, 1, 1), t_log2=T_LOG2, learning_rate=1e-2, seed=5)
    kappa = 20.0

    def batch(n):
        p = rng.random((n, 3)).astype(np.float32)
        m = p - 0.5 + np.array([0, 0, 0.6], np.float32)
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        u = rng.random((3, n)).astype(np.float32)
        w = G.sample_mixture(np.repeat(m[:, None, :], 8, 1), np.full((n, 8), kappa, np.float32),
                             np.eye(8, dtype=np.float32)[np.zeros(n, int)], u[0], u[1], u[2])
        return p, _unit(rng, n), w
    losses = []
    for step in range(150):
        p, nrm, w = batch(8192)
        losses.append(guided.guide_train(p, nrm, w, np.ones(len(p), np.float32), apply=True))
    print(f"[guide] fit: first {losses[0]:.3f} last {np.mean(losses[-10:]):.3f}", flush=True)
    assert abs(losses[0] - math.log(4 * math.pi)) < 0.3
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-10:]) < 0.6
