"""CPU FP32 restatement of the path-guiding network (TEST INFRASTRUCTURE ONLY).

The reference renderer has no path guiding; this module is the checker for
the HIP guiding kernels (xenon_amd/csrc/device/guide.hip), written from the
published methods it combines:
  * multiresolution hash encoding (Mueller et al. 2022, "Instant Neural
    Graphics Primitives"): L levels, T entries per level, F features,
    geometric resolutions, spatial hash x*1 ^ y*2654435761 ^ z*805459861,
    dense indexing where the level fits, trilinear interpolation;
  * real spherical harmonics up to degree 3 for the normal;
  * a bias-free ReLU MLP (3 hidden layers of 64);
  * an 8-lobe von Mises-Fisher mixture (mean, concentration, softmax weight)
    with C(k) = k / (2 pi (1 - exp(-2k))) and the inverse-CDF sampler of
    Wenzel Jakob's "Numerically stable sampling of the von Mises-Fisher
    distribution on S2" (2012);
  * the online KL objective of Mueller et al. 2019 ("Neural Importance
    Sampling"): loss = -lambda * log q(w), lambda = luminance(L_in)/p(w);
  * Adam (Kingma & Ba 2015) with bias correction.
Parameters and activations are rounded to FP16 where the device stores them
in FP16; products accumulate in FP32.
"""
from __future__ import annotations

import math

import numpy as np

L_LEVELS = 16
T_LOG2 = 19
F_FEAT = 2
N_MIN = 16
N_MAX = 2048
N_IN = 48          # 32 grid features + 16 SH coefficients
N_HID = 64
N_OUT = 48         # 8 lobes x 5 raw outputs, padded
N_LOBES = 8
KAPPA_LOG_MIN, KAPPA_LOG_MAX = -4.0, 8.0
GRAD_CLIP = 1024.0  # per-sample output-gradient clip (keeps FP16 back-propagation finite)
PRIMES = (1, 2654435761, 805459861)


def level_resolutions(levels=L_LEVELS, nmin=N_MIN, nmax=N_MAX):
    b = math.exp((math.log(nmax) - math.log(nmin)) / (levels - 1))
    return np.array([int(math.floor(nmin * b ** l)) for l in range(levels)], np.int64)


def corner_index(ix, iy, iz, res, t_log2=T_LOG2):
    """Table slot of integer lattice corner (ix, iy, iz) at a level of
    resolution `res`: dense when (res+1)^3 fits the table, hashed otherwise."""
    T = 1 << t_log2
    n = res + 1
    if n ** 3 <= T:
        return (ix + iy * n + iz * n * n).astype(np.int64)
    h = (ix.astype(np.uint64) * np.uint64(PRIMES[0])) ^ (iy.astype(np.uint64) * np.uint64(PRIMES[1])) ^ \
        (iz.astype(np.uint64) * np.uint64(PRIMES[2]))
    return (h & np.uint64(0xFFFFFFFF) & np.uint64(T - 1)).astype(np.int64)


def sh4(n):
    """Real spherical harmonics, degrees 0..3 (16 values), of unit vectors."""
    x, y, z = n[:, 0], n[:, 1], n[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    out = np.empty((n.shape[0], 16), np.float32)
    out[:, 0] = 0.28209479177387814
    out[:, 1] = -0.48860251190291987 * y
    out[:, 2] = 0.48860251190291987 * z
    out[:, 3] = -0.48860251190291987 * x
    out[:, 4] = 1.0925484305920792 * x * y
    out[:, 5] = -1.0925484305920792 * y * z
    out[:, 6] = 0.94617469575755997 * zz - 0.31539156525251999
    out[:, 7] = -1.0925484305920792 * x * z
    out[:, 8] = 0.54627421529603959 * (xx - yy)
    out[:, 9] = 0.59004358992664352 * y * (-3.0 * xx + yy)
    out[:, 10] = 2.8906114426405538 * x * y * z
    out[:, 11] = 0.45704579946446572 * y * (1.0 - 5.0 * zz)
    out[:, 12] = 0.3731763325901154 * z * (5.0 * zz - 3.0)
    out[:, 13] = 0.45704579946446572 * x * (1.0 - 5.0 * zz)
    out[:, 14] = 1.4453057213202769 * z * (xx - yy)
    out[:, 15] = 0.59004358992664352 * x * (-xx + 3.0 * yy)
    return out


def f16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


class GuideParams:
    def __init__(self, seed=0, t_log2=T_LOG2):
        rng = np.random.default_rng(seed)
        self.t_log2 = t_log2
        self.res = level_resolutions()
        self.grid = rng.uniform(-1e-4, 1e-4, (L_LEVELS, 1 << t_log2, F_FEAT)).astype(np.float32)
        dims = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]
        self.W = [rng.uniform(-1, 1, d).astype(np.float32) * np.float32(math.sqrt(6.0 / (d[0] + d[1])))
                  for d in dims]


POS_MAX = np.float32(0.99999994)  # largest float below 1


def unit_pos(p, lo, hi):
    """World positions -> the unit cube of the grid, clamped to [0, POS_MAX]."""
    lo = np.asarray(lo, np.float32)
    e = np.asarray(hi, np.float32) - lo
    inv = np.where(e > 0, np.float32(1) / np.where(e > 0, e, 1).astype(np.float32), np.float32(1)).astype(np.float32)
    q = ((np.asarray(p, np.float32) - lo) * inv).astype(np.float32)
    return np.clip(q, np.float32(0), POS_MAX).astype(np.float32)


def encode(P, p01, n):
    """Returns (enc (N,48) fp16-rounded, per-level corner slots and weights)."""
    p01 = np.clip(np.asarray(p01, np.float32), np.float32(0), POS_MAX)
    N = p01.shape[0]
    feats = np.zeros((N, L_LEVELS * F_FEAT), np.float32)
    slots, weights = [], []
    g16 = f16(P.grid)
    for l, res in enumerate(P.res):
        x = p01.astype(np.float32) * np.float32(res)
        i0 = np.floor(x).astype(np.int64)
        fr = (x - i0).astype(np.float32)
        lv_s, lv_w = [], []
        acc = np.zeros((N, F_FEAT), np.float32)
        for c in range(8):
            dx, dy, dz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            w = ((fr[:, 0] if dx else 1 - fr[:, 0]) * (fr[:, 1] if dy else 1 - fr[:, 1])
                 * (fr[:, 2] if dz else 1 - fr[:, 2])).astype(np.float32)
            s = corner_index(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, int(res), P.t_log2)
            acc += w[:, None] * g16[l, s]
            lv_s.append(s)
            lv_w.append(w)
        feats[:, 2 * l:2 * l + 2] = acc
        slots.append(np.stack(lv_s, 1))
        weights.append(np.stack(lv_w, 1))
    enc = np.concatenate([feats, sh4(n)], 1)
    return f16(enc), slots, weights


def mlp_forward(P, enc):
    acts = [enc]
    h = enc
    for k, W in enumerate(P.W):
        z = h @ f16(W)
        h = f16(np.maximum(z, 0.0)) if k < 3 else z.astype(np.float32)
        acts.append(h)
    return h, acts


def decode(out):
    """Raw outputs (N,48) -> mu (N,8,3), kappa (N,8), pi (N,8), |m| (N,8)."""
    raw = out[:, :N_LOBES * 5].reshape(-1, N_LOBES, 5)
    m = raw[:, :, 0:3]
    nm = np.sqrt((m * m).sum(2) + np.float32(1e-12)).astype(np.float32)
    mu = m / nm[:, :, None]
    lk = np.clip(raw[:, :, 3], KAPPA_LOG_MIN, KAPPA_LOG_MAX)
    kappa = np.exp(lk).astype(np.float32)
    a = raw[:, :, 4]
    a = a - a.max(1, keepdims=True)
    e = np.exp(a)
    pi = (e / e.sum(1, keepdims=True)).astype(np.float32)
    return mu.astype(np.float32), kappa, pi, nm


def vmf_norm(k):
    """kappa / (2 pi (1 - exp(-2 kappa))), stable for small kappa."""
    k = np.asarray(k, np.float64)
    return (k / (2 * np.pi * -np.expm1(-2 * k))).astype(np.float32)


def mixture_pdf(mu, kappa, pi, w):
    dots = (mu * w[:, None, :]).sum(2)
    lobes = vmf_norm(kappa) * np.exp(kappa * (dots - 1.0))
    return (pi * lobes).sum(1).astype(np.float32), lobes.astype(np.float32)


def loss_and_grad_out(out, w, lam):
    """Per-sample loss -lam*log q(w) and its gradient wrt the raw outputs."""
    mu, kappa, pi, nm = decode(out)
    q, lobes = mixture_pdf(mu, kappa, pi, w)
    q = np.maximum(q, np.float32(1e-20))
    loss = -lam * np.log(q)
    r = pi * lobes / q[:, None]                        # responsibilities
    g = np.zeros_like(out)
    G = g[:, :N_LOBES * 5].reshape(-1, N_LOBES, 5)
    dots = (mu * w[:, None, :]).sum(2)
    k = kappa.astype(np.float64)
    dlogv_dk = (1.0 / k - 2.0 * np.exp(-2 * k) / -np.expm1(-2 * k) + (dots - 1.0)).astype(np.float32)
    raw_lk = out[:, :N_LOBES * 5].reshape(-1, N_LOBES, 5)[:, :, 3]
    inside = ((raw_lk > KAPPA_LOG_MIN) & (raw_lk < KAPPA_LOG_MAX)).astype(np.float32)
    # d(-lam log q)/d raw
    G[:, :, 4] = -lam[:, None] * (r - pi)
    G[:, :, 3] = -lam[:, None] * r * dlogv_dk * kappa * inside
    tang = w[:, None, :] - dots[:, :, None] * mu
    G[:, :, 0:3] = -lam[:, None, None] * r[:, :, None] * kappa[:, :, None] * tang / nm[:, :, None]
    np.clip(g, -GRAD_CLIP, GRAD_CLIP, out=g)
    return loss.astype(np.float32), g


def mlp_backward(P, acts, g_out):
    """Gradients wrt the weights and the encoding (FP32)."""
    gW = [None] * 4
    g = g_out
    for k in range(3, -1, -1):
        gW[k] = acts[k].T @ g
        g = g @ f16(P.W[k]).T
        if k > 0:
            g = g * (acts[k] > 0)
    return gW, g


def grid_backward(P, g_enc, slots, weights):
    gg = np.zeros_like(P.grid)
    for l in range(L_LEVELS):
        gl = g_enc[:, 2 * l:2 * l + 2]
        for c in range(8):
            np.add.at(gg[l], slots[l][:, c], weights[l][:, c, None] * gl)
    return gg


def train_step_grads(P, p01, n, w, lam):
    """Mean loss and gradients (grid, [W1..W4]) over a batch."""
    enc, slots, weights = encode(P, p01, n)
    out, acts = mlp_forward(P, enc)
    loss, g_out = loss_and_grad_out(out, w, lam)
    N = p01.shape[0]
    g_out = g_out / np.float32(N)
    gW, g_enc = mlp_backward(P, acts, g_out)
    gg = grid_backward(P, g_enc[:, :L_LEVELS * F_FEAT], slots, weights)
    return float(loss.mean()), gg, gW


class Adam:
    def __init__(self, lr=5e-3, b1=0.9, b2=0.99, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps, self.t = lr, b1, b2, eps, 0
        self.m, self.v = {}, {}

    def step(self, name, p, g):
        if name not in self.m:
            self.m[name] = np.zeros_like(p)
            self.v[name] = np.zeros_like(p)
        m, v = self.m[name], self.v[name]
        m *= self.b1
        m += (1 - self.b1) * g
        v *= self.b2
        v += (1 - self.b2) * g * g
        mh = m / (1 - self.b1 ** self.t)
        vh = v / (1 - self.b2 ** self.t)
        p -= self.lr * mh / (np.sqrt(vh) + self.eps)


def sample_mixture(mu, kappa, pi, u0, u1, u2):
    """Inverse-CDF lobe choice and vMF sample per row."""
    cdf = np.cumsum(pi, 1)
    j = np.minimum((cdf < u0[:, None]).sum(1), N_LOBES - 1)
    idx = np.arange(mu.shape[0])
    m, k = mu[idx, j], kappa[idx, j].astype(np.float64)
    wz = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * k)) / k
    st = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    phi = 2 * np.pi * u2
    local = np.stack([st * np.cos(phi), st * np.sin(phi), wz], 1)
    s = np.where(m[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + m[:, 2])
    b = m[:, 0] * m[:, 1] * a
    t1 = np.stack([1 + s * m[:, 0] ** 2 * a, s * b, -s * m[:, 0]], 1)
    t2 = np.stack([b, s + m[:, 1] ** 2 * a, -m[:, 1]], 1)
    return (local[:, :1] * t1 + local[:, 1:2] * t2 + local[:, 2:3] * m).astype(np.float32)
