"""FP64 CPU oracle of the wave-term surrogate (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference for the device
sampler, network forward pass and training gradients. The product
(bemuse_amd) never imports it.

The surrogate has no counterpart in the reference. What it follows from the
reference is its training target, HavelockGF (Kernels.cpp:342-412), taken
from the oracle's restatement (oracle.havelock). Definitions (the same as
bemuse_amd/csrc/bem_surrogate.h):

  features  f0 = 2 log1p(h)/log1p(200) - 1, f1 = 2 log1p(-v)/log1p(50) - 1,
            f2 = sin h, f3 = cos h
  scale     s = 1/(1 + dd) + e^v / sqrt(1 + h),  dd = sqrt(h^2 + v^2)
  singular  SG = 2 e^v ln((dd - v)/2) / (1 + dd^3)
            SGh = 2 (h/dd + h) / ((dd - v)(1 + dd^3))
  targets   ((Re GF - SG)/s, (Re GFh - SGh)/s); the imaginary parts are
            the closed forms Im GF = -2 pi e^v J0(h), Im GFh = 2 pi e^v J1(h)
            (Kernels.cpp:284-295, 327-338), evaluated exactly
  network   4 -> 128 -> 128 -> 128 -> 128 -> 2, tanh hidden, flat parameter
            layout W1 b1 W2 b2 W3 b3 W4 b4 W5 b5 (row-major (out, in))
  loss      sum over batch and channels of (y - t)^2 / (2 n)
  sampler   sample i: u0 = u(seed + 2i), u1 = u(seed + 2i + 1),
            u(k) = ((splitmix64(k) >> 11) + 1) 2^-53; sampling 0:
            h = 200 u0, v = -50 u1; sampling 1: h = 201^u0 - 1, v = 1 - 51^u1

Pinning (tests/test_surrogate.py):
  splitmix64         published splitmix64 output sequence (seed 1234567)
  singular terms     Re GF - SG and Re GFh - SGh stay bounded towards the
                     origin while Re GF and Re GFh diverge (a wrong
                     coefficient leaves a log or 1/dd residual)
  forward, gradient  torch float64 autograd of the same network, and central
                     finite differences of loss()
"""
from __future__ import annotations

import numpy as np

from . import bessel, havelock

IN, HID, OUT = 4, 128, 2
HMAX, VMIN = 200.0, -50.0
_SIZES = [(HID, IN), (HID, HID), (HID, HID), (HID, HID), (OUT, HID)]


def layout():
    """[(w_offset, b_offset, out, in)] per layer of the flat parameter vector."""
    out, o = [], 0
    for (n_out, n_in) in _SIZES:
        out.append((o, o + n_out * n_in, n_out, n_in))
        o += n_out * n_in + n_out
    return out


N_PARAMS = sum(a * b + a for a, b in _SIZES)

def splitmix64(x):
    """splitmix64 finaliser of x + golden gamma (vectorised, uint64 wrap-around)."""
    with np.errstate(over="ignore"):
        x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def unit_open0(key):
    return ((splitmix64(key) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53


def sample_points(seed, first, n, sampling=1):
    i = np.arange(first, first + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = np.uint64(seed) + np.uint64(2) * i
        u0 = unit_open0(k)
        u1 = unit_open0(k + np.uint64(1))
    if sampling == 1:
        return np.expm1(u0 * np.log1p(HMAX)), -np.expm1(u1 * np.log1p(-VMIN))
    return HMAX * u0, VMIN * u1


def features(h, v):
    h = np.asarray(h, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    return np.stack([2.0 * np.log1p(h) / np.log1p(HMAX) - 1.0,
                     2.0 * np.log1p(-v) / np.log1p(-VMIN) - 1.0,
                     np.sin(h), np.cos(h)], axis=-1)


def scale_singular(h, v):
    """(s, SG, SGh) at (h, v)."""
    h = np.asarray(h, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    dd = np.sqrt(h * h + v * v)
    inv1 = 1.0 / (1.0 + dd ** 3)
    s = 1.0 / (1.0 + dd) + np.exp(v) / np.sqrt(1.0 + h)
    sg = 2.0 * np.exp(v) * np.log(0.5 * (dd - v)) * inv1
    sgh = 2.0 * (h / dd + h) / (dd - v) * inv1
    return s, sg, sgh


def exact_gf(h, v):
    """(GF, GFh) complex from the oracle's HavelockGF at X = h, Y = 0, V = v
    (GFh = DGFX there; valid for h > 1e-6)."""
    out = havelock(np.asarray(h, dtype=np.float64), np.zeros(len(h)), np.asarray(v, dtype=np.float64))
    return out[:, 0], out[:, 1]


def targets(h, v):
    gf, gfh = exact_gf(h, v)
    s, sg, sgh = scale_singular(h, v)
    return np.stack([(gf.real - sg) / s, (gfh.real - sgh) / s], axis=-1)


def sample(seed, first, n, sampling=1):
    """(h, v, X, T) of the device sampler's samples first .. first + n - 1
    (X, T in float64; the device stores them rounded to float32)."""
    h, v = sample_points(seed, first, n, sampling)
    return h, v, features(h, v), targets(h, v)


def reconstruct(h, v, y):
    """(GF, GFh) complex from network outputs y (n x 2) and the exact
    imaginary parts."""
    s, sg, sgh = scale_singular(h, v)
    y = np.asarray(y, dtype=np.float64)
    j0, j1, _, _ = bessel(np.asarray(h, dtype=np.float64))
    E = 2.0 * np.pi * np.exp(np.asarray(v, dtype=np.float64))
    return (sg + s * y[:, 0]) - 1j * (E * j0), (sgh + s * y[:, 1]) + 1j * (E * j1)


def init_params(rng, scale=1.0):
    """Glorot-uniform weights (times scale) and small random biases."""
    p = np.zeros(N_PARAMS)
    for (wo, bo, n_out, n_in) in layout():
        a = np.sqrt(6.0 / (n_in + n_out)) * scale
        p[wo:wo + n_out * n_in] = rng.uniform(-a, a, n_out * n_in)
        p[bo:bo + n_out] = rng.uniform(-0.1, 0.1, n_out)
    return p


def _unpack(p):
    return [(p[wo:bo].reshape(n_out, n_in), p[bo:bo + n_out]) for (wo, bo, n_out, n_in) in layout()]


def forward(p, X, keep=False):
    """Network outputs (n x 4); keep=True also returns the activations."""
    layers = _unpack(np.asarray(p, dtype=np.float64))
    a = np.asarray(X, dtype=np.float64)
    acts = [a]
    for k, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = np.tanh(z) if k < len(layers) - 1 else z
        acts.append(a)
    return (a, acts) if keep else a


def loss(p, X, T):
    y = forward(p, X)
    return float(np.sum((y - np.asarray(T, dtype=np.float64)) ** 2) / (OUT * len(y)))


def loss_grad(p, X, T):
    """(loss, gradient) by reverse-mode differentiation written out layer by layer."""
    p = np.asarray(p, dtype=np.float64)
    layers = _unpack(p)
    y, acts = forward(p, X, keep=True)
    n = len(y)
    e = y - np.asarray(T, dtype=np.float64)
    L = float(np.sum(e * e) / (OUT * n))
    g = np.zeros_like(p)
    d = 2.0 * e / (OUT * n)          # dL/dz of the output layer
    lay = layout()
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        wo, bo, n_out, n_in = lay[k]
        a_in = acts[k]
        g[wo:bo] = (d.T @ a_in).ravel()
        g[bo:bo + n_out] = d.sum(axis=0)
        if k > 0:
            d = (d @ W) * (1.0 - a_in * a_in)
    return L, g


def wave_matrices(geom, K, gf=None):
    """Wave influence sums ws (S part) and wd (D part), complex N x N, of the
    geometry returned by oracle.Body.geometry(), as FreeSurface_Wave_Kernel +
    CInf_SingleDoubleLayerQuad sum them (Panel.cpp:601-636, 776-794): receiver
    R = BC point, nodes q of the mirrored panel S,
      ws += w_q G K/4pi,  wd += w_q (grad G . n_q) K^2/4pi,
      grad G = (GFh X/h, GFh Y/h, GF) (0 horizontal part for h <= 1e-6).
    gf(h, v) -> (GF, GFh) replaces HavelockGF; None uses the oracle's exact
    kernel."""
    bc, qpos, qnrm, qw, nq = geom["bc"], geom["qpos"], geom["qnrm"], geom["qw"], geom["nq"]
    N = len(bc)
    ws = np.zeros((N, N), dtype=np.complex128)
    wd = np.zeros((N, N), dtype=np.complex128)
    s1, s2 = K / (4.0 * np.pi), K * K / (4.0 * np.pi)
    for q in range(4):
        cols = np.nonzero(nq > q)[0]
        if len(cols) == 0:
            continue
        d = bc[:, None, :] - qpos[None, cols, q, :]
        X, Y, V = (K * d[..., 0]).ravel(), (K * d[..., 1]).ravel(), (K * d[..., 2]).ravel()
        h = np.sqrt(X * X + Y * Y)
        if gf is None:
            out = havelock(X, Y, V)
            G, GX, GY, GZ = out[:, 0], out[:, 1], out[:, 2], out[:, 3]
        else:
            G, Gh = gf(h, V)
            big = h > 1e-6
            cx = np.where(big, X / np.where(big, h, 1.0), 0.0)
            cy = np.where(big, Y / np.where(big, h, 1.0), 0.0)
            GX, GY, GZ = Gh * cx, Gh * cy, G
        n = qnrm[cols, q, :]
        w = qw[cols, q][None, :]
        G = G.reshape(N, len(cols))
        T = (GX.reshape(N, -1) * n[None, :, 0] + GY.reshape(N, -1) * n[None, :, 1]
             + GZ.reshape(N, -1) * n[None, :, 2])
        ws[:, cols] += w * (G * s1)
        wd[:, cols] += w * (T * s2)
    return ws, wd


def surrogate_gf(params):
    """gf(h, v) of the network with float32 parameters `params`: features
    rounded to float32 as on the device, network in float64; the exact kernel
    outside the domain h <= 200, -50 <= v <= 0."""
    p = np.asarray(params, dtype=np.float32).astype(np.float64)

    def gf(h, v):
        h = np.asarray(h, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        inside = (h <= HMAX) & (v >= VMIN) & (v <= 0.0) & ((h > 0.0) | (v < 0.0))
        G = np.zeros(len(h), dtype=np.complex128)
        Gh = np.zeros(len(h), dtype=np.complex128)
        if inside.any():
            hi, vi = h[inside], v[inside]
            y = forward(p, features(hi, vi).astype(np.float32).astype(np.float64))
            G[inside], Gh[inside] = reconstruct(hi, vi, y)
        if (~inside).any():
            ho, vo = h[~inside], v[~inside]
            G[~inside], Gh[~inside] = exact_gf(ho, vo)
        return G, Gh

    return gf
