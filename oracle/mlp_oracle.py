"""CPU oracle of the learned Davidson preconditioner — TEST INFRASTRUCTURE ONLY.

numpy (FP64) restatement of the network of rest_tensors_amd/csrc/mlp.hpp:
basis and features, forward pass, the squared-cosine misalignment loss of
the subspace-projected correction and its analytic gradient.  Used to check
the GPU forward/backward kernels (which run the network in FP32) and, through
central finite differences, to pin its own analytic gradient.  No reference
counterpart: the learned preconditioner is this build's addition to
davidson.rs.  The product never imports this module.

Parameter vector layout (837 floats): W1[64][7], c1[64], W2[5][64], c2[5].
"""
from __future__ import annotations

import numpy as np

F, H, O = 7, 64, 5
NPARAM = H * F + H + O * H + O
SHIFT_SCALE = 0.25


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[: H * F].reshape(H, F)
    c1 = p[H * F: H * F + H]
    W2 = p[H * F + H: H * F + H + O * H].reshape(O, H)
    c2 = p[H * F + H + O * H:]
    return W1, c1, W2, c2


def safe(h):
    return np.where(np.abs(h) < 1e-8, 1e-8, h)


def inputs(r, x0, hd, ek, e0, a=None):
    """basis (n, 5) and features (n, 7) for one sample (a: the auxiliary
    correction, or None for a zero basis vector)."""
    n = len(r)
    hk, h0 = hd - ek, hd - e0
    gk, g0 = r / safe(hk), r / safe(h0)
    a = np.zeros(n) if a is None else a
    rms = lambda v: np.sqrt(np.sum(v * v) / n)
    rr, rk, r0, ra = rms(r), rms(gk), rms(g0), rms(a)
    b = np.stack([gk / rk if rk > 0 else 0 * gk, g0 / r0 if r0 > 0 else 0 * g0,
                  x0 * np.sqrt(n), r / rr if rr > 0 else 0 * r,
                  a / ra if ra > 0 else 0 * a], axis=1)
    f = np.concatenate([np.arcsinh(b), np.tanh(hk / SHIFT_SCALE)[:, None],
                        np.tanh(h0 / SHIFT_SCALE)[:, None]], axis=1)
    return b, f


def forward(p, r, x0, hd, ek, e0, a=None):
    W1, c1, W2, c2 = unpack(p)
    b, f = inputs(r, x0, hd, ek, e0, a)
    z = np.tanh(f @ W1.T + c1)
    w = z @ W2.T + c2
    return np.sum(w * b, axis=1), (b, f, z, w)


def loss_grad(p, samples, taus, wts, Qs=None):
    """Mean weighted misalignment over samples [(r, x0, hd, ek, e0[, a])] and
    its gradient with respect to p.  Qs[j] (n x m orthonormal, or None): the
    subspace projected out of sample j's correction before comparing."""
    W1, c1, W2, c2 = unpack(p)
    Wsum = float(np.sum(wts))
    L = 0.0
    g = np.zeros(NPARAM)
    for j, (smp, tau, wt) in enumerate(zip(samples, taus, wts)):
        if wt <= 0:
            continue
        t, (b, f, z, w) = forward(p, *smp)
        Q = None if Qs is None else Qs[j]
        tp = t if Q is None else t - Q @ (Q.T @ t)
        a, bb, cc = tp @ tau, tp @ tp, tau @ tau
        wj = wt / Wsum
        L += wj * (1 - a * a / (bb * cc))
        gt = wj * (-2 * a * tau / (bb * cc) + 2 * a * a * tp / (bb * bb * cc))
        dw = gt[:, None] * b                   # (n, 4)
        dz = dw @ W2                           # (n, 64)
        du = dz * (1 - z * z)
        gW1 = du.T @ f
        gc1 = du.sum(0)
        gW2 = dw.T @ z
        gc2 = dw.sum(0)
        g += np.concatenate([gW1.ravel(), gc1, gW2.ravel(), gc2])
    return L, g


def init_params(seed: int = 1):
    """Same initial network as rt_mlp_create: W1 uniform(+-sqrt(3/F)) from the
    SplitMix64 hash, c1 = 0, W2 = 0, c2 = (0, 1, 0, 0, 0)."""
    from oracle.ri_oracle import splitmix64, rng_key
    key = rng_key(seed)
    with np.errstate(over="ignore"):
        h = splitmix64(key + np.arange(H * F, dtype=np.uint64))
    u = (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    p = np.zeros(NPARAM, dtype=np.float32)
    p[: H * F] = ((2 * u - 1) * np.sqrt(3.0 / F)).astype(np.float32)
    p[H * F + H + O * H + 1] = 1.0
    return p
