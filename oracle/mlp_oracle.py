"""CPU oracle of the learned NoSlipWall warm start's network -- test infrastructure
only (tests/ and the GPU gradient check use it as the checker; the product path is
the HIP kernels in rbc3d_amd/csrc/mlp.hip).

The network has no counterpart in the reference (its NoSlipWall starts GMRES from
zero, common/ModNoSlip.F90:117); this module restates the build's own definition:
9 -> 64 -> 64 -> 64 -> 3, GELU (tanh form, Hendrycks and Gimpel 2016), inputs
standardised as (x - mean) / std, outputs scaled per component, mean squared error
of the scaled increments, Adam (Kingma and Ba 2015).

`forward`/`loss_grad` with bf16=True round to BF16 (nearest even) at the same points
as the GPU kernels -- weights, the inputs of every matrix product (activations and
back-propagated gradients) -- and accumulate in FP64, so they differ from the GPU's
FP32 accumulation only by accumulation rounding.  With bf16=False everything is FP64;
that variant is pinned against central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 9, 64, 3
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * x * (1.0 + t)


def gelu_d(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def init_params(rng):
    """He-style initialisation: W ~ N(0, 2 / fan_in), b = 0; last layer scaled by 0.1."""
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shp) * np.sqrt(2.0 / shp[1])
        else:
            d[name] = np.zeros(shp)
    d["W4"] *= 0.1
    return pack(d).astype(np.float32)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(p, mean, std, scale, X, bf=True):
    """Returns the scaled-space outputs y (n,3) and the intermediates for backward."""
    P = unpack(p)
    W = {k: _r(P[k], bf) for k in ("W1", "W2", "W3", "W4")}
    if bf:
        xn = (np.asarray(X, np.float32) - np.asarray(mean, np.float32)) * \
            (np.float32(1.0) / np.asarray(std, np.float32))
    else:
        xn = (np.asarray(X, np.float64) - mean) / std
    h0 = _r(xn, bf)
    z1 = h0 @ W["W1"].T + P["b1"]
    h1 = _r(gelu(z1), bf)
    z2 = h1 @ W["W2"].T + P["b2"]
    h2 = _r(gelu(z2), bf)
    z3 = h2 @ W["W3"].T + P["b3"]
    h3 = _r(gelu(z3), bf)
    y = h3 @ W["W4"].T + P["b4"]
    return y, dict(W=W, h=(h0, h1, h2, h3), z=(z1, z2, z3))


def predict(p, mean, std, scale, X, bf=True):
    """The predicted increments (n,3) = scale * y."""
    return forward(p, mean, std, scale, X, bf)[0] * np.asarray(scale, np.float64)


def loss_grad(p, mean, std, scale, X, T, bf=True):
    """Loss mean((y - T/scale)^2) over the 3n outputs and its gradient (param layout)."""
    n = X.shape[0]
    y, c = forward(p, mean, std, scale, X, bf)
    e = y - np.asarray(T, np.float64) / np.asarray(scale, np.float64)
    loss = float((e * e).sum() / (3 * n))
    dy = 2.0 * e / (3 * n)
    W, (h0, h1, h2, h3), (z1, z2, z3) = c["W"], c["h"], c["z"]
    g = {}
    dyb = _r(dy, bf)
    g["b4"] = dy.sum(0)
    g["W4"] = dyb.T @ h3
    dz3 = (dyb @ W["W4"]) * gelu_d(z3)
    dz3b = _r(dz3, bf)
    g["b3"] = dz3.sum(0)
    g["W3"] = dz3b.T @ h2
    dz2 = (dz3b @ W["W3"]) * gelu_d(z2)
    dz2b = _r(dz2, bf)
    g["b2"] = dz2.sum(0)
    g["W2"] = dz2b.T @ h1
    dz1 = (dz2b @ W["W2"]) * gelu_d(z1)
    dz1b = _r(dz1, bf)
    g["b1"] = dz1.sum(0)
    g["W1"] = dz1b.T @ h0
    return loss, pack(g)


def adam(p, m1, m2, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (returns new p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
