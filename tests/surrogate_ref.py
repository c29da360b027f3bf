"""CPU reference of the surrogate (test infrastructure).

- ``inputs``: the network inputs and closed-form baseline, restated in numpy
  from their definition (Hull-White bond-option formula on the tree's dates,
  lattice phase of the strike).
- ``forward_bf16``: the network forward with the GPU's rounding points
  emulated (BF16 round-to-nearest-even of the hidden activations and of the
  256x256 weights; FP32 input and output layers).
- ``backward``: the loss gradient in float64 through the same rounded
  forward (the GPU additionally rounds the back-propagated dz to BF16).
- ``forward_exact`` / ``grad_exact``: the unrounded float64 network, whose
  backward is pinned by finite differences.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc

OFF = {"W1": 0, "b1": 2048, "W2": 2304, "b2": 67840, "W3": 68096, "b3": 133632,
       "W4": 133888, "b4": 199424, "W5": 199680, "b5": 199936}
SHAPE = {"W1": (256, 8), "b1": (256,), "W2": (256, 256), "b2": (256,), "W3": (256, 256), "b3": (256,),
         "W4": (256, 256), "b4": (256,), "W5": (1, 256), "b5": (1,)}
NPAR = 199937
RES = 0.01


def unpack(flat):
    return {k: np.asarray(flat[o:o + int(np.prod(SHAPE[k]))]).reshape(SHAPE[k]) for k, o in OFF.items()}


def pack(d, dtype=np.float64):
    out = np.zeros(NPAR, dtype)
    for k, o in OFF.items():
        out[o:o + int(np.prod(SHAPE[k]))] = np.asarray(d[k]).ravel()
    return out


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward_bf16(flat, x):
    """GPU-shaped forward; returns y and the intermediates backward needs."""
    p = {k: v.astype(np.float32) for k, v in unpack(flat).items()}
    x = np.asarray(x, np.float32)
    z1 = (x.astype(np.float64) @ p["W1"].T.astype(np.float64) + p["b1"]).astype(np.float32)
    a1 = bf16(silu(z1.astype(np.float64)).astype(np.float32))
    zs, acts = [z1], [a1]
    a = a1
    for name in ("2", "3", "4"):
        W = bf16(p["W" + name]).astype(np.float64)
        z = (a.astype(np.float64) @ W.T + p["b" + name]).astype(np.float32)
        zs.append(z)
        h = silu(z.astype(np.float64)).astype(np.float32)
        a = bf16(h) if name != "4" else h
        acts.append(a)
    y = acts[-1].astype(np.float64) @ p["W5"][0].astype(np.float64) + float(p["b5"][0])
    return y.astype(np.float32), {"x": x, "z": zs, "a": acts, "p": p}


def backward(cache, y, t):
    """Gradient (float64, flat) of mean (y - t)^2 through the rounded forward."""
    p, x, zs, acts = cache["p"], cache["x"].astype(np.float64), cache["z"], cache["a"]
    n = x.shape[0]
    dy = 2.0 * (y.astype(np.float64) - t) / n
    g = {"W5": (dy @ acts[3].astype(np.float64))[None, :], "b5": np.array([dy.sum()])}
    dz = np.outer(dy, p["W5"][0].astype(np.float64)) * dsilu(zs[3].astype(np.float64))
    for li in (4, 3, 2):
        aprev = acts[li - 2].astype(np.float64)
        g["W%d" % li] = dz.T @ aprev
        g["b%d" % li] = dz.sum(0)
        da = dz @ bf16(p["W%d" % li]).astype(np.float64)
        dz = da * dsilu(zs[li - 2].astype(np.float64))
    g["W1"] = dz.T @ x
    g["b1"] = dz.sum(0)
    return pack(g)


def forward_exact(flat, x):
    p = unpack(np.asarray(flat, np.float64))
    h = np.asarray(x, np.float64)
    zs, hs = [], [h]
    for li in (1, 2, 3, 4):
        z = h @ p["W%d" % li].T + p["b%d" % li]
        h = silu(z)
        zs.append(z)
        hs.append(h)
    return h @ p["W5"][0] + p["b5"][0], (p, zs, hs)


def grad_exact(flat, x, t):
    y, (p, zs, hs) = forward_exact(flat, x)
    n = len(y)
    dy = 2.0 * (y - t) / n
    g = {"W5": (dy @ hs[4])[None, :], "b5": np.array([dy.sum()])}
    dz = np.outer(dy, p["W5"][0]) * dsilu(zs[3])
    for li in (4, 3, 2, 1):
        g["W%d" % li] = dz.T @ hs[li - 1]
        g["b%d" % li] = dz.sum(0)
        if li > 1:
            dz = (dz @ p["W%d" % li]) * dsilu(zs[li - 2])
    return pack(g)


def loss_exact(flat, x, t):
    y, _ = forward_exact(flat, x)
    return float(np.mean((y - t) ** 2))


# ------------------------------------------------------------------ inputs
def _yield_at(t, rates, days, tu=365):
    d = math.floor(t * tu + 0.5)  # round() half away from zero, t >= 0
    if d <= days[0]:
        return rates[0]
    if d > days[-1]:
        return rates[-1]
    i = int(np.searchsorted(days, d))
    return rates[i - 1] + (d - days[i - 1]) / (days[i] - days[i - 1]) * (rates[i] - rates[i - 1])


def closed_form(X, T, S, a, sg, typ, y):
    """Hull-White price of a European option at T on a zero-coupon bond
    maturing at S (face 100), on the curve's discount bonds."""
    PT = math.exp(-_yield_at(T, y.rates, y.days) * T)
    PS = math.exp(-_yield_at(S, y.rates, y.days) * S)
    Bc = (1 - math.exp(-a * (S - T))) / a
    sp = sg * Bc * math.sqrt((1 - math.exp(-2 * a * T)) / (2 * a))
    h = math.log(100 * PS / (X * PT)) / sp + 0.5 * sp
    r2 = 1 / math.sqrt(2)
    call = 100 * PS * 0.5 * erfc(-h * r2) - X * PT * 0.5 * erfc(-(h - sp) * r2)
    put = X * PT * 0.5 * erfc((h - sp) * r2) - 100 * PS * 0.5 * erfc(h * r2)
    return put if typ == 1 else call


def inputs(opts, i, y):
    """(x[8], baseline price, in_box) of option i."""
    X, Tb, L, a, sg = (float(opts.strike[i]), float(opts.maturity[i]), float(opts.length[i]),
                       float(opts.a[i]), float(opts.sigma[i]))
    tu, st, ty = int(opts.term_unit[i]), int(opts.term_steps[i]), int(opts.type[i])
    box = (tu == 365 and st in (12, 36) and 60 <= X <= 100 and 1 <= L <= 30 and 1 <= Tb - L <= 10
           and 0.01 <= a <= 0.2 and 0.005 <= sg <= 0.02 and ty in (0, 1))
    if not box:
        return None, None, False
    dt = 1.0 / st
    n = int(st * Tb)
    e = int(L / dt)
    T, S = e * dt, n * dt
    PT = math.exp(-_yield_at(T, y.rates, y.days) * T)
    PS = math.exp(-_yield_at(S, y.rates, y.days) * S)
    Bc = (1 - math.exp(-a * (S - T))) / a
    V = sg * sg * (1 - math.exp(-2 * a * dt)) / (2 * a)
    dr = math.sqrt(3 * V)
    q = math.exp(-a * dt)
    Bd = dt * (1 - q ** (n - e)) / (1 - q)
    e1 = 1 - math.exp(-a * T)
    c0 = (math.log(100 * PS / PT) - Bc * sg * sg * e1 * e1 / (2 * a * a)
          - Bc * Bc * sg * sg * (1 - math.exp(-2 * a * T)) / (4 * a))
    u = (c0 - math.log(X)) / (Bd * dr)
    ph = 2 * (u - math.floor(u))
    x = np.array([(X - 80) / 20, (T - 15.5) / 14.5, (S - T - 5.5) / 4.5, (st - 24) / 12,
                  (a - 0.105) / 0.095, (sg - 0.0125) / 0.0075, math.sin(math.pi * ph), math.cos(math.pi * ph)])
    return x, closed_form(X, T, S, a, sg, ty, y), True
