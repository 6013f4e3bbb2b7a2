"""CPU oracle of the learned Poisson warm start (test infrastructure only).

The U-Net has no counterpart in the reference: it is the learned part the
build adds on top of the Poisson solve (iterative_solver.f90:1-69 stays the
solver of record; the net only proposes the starting guess).  This module
restates the network the GPU kernels implement, in numpy, with the same
rounding points, so the GPU forward pass and gradients can be checked
against it; `finite_difference_check` pins the oracle's own backward pass
against central differences in float64.

Architecture (4 levels, 16-32-64-128 channels, 3x3 convolutions, no bias):
  enc0: 2->16, 16->16          (skip0)  | avg-pool 2x2
  enc1: 16->32, 32->32         (skip1)  | avg-pool
  enc2: 32->64, 64->64         (skip2)  | avg-pool
  mid : 64->128, 128->128
  dec2: up(128) ++ skip2 -> 192->64, 64->64
  dec1: up(64)  ++ skip1 ->  96->32, 32->32
  dec0: up(32)  ++ skip0 ->  48->16, 16->16
  out : 1x1 16->1 (FP32 weights, no activation)
Every 3x3 convolution is followed by ReLU.  Inputs, 3x3 weights, stored
activations and stored activation gradients are BF16 (round to nearest
even); products accumulate in FP32.  Tensors are NHWC; the spatial size must
be a multiple of 8.  Weight layout of a 3x3 layer: [cout][tap][cin] with
tap = 3*(dy+1) + (dx+1).
"""
from __future__ import annotations

import numpy as np

CONV_LAYERS = [  # (name, cin, cout)
    ("enc0a", 2, 16), ("enc0b", 16, 16),
    ("enc1a", 16, 32), ("enc1b", 32, 32),
    ("enc2a", 32, 64), ("enc2b", 64, 64),
    ("mida", 64, 128), ("midb", 128, 128),
    ("dec2a", 192, 64), ("dec2b", 64, 64),
    ("dec1a", 96, 32), ("dec1b", 32, 32),
    ("dec0a", 48, 16), ("dec0b", 16, 16),
]
OUT_CIN = 16
_DT = np.float32     # arithmetic type; the finite-difference check runs in float64


def n_params():
    return sum(9 * ci * co for _, ci, co in CONV_LAYERS) + OUT_CIN


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a).astype(np.float32)


def init_params(seed=0):
    """He-normal 3x3 weights and a small output layer (float32, flat)."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, ci, co in CONV_LAYERS:
        parts.append((rng.standard_normal((co, 9, ci)) * np.sqrt(2.0 / (9 * ci))).astype(np.float32).ravel())
    parts.append((rng.standard_normal(OUT_CIN) * np.sqrt(1.0 / OUT_CIN)).astype(np.float32))
    return np.concatenate(parts)


def split_params(flat):
    out, o = {}, 0
    for name, ci, co in CONV_LAYERS:
        n = 9 * ci * co
        out[name] = flat[o:o + n].reshape(co, 9, ci)
        o += n
    out["out"] = flat[o:o + OUT_CIN]
    return out


def _shifted(x, dy, dx):
    """x[b, y+dy, x+dx, c] with zero padding (NHWC)."""
    B, H, W, C = x.shape
    p = np.zeros((B, H + 2, W + 2, C), _DT)
    p[:, 1:H + 1, 1:W + 1] = x
    return p[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]


def conv3x3(x, w):
    """y[b,y,x,co] = sum_{tap,ci} w[co,tap,ci] x[b,y+dy,x+dx,ci] (FP32 sums)."""
    B, H, W, _ = x.shape
    co = w.shape[0]
    y = np.zeros((B, H, W, co), _DT)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += np.einsum("bhwc,oc->bhwo", _shifted(x, dy, dx), w[:, t, :], dtype=_DT, optimize=True)
    return y


def conv3x3_dgrad(dy_, w):
    """Gradient w.r.t. the input of conv3x3."""
    B, H, W, _ = dy_.shape
    ci = w.shape[2]
    dx_ = np.zeros((B, H, W, ci), _DT)
    for t in range(9):
        ddy, ddx = t // 3 - 1, t % 3 - 1
        dx_ += np.einsum("bhwo,oc->bhwc", _shifted(dy_, -ddy, -ddx), w[:, t, :], dtype=_DT, optimize=True)
    return dx_


def conv3x3_wgrad(x, dy_):
    co, ci = dy_.shape[3], x.shape[3]
    dw = np.zeros((co, 9, ci), _DT)
    for t in range(9):
        ddy, ddx = t // 3 - 1, t % 3 - 1
        dw[:, t, :] = np.einsum("bhwo,bhwc->oc", dy_, _shifted(x, ddy, ddx), dtype=_DT, optimize=True)
    return dw


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4), dtype=_DT)


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) * _DT(0.25)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4), dtype=_DT)


def forward(x, flat, rounding=True):
    """x: (B,H,W,2) float32.  Returns (y (B,H,W) float32, cache)."""
    R = bf16 if rounding else (lambda a: np.asarray(a, _DT))
    P = split_params(flat)
    Wt = {k: R(v) for k, v in P.items() if k != "out"}
    cache = {"x": R(x)}

    def conv(name, a):
        z = conv3x3(a, Wt[name])
        o = R(np.maximum(z, 0.0))
        cache[name + ":in"] = a
        cache[name + ":out"] = o
        return o

    a = conv("enc0a", cache["x"]); s0 = conv("enc0b", a)
    a = R(pool(s0)); a = conv("enc1a", a); s1 = conv("enc1b", a)
    a = R(pool(s1)); a = conv("enc2a", a); s2 = conv("enc2b", a)
    a = R(pool(s2)); a = conv("mida", a); m = conv("midb", a)
    a = np.concatenate([up(m), s2], axis=3); a = conv("dec2a", a); d2 = conv("dec2b", a)
    a = np.concatenate([up(d2), s1], axis=3); a = conv("dec1a", a); d1 = conv("dec1b", a)
    a = np.concatenate([up(d1), s0], axis=3); a = conv("dec0a", a); d0 = conv("dec0b", a)
    cache["d0"] = d0
    y = np.einsum("bhwc,c->bhw", d0, P["out"].astype(_DT), dtype=_DT)
    return y, cache


def stencil(u, cw, cn):
    """cw*(u[x-1]+u[x+1]) + cn*(u[y-1]+u[y+1]) - 2*(cw+cn)*u, zero outside (B,H,W)."""
    p = np.zeros((u.shape[0], u.shape[1] + 2, u.shape[2] + 2), _DT)
    p[:, 1:-1, 1:-1] = u
    return (cw * (p[:, 1:-1, :-2] + p[:, 1:-1, 2:]) + cn * (p[:, :-2, 1:-1] + p[:, 2:, 1:-1])
            - 2.0 * (cw + cn) * u).astype(_DT)


def loss_and_grad(x, target, mask, flat, rounding=True, lam=0.0, cw=0.0, cn=0.0):
    """Loss over mask nodes: L = sum(mask*(y-t)^2)/sum(mask) plus, when lam > 0,
    lam * sum(r^2)/sum(mask) with the Poisson residual r = mask*(A(mask*y) -
    x[...,0]) of the prediction (A: the 5-point stencil with coefficients cw,
    cn in the network's units).  Returns (L, dflat, y)."""
    R = bf16 if rounding else (lambda a: np.asarray(a, _DT))
    y, c = forward(x, flat, rounding)
    P = split_params(flat)
    Wt = {k: R(v) for k, v in P.items() if k != "out"}
    nm = _DT(mask.sum())
    diff = (y - target) * mask
    L = float((diff.astype(np.float64) ** 2).sum() / nm)
    dy = (2.0 * diff / nm).astype(_DT)
    if lam > 0.0:
        r = (stencil(y * mask, cw, cn) - x[..., 0]) * mask
        L += lam * float((r.astype(np.float64) ** 2).sum() / nm)
        dy = (dy + mask * stencil(r, cw, cn) * _DT(2.0 * lam / nm)).astype(_DT)
    g = {"out": np.einsum("bhwc,bhw->c", c["d0"], dy, dtype=_DT)}
    dd0 = R(dy[..., None] * P["out"].astype(_DT)[None, None, None, :]) * (c["d0"] > 0)

    def back(name, dout):
        """dout: gradient w.r.t. the layer's post-ReLU output, already masked by
        its ReLU.  Returns the gradient w.r.t. the layer input (unmasked)."""
        g[name] = conv3x3_wgrad(c[name + ":in"], dout)
        return conv3x3_dgrad(dout, Wt[name])

    def relu_mask(name, d):
        return R(d) * (c[name + ":out"] > 0)

    da = back("dec0b", dd0)
    da = back("dec0a", relu_mask("dec0a", da))          # gradient w.r.t. concat(up(d1), s0)
    dup, ds0 = da[..., :32], da[..., 32:]
    dd1 = up_bwd(R(dup))
    da = back("dec1b", relu_mask("dec1b", dd1))
    da = back("dec1a", relu_mask("dec1a", da))
    dup, ds1 = da[..., :64], da[..., 64:]
    dd2 = up_bwd(R(dup))
    da = back("dec2b", relu_mask("dec2b", dd2))
    da = back("dec2a", relu_mask("dec2a", da))
    dup, ds2 = da[..., :128], da[..., 128:]
    dm = up_bwd(R(dup))
    da = back("midb", relu_mask("midb", dm))
    da = back("mida", relu_mask("mida", da))
    ds2 = R(ds2) + pool_bwd(R(da))
    da = back("enc2b", relu_mask("enc2b", ds2))
    da = back("enc2a", relu_mask("enc2a", da))
    ds1 = R(ds1) + pool_bwd(R(da))
    da = back("enc1b", relu_mask("enc1b", ds1))
    da = back("enc1a", relu_mask("enc1a", da))
    ds0 = R(ds0) + pool_bwd(R(da))
    da = back("enc0b", relu_mask("enc0b", ds0))
    back("enc0a", relu_mask("enc0a", da))
    dflat = np.concatenate([g[name].ravel() for name, _, _ in CONV_LAYERS] + [g["out"]])
    return L, dflat.astype(_DT), y


def finite_difference_check(H=8, W=8, B=1, n_probe=12, seed=3):
    """Largest error between the analytic gradient (no rounding, the same code
    path, float64 arithmetic) and central differences of the loss, relative to
    the largest probed gradient entry, on a small random problem; pins the
    conv/pool/up/concat/ReLU backward pass."""
    global _DT
    saved, _DT = _DT, np.float64
    try:
        return _fd_check(H, W, B, n_probe, seed)
    finally:
        _DT = saved


def _fd_check(H, W, B, n_probe, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, H, W, 2))
    t = rng.standard_normal((B, H, W))
    m = (rng.uniform(size=(B, H, W)) > 0.2).astype(np.float64)
    flat = init_params(seed).astype(np.float64)
    kw = dict(rounding=False, lam=0.7, cw=0.3, cn=0.5)   # both loss terms
    _, g, _ = loss_and_grad(x, t, m, flat, **kw)
    idx = rng.choice(flat.size, n_probe, replace=False)
    idx = np.concatenate([idx, [flat.size - 1, 0]])
    num, ana = [], []
    for i in idx:
        h = 1e-6 * max(abs(float(flat[i])), 1e-2)
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        Lp, _, _ = loss_and_grad(x, t, m, fp, **kw)
        Lm, _, _ = loss_and_grad(x, t, m, fm, **kw)
        num.append((Lp - Lm) / (fp[i] - fm[i]))
        ana.append(g[i])
    num, ana = np.array(num), np.array(ana)
    # error relative to the scale of the probed gradient entries
    return float(np.abs(num - ana).max() / max(np.abs(ana).max(), 1e-30))
