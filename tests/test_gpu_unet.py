"""GPU tests of the learned warm start: the BF16 MFMA forward pass and
gradients against the oracle (oracle/unet_oracle.py, pinned by finite
differences), training on the simulation's own solves, and what the trained
network gains over the previous-step start."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import jd_lepic_amd as L  # noqa: E402
import unet_oracle as U  # noqa: E402

pytestmark = pytest.mark.gpu


def box_ctx(n, precision=L.FP64, ng=None):
    h = 5e-5
    ng = ng or int(np.log2(n))
    ctx = L.Context(n, n, h, h, 5e-12, [L.Species(-L.QE, L.ME, 1.0, 16)], ng=ng, precision=precision)
    bc = -np.ones((n + 3, n + 3), np.int32)
    bc[:2, :] = bc[n + 1:, :] = bc[:, :2] = bc[:, n + 1:] = 1
    ctx.set_boundary(bc, np.zeros((n + 3, n + 3)))
    return ctx, bc


def rand_batch(B, H, W, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, H, W, 2), np.float32)
    x[..., 0] = rng.standard_normal((B, H, W))
    x[..., 1] = (rng.uniform(size=(B, H, W)) > 0.1).astype(np.float32)
    t = rng.standard_normal((B, H, W)).astype(np.float32)
    return x, t, x[..., 1].copy()


@pytest.mark.parametrize("B,H,W", [(1, 8, 8), (2, 16, 24), (1, 40, 72)])
def test_forward_matches_oracle(B, H, W):
    ctx, _ = box_ctx(16)
    with ctx:
        ctx.net_create(capacity=0, seed=1)
        flat = U.init_params(7)
        ctx.net_set_params(flat)
        assert np.array_equal(ctx.net_get_params(U.n_params()), flat)
        x, _, _ = rand_batch(B, H, W, 11)
        y = ctx.net_forward(x)
    ref, _ = U.forward(x, flat)
    # BF16 activations: sums in another order can round a stored activation
    # to the neighbouring BF16 value; the tolerance allows a few such steps
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 2e-2 * scale
    assert np.abs(y - ref).mean() <= 2e-3 * scale


@pytest.mark.parametrize("B,H,W,lam", [(1, 8, 8, 0.0), (2, 16, 24, 0.0), (2, 16, 24, 0.7)])
def test_gradient_matches_oracle(B, H, W, lam):
    ctx, _ = box_ctx(16)
    n = U.n_params()
    kw = dict(lam=lam, cw=0.3, cn=0.5)
    with ctx:
        ctx.net_create(capacity=0, seed=1)
        flat = U.init_params(3)
        ctx.net_set_params(flat)
        x, t, m = rand_batch(B, H, W, 5)
        loss, g = ctx.net_loss_grad(x, t, m, n, **kw)
    ref_loss, ref_g, _ = U.loss_and_grad(x, t, m, flat, **kw)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    cos = float(g @ ref_g / (np.linalg.norm(g) * np.linalg.norm(ref_g)))
    assert rel <= 5e-2 and cos >= 0.998, (rel, cos)
    # per layer, so a wrong layer is not hidden by the others
    o = 0
    for name, ci, co in U.CONV_LAYERS:
        k = 9 * ci * co
        gl, rl = g[o:o + k], ref_g[o:o + k]
        assert np.linalg.norm(gl - rl) <= 0.1 * np.linalg.norm(rl) + 1e-12, name
        o += k


def moving_rhs(n, step, rng_phase):
    """Charge of two Gaussian blobs drifting across the box (a smooth,
    time-dependent source like a plasma's), on the rhs grid (n+1, n+1)."""
    x = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(x, x)
    b = np.zeros((n + 1, n + 1))
    for k, (cx, cy, vx, vy, s) in enumerate([(0.3, 0.4, 0.004, 0.002, 1.0), (0.7, 0.6, -0.003, 0.003, -0.8)]):
        px, py = cx + vx * step + 0.02 * np.sin(0.3 * step + rng_phase[k]), cy + vy * step
        b += s * np.exp(-((X - px) ** 2 + (Y - py) ** 2) / (2 * 0.08 ** 2))
    b *= 1.6e-19 * 1e16 * 1e-3
    b[0, :] = b[-1, :] = b[:, 0] = b[:, -1] = 0.0
    return b


def test_training_and_warm_start_gain():
    n = 64
    ctx, _ = box_ctx(n)
    phase = [0.1, 1.3]
    with ctx:
        ctx.net_create(capacity=64, seed=2)
        ctx.net_recording(True)
        ctx.set_grid(L.GRID_RHS, moving_rhs(n, 0, phase))
        ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        for it in range(1, 41):
            ctx.set_grid(L.GRID_RHS, moving_rhs(n, it, phase))
            ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        assert ctx.net_samples() == 40
        ctx.net_recording(False)
        ctx.net_policy(min_gain=0.0, max_backoff=0)    # evaluate the network on every solve
        l0, l1 = ctx.net_train(300, batch=4, lr=2e-3)
        print(f"training loss {l0:.4g} -> {l1:.4g}")
        assert l1 < 0.3 * l0, (l0, l1)
        gains, used = [], 0
        for it in range(41, 51):
            ctx.set_grid(L.GRID_RHS, moving_rhs(n, it, phase))
            info = ctx.iterative_solver(eps=1e-10, start=L.START_NET)
            print(f"step {it}: start {info['start_used']} res0 {info['res0']:.3g} other {info['res0_alt']:.3g} "
                  f"vcycles {info['vcycles']}")
            assert info["res"] <= 1e-10
            if info["start_used"] == L.START_NET:
                used += 1
                gains.append(info["res0_alt"] / info["res0"])
            ref = ctx.get_grid(L.GRID_PHI)
        # the network's start is taken and lowers the initial residual
        assert used >= 8, used
        assert np.median(gains) > 2.0, gains
        assert np.isfinite(ref).all()


def test_probe_backoff():
    """A guess that misses min_gain makes the next 8, 16, ... max_backoff
    START_NET solves start from the previous phi without the network."""
    n = 32
    ctx, _ = box_ctx(n)
    phase = [0.4, 2.0]
    with ctx:
        ctx.net_create(capacity=8, seed=3)
        ctx.net_recording(True)
        ctx.set_grid(L.GRID_RHS, moving_rhs(n, 0, phase))
        ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        for it in range(1, 5):
            ctx.set_grid(L.GRID_RHS, moving_rhs(n, it, phase))
            ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        ctx.net_recording(False)
        ctx.net_train(1, batch=2, lr=1e-4)
        ctx.net_policy(min_gain=1e30, max_backoff=16)   # no guess can pass
        probed = []
        for it in range(5, 35):
            ctx.set_grid(L.GRID_RHS, moving_rhs(n, it, phase))
            info = ctx.iterative_solver(eps=1e-10, start=L.START_NET)
            assert info["res"] <= 1e-10
            if info["res0_alt"] >= 0.0:
                probed.append(it - 5)
            else:
                assert info["start_used"] == L.START_PREV
        print("probed", probed)
        assert probed == [0, 9, 26], probed
        ctx.net_policy(min_gain=0.0, max_backoff=256)   # every guess passes
        for it in range(35, 38):
            ctx.set_grid(L.GRID_RHS, moving_rhs(n, it, phase))
            info = ctx.iterative_solver(eps=1e-10, start=L.START_NET)
            assert info["res0_alt"] >= 0.0
