"""Offline warm-start training (jd_lepic_amd.warmstart): samples of several
runs gathered in one store, the network trained there and copied into a
context on another grid, where the solve started from its guess reaches the
same tolerance as the previous-phi start."""
import numpy as np
import pytest

L = pytest.importorskip("jd_lepic_amd")
from jd_lepic_amd import warmstart as WS, workloads as WL  # noqa: E402

pytestmark = pytest.mark.gpu


def test_samples_gathered_across_runs_and_network_copied():
    store, meta = WS.build_dataset(samples=12, runs=2, n=64, ppc=20, skip=3, seed=5)
    try:
        assert [m["samples"] for m in meta] == [6, 6]
        assert store.net_samples() == 12
        first, last, _ = WS.train(store, 40, batch=2, lr=2e-3)
        assert np.isfinite(first) and np.isfinite(last)
        W = WL.argon_box(128, 10)
        with WL.build_context(W, L.MIXED) as ctx:
            ctx.net_create(capacity=0)
            ctx.net_copy(store)
            n = 486448
            assert np.array_equal(store.net_get_params(n), ctx.net_get_params(n))
            ctx.net_policy(min_gain=1.0, max_backoff=0)
            WS.run_steps(ctx, W, 3)
            ctx.calc_rhs()
            rhs = ctx.get_grid(L.GRID_RHS)
            phi0 = ctx.get_grid(L.GRID_PHI)
            info = ctx.iterative_solver(eps=1e-10, start=L.START_NET)
            assert info["res0_alt"] >= 0.0          # the network was evaluated
            assert info["res"] <= 1e-10
            phi_net = ctx.get_grid(L.GRID_PHI)
            # the same right-hand side from the previous phi, without the network
            ctx.set_grid(L.GRID_PHI, phi0)
            ctx.set_grid(L.GRID_RHS, rhs)
            ref = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
            assert ref["res"] <= 1e-10
            phi_ref = ctx.get_grid(L.GRID_PHI)
            # both iterate to the same residual tolerance from different starts:
            # equal within the mixed form's phi tolerance
            assert np.max(np.abs(phi_net - phi_ref)) <= 1e-6 * max(np.max(np.abs(phi_ref)), 1e-30)
    finally:
        store.close()


def test_copy_samples_rejects_other_grid():
    store, _ = WS.build_dataset(samples=2, runs=1, n=64, ppc=10, skip=2, seed=6)
    try:
        W = WL.argon_box(128, 10)
        with WL.build_context(W, L.MIXED) as ctx:
            ctx.net_create(capacity=2)
            with pytest.raises(L.LepicError):
                store.net_copy_samples(ctx)
    finally:
        store.close()
