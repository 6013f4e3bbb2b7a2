"""Context-level backup files: particles saved from the GPU state and loaded
into another context come back unchanged (FP64 form bit-exact; the mixed
form stores cell + FP32 offset, whose round trip through x = (ix-1+fx) h is
exact to the offset's rounding), and the reloaded state redeposits the same
density, as restart_OMP does (restart.f90:157-226)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import jd_lepic_amd as L  # noqa: E402

pytestmark = pytest.mark.gpu


def make(prec, n=32):
    h = 5e-5
    sp = [L.Species(-L.QE, L.ME, 1e3, 200000), L.Species(L.QE, 40 * L.AMU, 1e3, 200000)]
    ctx = L.Context(n, n, h, h, 5e-12, sp, ng=5, precision=prec, track_z=0)
    bc = -np.ones((n + 3, n + 3), np.int32)
    bc[:2, :] = bc[n + 1:, :] = bc[:, :2] = bc[:, n + 1:] = 1
    ctx.set_boundary(bc, np.zeros((n + 3, n + 3)))
    return ctx


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_save_load_round_trip(tmp_path, prec):
    f = tmp_path / "particles0.bak"
    a = make(prec)
    with a:
        a.load_uniform(1, 20, 1e6, 0.0, 32 * 5e-5, 11)
        a.load_uniform(2, 20, 1e3, 0.0, 32 * 5e-5, 11)
        for p in (1, 2):
            a.part_mover(p, deposit=False)   # leaves unpacked segments behind
        before = [a.download(p) for p in (1, 2)]
        a.save_particles(f, time=2.5e-9, gtype=3, pabs_cor=0.5)
        for p in (1, 2):
            a.charge_deposition(p)
        rho_a = [a.get_grid(L.GRID_NP, p).copy() for p in (1, 2)]
    # the file is the reference's: one partition, vxp per species
    t, g, pc, counts = L.backup_header(f, 2)
    assert (t, g, pc) == (2.5e-9, 3, 0.5)
    assert counts.tolist() == [[len(before[0]), len(before[1])]]
    for p in (1, 2):
        assert np.array_equal(L.backup_read_species(f, 2, p), before[p - 1])
    b = make(prec)
    with b:
        assert b.load_particles(f) == (2.5e-9, 3, 0.5)
        after = [b.download(p) for p in (1, 2)]
        for p in (1, 2):
            b.charge_deposition(p)
        rho_b = [b.get_grid(L.GRID_NP, p).copy() for p in (1, 2)]
    for p in range(2):
        assert after[p].shape == before[p].shape
        if prec == L.FP64:
            assert np.array_equal(after[p], before[p])
            np.testing.assert_allclose(rho_b[p], rho_a[p], rtol=1e-12, atol=0)
        else:
            h = 5e-5
            assert np.abs(after[p][:, :2] - before[p][:, :2]).max() <= 1e-6 * h
            assert np.array_equal(after[p][:, 3:], before[p][:, 3:])
            scale = np.abs(rho_a[p]).max()
            assert np.abs(rho_b[p] - rho_a[p]).max() <= 1e-5 * scale
