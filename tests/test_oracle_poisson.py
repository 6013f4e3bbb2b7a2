"""Pins of the oracle's Poisson solver (mg.f90 / sors.f90 cannot be compiled
here: they need MPI).  The converged potential is pinned by a direct sparse
solve of the same 5-point system; a sweep by the Gauss-Seidel property of the
red/black ordering; restriction and prolongation by their linear maps.  The
per-level sweep counts remain parity unpinned."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def system(n1, n2, h, bcnd, eps0=O.EPS0):
    """A (on all nodes 1..n+1) and the interior mask, from mg.h coefficients."""
    aw = eps0 / (h[0] * h[0])
    an = eps0 / (h[1] * h[1])
    ac = -(2 * an + 2 * aw)
    W = n1 + 3
    idx = lambda i, j: j * W + i
    rows, cols, vals = [], [], []
    inner = bcnd == -1
    for j in range(n2 + 3):
        for i in range(n1 + 3):
            if inner[j, i]:
                r = idx(i, j)
                for (di, dj, a) in ((0, 0, ac), (-1, 0, aw), (1, 0, aw), (0, -1, an), (0, 1, an)):
                    rows.append(r); cols.append(idx(i + di, j + dj)); vals.append(a)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(W * (n2 + 3), W * (n2 + 3)))
    return A, inner


def full_rhs(b, n1, n2):
    B = np.zeros((n2 + 3, n1 + 3))
    B[1:n2 + 2, 1:n1 + 2] = b
    return B


def direct_solve(phi_walls, b, bcnd, n1, n2, h):
    A, inner = system(n1, n2, h, bcnd)
    flat_in = inner.ravel()
    rhs = full_rhs(b, n1, n2).ravel() - A @ np.where(flat_in, 0.0, phi_walls.ravel())
    Aii = A[flat_in][:, flat_in]
    u = phi_walls.ravel().copy()
    u[flat_in] = spla.spsolve(Aii.tocsc(), rhs[flat_in])
    return u.reshape(phi_walls.shape)


def case(n1=32, n2=32, seed=0, electrode=True):
    h = (5e-5, 5e-5)
    bc = O.box_bcnd(n1, n2)
    phi = np.zeros((n2 + 3, n1 + 3))
    if electrode:
        bc[n2 // 2 - 2:n2 // 2 + 2, n1 // 2] = 2
        phi[bc == 2] = 25.0
    rng = np.random.default_rng(seed)
    b = rng.normal(0, 1e-3, (n2 + 1, n1 + 1))  # rho/eps0 scale ~ V/m^2 * eps0
    b[bc[1:n2 + 2, 1:n1 + 2] != -1] = 0.0
    return h, bc, phi, b


@pytest.mark.parametrize("ng", [1, 3, 5])
def test_converged_phi_matches_direct_solve(ng):
    n = 32
    h, bc, phi, b = case(n, n)
    out = O.iterative_solver(phi, b, bc, h, n, n, ng, eps=1e-10)
    ref = direct_solve(phi, b, bc, n, n, h)
    scale = np.abs(ref).max()
    assert np.abs(out["phi"] - ref).max() / scale < 1e-7
    assert out["res"] <= 1e-10
    assert np.array_equal(out["phi"][bc >= 1], phi[bc >= 1])  # walls untouched


def test_red_black_sweep_gauss_seidel_property():
    # with omega = 1 the black half-sweep solves every black equation exactly
    # given the red values, so right after a sweep the residual vanishes on
    # black nodes (i+j even) and not on red ones.
    n = 16
    h, bc, phi, b = case(n, n, seed=3, electrode=False)
    rng = np.random.default_rng(4)
    u0 = phi.copy()
    u0[bc == -1] = rng.normal(0, 1.0, (bc == -1).sum())
    u1, _, _ = O.sor_rb(u0, b, h, O.EPS0, bc, n, n, omega=1.0, eps=0.0, kmg=2, dig=-1)
    A, inner = system(n, n, h, bc)
    r = (full_rhs(b, n, n).ravel() - A @ u1.ravel()).reshape(u1.shape)
    J, I = np.mgrid[0:n + 3, 0:n + 3]
    black = inner & ((I + J) % 2 == 0)
    red = inner & ((I + J) % 2 == 1)
    scale = np.abs(A @ u1.ravel()).max()
    assert np.abs(r[black]).max() < 1e-12 * scale
    assert np.abs(r[red]).max() > 1e-6 * scale


def test_sor_fixed_point():
    n = 16
    h, bc, phi, b = case(n, n, seed=5)
    exact = direct_solve(phi, b, bc, n, n, h)
    u1, res, ks = O.sor_rb(exact, b, h, O.EPS0, bc, n, n, omega=1.4, eps=0.0, kmg=2, dig=-1)
    assert np.abs(u1 - exact).max() < 1e-12 * np.abs(exact).max()


def test_restriction_is_full_weighting_of_residual():
    n = 16
    h, bc, phi, b = case(n, n, seed=6)
    rng = np.random.default_rng(7)
    u = phi.copy()
    u[bc == -1] = rng.normal(0, 1.0, (bc == -1).sum())
    r2, e2, bc2 = O.restriction(u, b, h, O.EPS0, bc, n, n)
    A, inner = system(n, n, h, bc)
    r = (full_rhs(b, n, n).ravel() - A @ u.ravel()).reshape(u.shape)
    r[~inner] = 0.0
    w = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]]) / 16.0
    m = n // 2
    for j in range(1, m + 2):
        for i in range(1, m + 2):
            ir, jr = 2 * i - 1, 2 * j - 1
            assert bc2[j, i] == bc[jr, ir]
            if bc2[j, i] <= 0:
                expect = (w * r[jr - 1:jr + 2, ir - 1:ir + 2]).sum()
                assert np.isclose(r2[j - 1, i - 1], expect, rtol=1e-12, atol=1e-12 * np.abs(r).max())
            else:
                assert r2[j - 1, i - 1] == 0.0
    assert not e2.any()


def test_prolongation_exact_for_bilinear_corrections():
    n = 16
    m = n // 2
    bc = O.box_bcnd(n, n)
    X, Y = np.meshgrid(np.arange(m + 3) * 2.0, np.arange(m + 3) * 2.0)  # fine index units
    f = lambda x, y: 0.3 + 0.1 * x - 0.2 * y + 0.05 * x * y
    e2 = f(X - 2.0, Y - 2.0)     # coarse node i sits at fine index 2i-1 -> x = 2i - 2
    u = np.zeros((n + 3, n + 3))
    out = O.prolongation(u, e2, bc, n, n)
    Xf, Yf = np.meshgrid(np.arange(n + 3) * 1.0, np.arange(n + 3) * 1.0)
    inner = bc == -1
    np.testing.assert_allclose(out[inner], f(Xf - 1.0, Yf - 1.0)[inner], rtol=1e-13, atol=1e-14)
    assert not out[~inner].any()


def test_mg_beats_sor_in_work():
    n = 64
    h, bc, phi, b = case(n, n, seed=8)
    mg = O.iterative_solver(phi, b, bc, h, n, n, 6, eps=1e-10)
    sor = O.iterative_solver(phi, b, bc, h, n, n, 1, eps=1e-10)
    assert mg["ktot"] < sor["ktot"]
    assert np.abs(mg["phi"] - sor["phi"]).max() < 1e-6 * np.abs(sor["phi"]).max()
