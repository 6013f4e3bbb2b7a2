"""Known-answer pins of the oracle's calc_rhs and dens_red (calc_rho.f90:37-77,
134-199).  Each case is built so that a plausible restatement mistake fails
it: a flipped sign, a species loop in another order, ghost nodes included,
the maximum taken over wall nodes, or a maximum that starts below zero."""
import numpy as np

import oracle as O


def test_calc_rhs_sign_and_range():
    # one species of charge +2 with density 3 at node (ix,iy) = (2,3):
    # rhs(2,3) = 0 - 2*3 = -6; every other rhs entry 0.  Ghost nodes
    # (index 0 and n+2) hold large values the sum must not read.
    n1, n2 = 5, 4
    nps = np.zeros((n2 + 3, n1 + 3))
    nps[0, :] = nps[-1, :] = nps[:, 0] = nps[:, -1] = 1e30
    nps[3, 2] = 3.0
    rhs = O.calc_rhs(n1, n2, [2.0], [nps])
    assert rhs.shape == (n2 + 1, n1 + 1)
    want = np.zeros((n2 + 1, n1 + 1))
    want[3 - 1, 2 - 1] = -6.0     # rhs(ix,iy) is stored at [iy-1, ix-1]
    assert np.array_equal(rhs, want)


def test_calc_rhs_species_order():
    # r = ((0 - c1 n1) - c2 n2) - c3 n3 in species order: with c*n = 1, 1e-16,
    # 1e-16 each small term is below half an ulp of 1 and rounds away, so
    # rhs = -1 exactly; summing the small terms first gives -1 - 2^-52.
    n1 = n2 = 2
    nps = [np.full((n2 + 3, n1 + 3), v) for v in (1.0, 1e-16, 1e-16)]
    rhs = O.calc_rhs(n1, n2, [1.0, 1.0, 1.0], nps)
    assert np.all(rhs == -1.0)
    rev = O.calc_rhs(n1, n2, [1.0, 1.0, 1.0], nps[::-1])
    assert np.all(rev == -1.0 - 2.0 ** -52)


def test_calc_rhs_neutral_plasma_and_linearity():
    # equal electron and ion densities cancel exactly; rhs is linear in np
    rng = np.random.default_rng(5)
    n1, n2 = 9, 7
    a = rng.uniform(0, 1e16, (n2 + 3, n1 + 3))
    b = rng.uniform(0, 1e16, (n2 + 3, n1 + 3))
    qe = O.QE
    assert np.all(O.calc_rhs(n1, n2, [-qe, qe], [a, a]) == 0.0)
    r1 = O.calc_rhs(n1, n2, [qe], [a])
    np.testing.assert_array_equal(r1, -qe * a[1:n2 + 2, 1:n1 + 2])
    r2 = O.calc_rhs(n1, n2, [-qe, qe], [a, b])
    np.testing.assert_allclose(r2, qe * (a - b)[1:n2 + 2, 1:n1 + 2], rtol=0, atol=4e-16 * qe * 1e16)


def test_dens_red_copy_ghosts_and_interior_max():
    n1, n2 = 6, 5
    bcnd = O.box_bcnd(n1, n2, 1)
    rng = np.random.default_rng(6)
    a = rng.uniform(0, 1, (n2 + 3, n1 + 3))
    b = rng.uniform(0, 1, (n2 + 3, n1 + 3))
    a[0, :] = a[-1, :] = 50.0            # ghosts: never copied
    a[1, 3] = 1e6                        # wall node (bcnd = 1): not in the maximum
    assert bcnd[1, 3] == 1
    b[:] = -1.0                          # no positive density: the maximum stays 0
    b[3, 3] = -0.5
    (ra, rb), mx = O.dens_red(n1, n2, bcnd, [a, b])
    for src, red in ((a, ra), (b, rb)):
        inner = np.zeros_like(src, dtype=bool)
        inner[1:n2 + 2, 1:n1 + 2] = True
        assert np.array_equal(red[inner], src[inner])
        assert np.all(red[~inner] == 0.0)
    interior = bcnd == -1
    assert mx[0] == a[interior].max() and mx[0] < 1e6
    assert mx[1] == 0.0
    # the maximum is attained at an interior node holding the largest value
    a2 = a.copy()
    iy, ix = np.argwhere(interior)[7]
    a2[iy, ix] = 7.5
    _, mx2 = O.dens_red(n1, n2, bcnd, [a2])
    assert mx2[0] == 7.5
