"""Pins of the oracle's part_moments / calc_avg (utils.f90:1-265) by closed
forms: the reference file needs mpif.h and cannot be compiled here."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

N, H, DT = 16, 5e-5, 5e-12
QE = O.QE


def dom():
    return O.Domain(n1=N, n2=N, h1=H, h2=H, dt=DT, bcnd=O.box_bcnd(N, N))


def test_single_particle_weights():
    """One particle at the half-step-back position (x0, y0): each of the four
    nodes gets k1 * kq * w * v with the bilinear weights of the left node."""
    d = dom()
    vx, vy, vz = 2e5, -1e5, 3e4
    x0, y0 = 5.3 * H, 7.6 * H
    v = np.array([[x0 + vx * DT / 2, y0 + vy * DT / 2, 0.0, vx, vy, vz]])
    m = O.part_moments(d, 2.0, v)
    k1 = 2.0 / (H * H)
    ix, iy = 6, 8                      # INT(x/h)+1
    px, py = ix - x0 / H, iy - y0 / H
    w = {(ix, iy): px * py, (ix + 1, iy): (1 - px) * py, (ix + 1, iy + 1): (1 - px) * (1 - py),
         (ix, iy + 1): px * (1 - py)}
    for (i, j), wt in w.items():
        np.testing.assert_allclose(m[j - 1, i - 1, 1:], k1 * wt * np.array([vx, vy, vz]), rtol=1e-12)
        np.testing.assert_allclose(m[j - 1, i - 1, 0], k1 * wt * (vx * vx + vy * vy + vz * vz), rtol=1e-12)
    assert np.count_nonzero(m[..., 0]) == 4


def stats_case(vt, drift, n_per_cell=400, seed=1, mass=O.AMU * 40):
    d = dom()
    rng = np.random.default_rng(seed)
    n = n_per_cell * N * N
    x = rng.uniform(2 * H, (N - 2) * H, n)
    y = rng.uniform(2 * H, (N - 2) * H, n)
    v = rng.normal(0, vt, (n, 3)) + np.asarray(drift)[None, :]
    vxp = np.column_stack([x, y, np.zeros(n), v])
    Nm = 1.0
    dens = O.charge_deposition(d, Nm, vxp)
    m = O.part_moments(d, Nm, vxp)
    avg = O.calc_avg(d, [mass], [Nm], [dens], [m], [np.zeros_like(dens)])
    return d, vxp, dens, m, avg[0]


def test_maxwellian_temperature_and_drift():
    """Tp = m/qe (<v^2> - |<v>|^2)/3 per node: a Maxwellian of thermal speed
    vt (per component) gives m vt^2/qe; a drift adds nothing to Tp and
    shows up as j/n."""
    mass = O.AMU * 40
    vt = 3e3
    T = mass * vt * vt / QE
    for drift in ([0.0, 0.0, 0.0], [4e3, -2e3, 1e3]):
        d, vxp, dens, m, avg = stats_case(vt, drift, mass=mass)
        inner = avg[5:12, 5:12]
        n = inner[..., 0]
        # the moments use the half-step-back position, the density the
        # current one; over a drift of v dt << h the two agree closely
        Tp = inner[..., 1]
        assert abs(np.median(Tp) / T - 1) < 0.05, np.median(Tp) / T
        for c in range(3):
            u = np.median(inner[..., 2 + c] / n)
            assert abs(u - drift[c]) < 0.1 * vt, (c, u, drift[c])


def test_cold_beam_has_no_temperature():
    """A cold beam: Tp is only what the density (current positions) and the
    moments (half a step back) disagree by, far below the beam energy."""
    mass = O.AMU * 40
    u = np.array([1e4, 2e4, -5e3])
    d, vxp, dens, m, avg = stats_case(0.0, u, mass=mass)
    inner = avg[5:12, 5:12]
    e_beam = mass * (u @ u) / QE
    assert np.all(inner[..., 1] >= 0.0) and inner[..., 1].max() < 1e-4 * e_beam
    np.testing.assert_allclose(inner[..., 2] / inner[..., 0], 1e4, rtol=1e-3)


def test_particles_outside_are_skipped_and_sink_scaled():
    d = dom()
    v = np.array([[-1e-9, 3 * H, 0, 0, 0, 0], [3 * H, 3 * H, 0, 1e5, 0, 0]])
    m = O.part_moments(d, 1.0, v)
    assert np.count_nonzero(m[..., 1]) == 4      # only the second particle
    sink = np.zeros((N + 3, N + 3))
    sink[4, 4] = 3.0
    src = np.zeros((N + 3, N + 3))
    src[6, 2] = 5.0
    avg = O.calc_avg(d, [O.ME], [2.0], [np.zeros_like(sink)], [m], [sink], srcs=[src])
    assert avg[0, 3, 3, 6] == 3.0 * 2.0 / (H * H)   # ss2D * Nm/Vc at node (4,4)
    assert np.count_nonzero(avg[0, ..., 6]) == 1
    assert avg[0, 5, 1, 5] == 5.0 * 2.0 / (H * H)   # the source column, node (2,6)
    assert np.count_nonzero(avg[0, ..., 5]) == 1


def test_coarse_grid_sums_fine_moments():
    """size_na = 2: moments on the coarse grid; every particle's weight is
    counted once, so the totals over nodes agree with size_na = 1 up to the
    k1 scaling 1/size_na^2 and the kq factors (interior here)."""
    d = dom()
    rng = np.random.default_rng(3)
    n = 2000
    vxp = np.column_stack([rng.uniform(3 * H, 12 * H, n), rng.uniform(3 * H, 12 * H, n), np.zeros(n),
                           rng.normal(0, 1e5, (n, 3))])
    m1 = O.part_moments(d, 1.0, vxp, size_na=1)
    m2 = O.part_moments(d, 1.0, vxp, size_na=2)
    np.testing.assert_allclose(m2.sum(axis=(0, 1)) * 4, m1.sum(axis=(0, 1)), rtol=1e-10)
