"""Seeded synthetic inputs shared by the oracle and GPU parity tests."""
import numpy as np

import oracle as O


def box_domain(n1=16, n2=12, h=5e-5, dt=5e-12, Bmode="map", Bpeak=0.01, flag_B_pos=0,
               ngrid=1, electrode=False):
    """A grounded box (generate_boundary.f90 drawn on its four sides) with a
    Gaussian Bz(x) filter field (Bfield.f90:41-63) on the E grid."""
    bcnd = O.box_bcnd(n1, n2, 1)
    if electrode:  # a grid electrode (label 2) across the middle, with a slit
        ix = n1 // 2 + 1
        bcnd[:, ix:ix + 2] = 2
        bcnd[n2 // 2:n2 // 2 + 3, ix:ix + 2] = -1
        ngrid = max(ngrid, 2)
    dom = O.Domain(n1=n1, n2=n2, h1=h, h2=h, dt=dt, bcnd=bcnd, ngrid=ngrid, flag_B_pos=flag_B_pos)
    if Bmode == "map":
        L = n1 * h
        x = (np.arange(n1 + 3) - 1) * h
        Bz = Bpeak * np.exp(-(x - L / 2) ** 2 / (2 * (L / 8) ** 2))
        Bi = np.zeros((n2 + 3, n1 + 3, 4))
        Bi[..., 2] = Bz[None, :]
        Bi[..., 3] = np.sqrt(Bi[..., 0] ** 2 + Bi[..., 1] ** 2 + Bi[..., 2] ** 2)
        dom.Bi, dom.nB, dom.hB = Bi, (n1, n2), (h, h)
    elif Bmode == "const":
        Bi = np.zeros((4, 4, 4))
        Bi[1, 1, :3] = (0.002, -0.001, Bpeak)
        Bi[1, 1, 3] = np.linalg.norm(Bi[1, 1, :3])
        dom.Bi, dom.nB, dom.hB = Bi, (1, 1), (n1 * h, n2 * h)
    elif Bmode == "coarse":  # B map on a grid of half the resolution
        m1, m2 = n1 // 2, n2 // 2
        hb = 2 * h
        rng = np.random.default_rng(7)
        Bi = np.zeros((m2 + 3, m1 + 3, 4))
        Bi[..., :3] = rng.normal(0, Bpeak, (m2 + 3, m1 + 3, 3))
        Bi[..., 3] = np.sqrt((Bi[..., :3] ** 2).sum(-1))
        Bi[1:4, 1:4, :] = 0.0   # a field-free patch: leap-frog there
        dom.Bi, dom.nB, dom.hB, dom.flag_gridB = Bi, (m1, m2), (hb, hb), 1
    return dom


def random_E(dom, scale=2e4, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, (dom.n2 + 3, dom.n1 + 3, 2))


def particles(dom, N, vt, seed=2, edge_frac=0.2):
    """Positions uniform in the box, a fraction pushed near the walls so some
    leave; Maxwellian-like velocities of scale vt."""
    rng = np.random.default_rng(seed)
    L1, L2 = dom.n1 * dom.h1, dom.n2 * dom.h2
    x = rng.uniform(0, L1, N)
    y = rng.uniform(0, L2, N)
    k = rng.random(N) < edge_frac
    x[k] = np.where(rng.random(k.sum()) < 0.5, rng.uniform(0, dom.h1 * 0.3, k.sum()),
                    L1 - rng.uniform(0, dom.h1 * 0.3, k.sum()))
    v = rng.normal(0, vt, (N, 3))
    v[k, 0] *= 20.0  # fast toward/away from the walls
    z = rng.uniform(0, 1.0, N)
    return np.column_stack([x, y, z, v])


def h2_source_domain(n1=64, n2=32, h=5e-5, dt=5e-12, Bpeak=0.01):
    """Negative-ion-source-like geometry (the 1024x512 four-species workload at
    reduced size): grounded box (label 1), a grid electrode of three bars a
    few cells before the extraction wall (label 2), the extraction wall
    itself (label 3), and the Gaussian Bz filter field.  Electrode edges sit
    on odd node indices so MG injection keeps them on coarse levels."""
    bcnd = O.box_bcnd(n1, n2, 1)
    bcnd[:, n1 + 1:] = 3                   # extraction wall and its ghost column
    c0 = n1 - 11 if (n1 - 11) % 2 == 1 else n1 - 10
    for r0 in range(5, n2 - 3, 8):
        bcnd[r0:r0 + 5, c0:c0 + 3] = 2
    dom = O.Domain(n1=n1, n2=n2, h1=h, h2=h, dt=dt, bcnd=bcnd, ngrid=3, flag_B_pos=1)
    L = n1 * h
    x = (np.arange(n1 + 3) - 1) * h
    Bi = np.zeros((n2 + 3, n1 + 3, 4))
    Bi[..., 2] = Bpeak * np.exp(-(x - L / 2) ** 2 / (2 * (L / 8) ** 2))[None, :]
    Bi[..., 3] = np.abs(Bi[..., 2])
    dom.Bi, dom.nB, dom.hB = Bi, (n1, n2), (h, h)
    return dom


def exp_profile_particles(dom, N, vt, n_hi=1e17, n_lo=1e15, seed=3):
    """Particles with density falling exponentially from n_hi at x=0 to n_lo
    at the extraction wall, uniform in y, none inside electrode cells."""
    rng = np.random.default_rng(seed)
    L1, L2 = dom.n1 * dom.h1, dom.n2 * dom.h2
    lam = L1 / np.log(n_hi / n_lo)
    out = []
    while sum(len(a) for a in out) < N:
        u = rng.uniform(size=2 * N)
        x = -lam * np.log(1.0 - u * (1.0 - np.exp(-L1 / lam)))
        y = rng.uniform(0, L2, 2 * N)
        ix = (x / dom.h1).astype(int) + 1
        iy = (y / dom.h2).astype(int) + 1
        b = dom.bcnd
        walls = (b[iy, ix] >= 1) & (b[iy, ix + 1] >= 1) & (b[iy + 1, ix] >= 1) & (b[iy + 1, ix + 1] >= 1)
        keep = ~walls
        out.append(np.column_stack([x[keep], y[keep]]))
    xy = np.concatenate(out)[:N]
    v = rng.normal(0, vt, (N, 3))
    z = rng.uniform(0, 1.0, N)
    return np.column_stack([xy, z, v])
