"""Cross-section table build (read_reactions.f90:511-697, ``ordering``).

Host-side setup: the union energy grid of all reaction tables, every table
interpolated onto it, the per-incident-species reaction lists and
(sigma*v)_max.  The same rules as the reference:
  * the union grid holds every distinct energy of every table (:546-585);
  * below a table's first point the first value is used (:613-617), between
    points linear interpolation (:621-627); past a table's last point the
    value of the previous union point is kept (the loop finds no interval
    and sig_p is not reassigned);
  * (sigma*v)_max ignores energies below 0.1 eV (:669);
  * reactions are listed per incident particle, the first reactant (:653-655).
"""
from __future__ import annotations

import numpy as np

from .reactions import QE, Chemistry


def _table_on_grid(E: np.ndarray, S: np.ndarray, Er: np.ndarray) -> np.ndarray:
    """One reaction's table on the union grid (read_reactions.f90:601-631),
    vectorised: union point Ekp takes the first interval [E(j), E(j+1)] that
    holds it -- the one ending at the first table point >= Ekp -- as
    sL + (Ekp-EkL)*(sR-sL)/(EkR-EkL), the first value below the table, and
    past the table the value of the previous union point.  Tables whose
    first points repeat, or with a single point, take the point-by-point
    scan of the reference (its divisions by a zero interval included)."""
    n = len(E)
    if n < 2 or E[0] == E[1]:
        return _table_on_grid_scan(E, S, Er)
    idx = np.searchsorted(E, Er, side="left")
    out = np.empty_like(Er)
    below = idx == 0
    out[below] = S[0]
    at0 = below & (Er == E[0])
    with np.errstate(divide="ignore", invalid="ignore"):
        out[at0] = S[0] + (Er[at0] - E[0]) * (S[1] - S[0]) / (E[1] - E[0])
        mid = (idx > 0) & (idx < n)
        j = idx[mid] - 1
        EkL, EkR, sL, sR = E[j], E[j + 1], S[j], S[j + 1]
        out[mid] = sL + (Er[mid] - EkL) * (sR - sL) / (EkR - EkL)
    past = idx == n
    if past.any():
        k_last = np.searchsorted(Er, E[-1], side="left")
        out[past] = out[k_last]
    return out


def _table_on_grid_scan(E, S, Er):
    out = np.zeros_like(Er)
    sig_p = 0.0
    for ip, Ekp in enumerate(Er):
        for j in range(len(E)):
            EkL, sL = E[j], S[j]
            EkR = E[j + 1] if j + 1 < len(E) else 0.0
            sR = S[j + 1] if j + 1 < len(E) else 0.0
            if Ekp < EkL:
                sig_p = sL
                break
            if EkL <= Ekp <= EkR:
                with np.errstate(divide="ignore", invalid="ignore"):
                    sig_p = sL + (Ekp - EkL) * (sR - sL) / (EkR - EkL)
                break
        out[ip] = sig_p
    return out


def ordering(chem: Chemistry) -> dict:
    rs = chem.reactions
    ncol = len(rs)
    Er = np.unique(np.concatenate([r.E for r in rs])) if ncol else np.zeros(2)
    npt = len(Er)
    sig = np.zeros((ncol, npt))
    for c, r in enumerate(rs):
        sig[c] = _table_on_grid(np.asarray(r.E, float), np.asarray(r.S, float), Er)
    nt = chem.ntype_all
    masses = chem.masses()
    sigv = np.zeros((nt + 1, max(ncol, 1)))
    p_ncol = np.zeros(nt + 1, np.int32)
    sig_list = np.zeros((nt + 1, max(ncol, 1)), np.int32)
    for c, r in enumerate(rs):
        r1, r2 = r.reactants[0], r.reactants[1]
        sig_list[r1, p_ncol[r1]] = c + 1
        p_ncol[r1] += 1
        mu = abs(masses[r1]) * abs(masses[r2]) / (abs(masses[r1]) + abs(masses[r2]))
        vr = np.sqrt(2.0 * Er * QE / mu)
        keep = Er >= 0.1
        if keep.any():
            sigv[r1, c] = np.max(vr[keep] * sig[c][keep])
    col_info = np.array([r.col_info() for r in rs], np.int32).reshape(ncol, 10)
    Eth = np.array([r.Eth for r in rs], float)
    return dict(sig_Er=Er, sig=sig, sigv_mx=sigv, p_ncol=p_ncol, sig_list=sig_list,
                col_info=col_info, Eth=Eth, npt=npt, ncol=ncol)
