"""Pair geometries placed on purpose for the per-pair parity tests.

Pair t is particles (2t, 2t+1).  Particle 2t+1 sits at distance d_t from
particle 2t along a random direction, with d_t drawn across the range where
the kind's terms change character: just inside contact, the well, the
switch band of the smoothed kinds, the weld band of the welded kinds, and a
little beyond the cutoff.  For the anisotropic kinds d_t is measured from
the pair's own contact distance for its orientations and direction: the
Gay-Berne sigma of particle i's shape (force.f90:351-366) or the converged
Perram-Wertheim contact distance (the oracle's contact solve).  A quarter of
the pairs straddle the periodic boundary.

Tolerance helper: per element |a - b| <= tol * (|b| + median |b|), the median
over the non-zero reference values, so that every ordinary pair is checked
at its own magnitude.
"""
from __future__ import annotations

import numpy as np

from oracle.oracle import (COULOMB, DH, DIPOLE, ECP, ECPC, ECPR, ECPW, ECPWC, GB, GBC, GBR, GBW, GBWC, LJ, LJ12, LJW,
                           MORSE, WCA, State)

GB_KINDS = (GB, GBR, GBW, GBC, GBWC)
ECP_KINDS = (ECP, ECPR, ECPW, ECPC, ECPWC)
WELDED = (LJW, GBW, GBWC, ECPW, ECPWC)


def unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def body_z(q):
    """The body z axis of quaternions q (w x y z), as quat_u (force.f90:343-349)."""
    w, x, y, z = q.T
    return np.stack([2.0 * (x * z + w * y), 2.0 * (y * z - w * x), w * w - x * x - y * y + z * z], 1)


def gb_sigma(sig_x, sig_z, u1, u2, rhat):
    """Gay-Berne range parameter sigma of particle i's shape for unit
    orientation vectors u1, u2 and unit direction rhat."""
    chi = (sig_z ** 2 - sig_x ** 2) / (sig_z ** 2 + sig_x ** 2)
    ru1 = np.sum(rhat * u1, 1)
    ru2 = np.sum(rhat * u2, 1)
    uu = np.sum(u1 * u2, 1)
    g = 1.0 - chi / 2.0 * ((ru1 + ru2) ** 2 / (1 + chi * uu) + (ru1 - ru2) ** 2 / (1 - chi * uu))
    return sig_x / np.sqrt(g)


def _wrap(x, L):
    return x - np.round(x / L) * L


def _range(kind, params):
    """(offset low, offset high) of the pair distance about contact (about 0
    for the isotropic kinds, where it is the distance itself)."""
    if kind in (LJ, LJ12):
        return 0.9 * params[0], params[2] + 0.2
    if kind == LJW:
        return 0.9 * params[0], params[2] + params[3] + 0.2
    if kind == WCA:            # cut at r = sigma (force.f90:155-190)
        return 0.86 * params[0], 1.08 * params[0]
    if kind == MORSE:
        return 0.75 * params[2], params[3] + 0.2
    if kind in (DH, COULOMB, DIPOLE):
        return 0.6, params[0] + 0.3
    wf = 0.0
    if kind in (GBW, GBWC):
        wf = params[3]
    if kind in (ECPW, ECPWC):
        wf = params[1]
    if kind == GBR:
        return -0.06, 0.05
    if kind == ECPR:
        return -0.05, 0.04
    return -0.06, params[0] - 1.0 + wf + 0.3   # reach sigma + cutoff (+ wf) - sigma_0, sigma_0 = 1


def pair_state(oracle, rng, kind, params, npairs, base, L=40.0):
    """A State of 2*npairs particles (shapes and the other per-particle
    values from base(rng, n, L)) with each pair placed as described above.
    Returns (state, i, j, d, contact) with 0-based indices."""
    n = 2 * npairs
    st = base(rng, n, L)
    i = np.arange(0, n, 2)
    j = i + 1
    rhat = unit_vectors(rng, npairs)
    pi = rng.uniform(-L / 2, L / 2, (npairs, 3))
    edge = np.arange(npairs) % 4 == 0      # a quarter of the pairs across the boundary
    pi[edge, 0] = L / 2 - rng.uniform(0.0, 1.0, edge.sum())
    rhat[edge, 0] = np.abs(rhat[edge, 0])
    q = np.stack([st[f] for f in ("qw", "qx", "qy", "qz")], 1)
    if kind in GB_KINDS:
        contact = gb_sigma(st["sig_x"][i], st["sig_z"][i], body_z(q[i]), body_z(q[j]), rhat)
    elif kind in ECP_KINDS:
        contact = np.zeros(npairs)
        for k, f in enumerate(("rx", "ry", "rz")):
            st[f][i] = 0.0
            st[f][j] = rhat[:, k]
        for t in range(npairs):
            ar, br, r = oracle.ecp_shape(st, i[t], j[t])
            lam, _ = oracle.lambda_converged(ar, br, r)
            contact[t] = oracle.contact_distance(ar, br, r, lam)
    else:
        contact = np.zeros(npairs)
    lo, hi = _range(kind, params)
    d = contact + rng.uniform(lo, hi, npairs)
    pj = pi + d[:, None] * rhat
    for k, f in enumerate(("rx", "ry", "rz")):
        st[f][i] = _wrap(pi[:, k], L)
        st[f][j] = _wrap(pj[:, k], L)
    return st, i, j, d, contact


def band_counts(kind, params, d, contact):
    """Pairs in the switch band and in the weld band, by the placement
    distances (the Gay-Berne contact is exact; the Perram-Wertheim one is the
    converged contact, which the reference's Brent solve approximates)."""
    out = {}
    wf = 0.0
    if kind in (GBW, GBWC):
        wf = params[3]
    elif kind in (ECPW, ECPWC):
        wf = params[1]
    elif kind == LJW:
        wf = params[3]
    if kind in GB_KINDS + ECP_KINDS and kind not in (GBR, ECPR):
        rlim = contact + params[0] - 1.0 + wf
        out["switch"] = int(np.sum((d >= 0.9 * rlim) & (d < rlim)))
    if kind == COULOMB:
        out["switch"] = int(np.sum((d >= params[0] - 2.0) & (d < params[0])))
    if kind == LJW:
        out["weld"] = int(np.sum((d >= params[0]) & (d <= params[0] + wf)))
    elif kind in WELDED:
        out["weld"] = int(np.sum((d >= contact) & (d <= contact + wf)))
    return out


def close(a, b):
    """Worst |a - b| / (|b| + median of the non-zero |b|) over the elements."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.size == 0:
        return 0.0
    nz = np.abs(b[b != 0])
    med = float(np.median(nz)) if nz.size else 1.0
    return float(np.max(np.abs(a - b) / (np.abs(b) + med)))
