"""Surface sources of the SN solve: particle_transport/surface_source.jl:22-408 (the SN branch)
and their accumulation in structures/Source.jl:152-190.

A surface source is a beam of the given intensity and direction entering through one face of the
Cartesian grid, in one energy group.  The solve receives it as half-range moments of that face:
Q[g, p, pos] = intensity * Y_p(|mu_s|, phi_s) (times the face area of the position's cell in 2D
and 3D), where Y_p is the half-range Legendre polynomial (1D quadratures) or real half-range
spherical harmonic of the face basis's p-th moment, and phi_s = atan(xi_s, eta_s).  As in the
reference, the harmonic is evaluated at the x-polar angles of the beam for every face.  Positions
enter when their cell centre lies within the source's transverse bounds.
"""
import math

import numpy as np

from .angular import angular_set, half_range_legendre_up_to_L, quadrature_dimension, real_half_range_sh_up_to_L

FACES = {"X-": 0, "X+": 1, "Y-": 2, "Y+": 3, "Z-": 4, "Z+": 5}


def _centres(widths):
    w = np.asarray(widths, np.float64)
    return np.cumsum(w) - w / 2


def surface_source_moments(prob, location, direction, intensity=1.0, group=0, bounds=None, legendre_order=None):
    """(Q [G, Np, npos], norm) of one surface source (surface_source.jl:22-408).

    location: "X-", "X+", "Y-", "Y+", "Z-" or "Z+".  direction: (mu, eta, xi) of the beam.
    group: 0-based energy group.  bounds: {"x": (lo, hi), "y": ..., "z": ...} for the transverse
    axes of the face (cell centres measured from the grid's lower corner).  legendre_order: the
    source's order (default: the solver's)."""
    surface = location.upper()
    if surface not in FACES:
        raise ValueError(f"Unknown surface {location}")
    f = FACES[surface]
    nd = prob.ndims
    if f >= 2 * nd:
        raise ValueError(f"Surface {surface} does not exist in {nd}D")
    qd = quadrature_dimension(prob.quad_type, nd, prob.qdims)
    ang = angular_set(prob.quad_type, prob.quad_order, nd, qd, prob.legendre_order, prob.sn_type)
    pl, pm = ang.face_lm[f]
    Lsrc = prob.legendre_order if legendre_order is None else min(legendre_order, prob.legendre_order)
    if prob.sn_type == "standard":
        keep = pl <= Lsrc
        pl, pm = pl[keep], pm[keep]
    Np = len(pl)
    Lmax = int(np.max(pl))
    mus, etas, xis = (float(v) for v in direction)
    G = prob.G
    # beam entering through this face?
    comp = {0: mus, 1: mus, 2: etas, 3: etas, 4: xis, 5: xis}[f]
    enters = comp > 0 if f % 2 == 0 else comp < 0
    if qd == 1:
        Y = half_range_legendre_up_to_L(Lmax, abs(mus))
        vals = np.array([Y[l] for l in pl])
    else:
        phis = math.atan2(xis, etas)
        R = real_half_range_sh_up_to_L(Lmax, abs(mus), phis)
        vals = np.array([R[l][l + m] for l, m in zip(pl, pm)])
    if nd == 1:
        npos = 1
        area = np.ones(1)
        inside = np.ones(1, bool)
    else:
        ax = f // 2
        x = _centres(prob.dx)
        y = _centres(prob.dy) if nd >= 2 else np.zeros(1)
        z = _centres(prob.dz) if nd >= 3 else np.zeros(1)
        wid = {"x": np.asarray(prob.dx, float), "y": np.asarray(prob.dy, float) if nd >= 2 else np.ones(1),
               "z": np.asarray(prob.dz, float) if nd >= 3 else np.ones(1)}
        cen = {"x": x, "y": y, "z": z}
        trans = [a for a in ("x", "y", "z")[:nd] if a != "xyz"[ax]]
        b = bounds or {}
        for a in trans:
            if a not in b:
                raise ValueError(f"surface source on {surface} needs bounds for the {a} axis")
        # positions: x faces iy + Ny*iz, y faces ix + Nx*iz, z faces ix + Nx*iy
        a1, a2 = (trans + [None])[:2]
        c1, w1 = cen[a1], wid[a1]
        in1 = (c1 >= b[a1][0]) & (c1 <= b[a1][1])
        if a2 is None:
            area = w1.copy()
            inside = in1
        else:
            c2, w2 = cen[a2], wid[a2]
            in2 = (c2 >= b[a2][0]) & (c2 <= b[a2][1])
            area = (w1[:, None] * w2[None, :]).reshape(-1, order="F")
            inside = (in1[:, None] & in2[None, :]).reshape(-1, order="F")
        npos = len(area)
    Q = np.zeros((G, Np, npos))
    norm = 0.0
    if enters:
        Q[group] = intensity * vals[:, None] * np.where(inside, area, 0.0)[None, :]
    if nd == 1:
        norm = intensity
    else:
        norm = intensity * float(np.sum(np.where(inside, area, 0.0)))
    return Q, norm


def add_surface_source(prob, location, direction, intensity=1.0, group=0, bounds=None, legendre_order=None):
    """Adds a surface source to prob (Source.jl:152-190: moments of the same face add up; a
    source with more moments pads the others with zeros).  Returns its normalisation factor."""
    Q, norm = surface_source_moments(prob, location, direction, intensity, group, bounds, legendre_order)
    f = FACES[location.upper()]
    if prob.ssrc is None:
        prob.ssrc = {}
    n = max(prob.Np_source, Q.shape[1])
    for k in list(prob.ssrc):
        if prob.ssrc[k].shape[1] < n:
            pad = np.zeros((prob.G, n - prob.ssrc[k].shape[1], prob.ssrc[k].shape[2]))
            prob.ssrc[k] = np.concatenate([prob.ssrc[k], pad], axis=1)
    if Q.shape[1] < n:
        Q = np.concatenate([Q, np.zeros((prob.G, n - Q.shape[1], Q.shape[2]))], axis=1)
    prob.ssrc[f] = prob.ssrc[f] + Q if f in prob.ssrc else Q
    prob.Np_source = n
    return norm
