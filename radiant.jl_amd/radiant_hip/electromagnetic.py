"""Lorentz-force operator of an external magnetic field, in moment space: the part of Radiant's
``compute_flux(::SN)`` set-up that builds it (sn_flux.jl:116-120 with
particle_transport/electromagnetic_scattering_matrix.jl), so the caller side of the C ABI, which
receives ℳ_EM ready to use (``rt_problem_desc.Mem``).

The pass adds ``ℳ_EM[g] @ Φ`` to the source of every cell and spatial moment
(sn_inner_pass.jl:26-32).  ℳ_EM[g] = -𝒯_g D M, where

* M[n, p] is the derivative of the real harmonic p at direction n along the rotation
  dΩ/dt = Ω × B of a charged particle, times charge c (2l + 1)/(4π): with the polar axis x,
  dμ/dt = B_z η - B_y ξ and dφ/dt = (B_y μ η + B_z μ ξ - B_x (1 - μ²)) / (1 - μ²), and the limits
  of the reference at |μ| = 1;
* 𝒯_g = 2/ΔE_g log((sqrt(E_g + 2 mc²) + sqrt(E_g)) / (sqrt(E_g+1 + 2 mc²) + sqrt(E_g+1))) is the
  group average of 1/(pc), the inverse momentum (E in MeV, mc² = 0.510999, c = 2.99792458).

Only magnetic fields are supported, as in the reference; λ₀_EM is zero (Σtot unchanged).
"""
import math

import numpy as np

C_LIGHT = 2.99792458
MC2 = 0.510999


def ferrer(l, m, x):
    """ferrer_associated_legendre.jl: P_l^m(x) without the (-1)^m factor."""
    if l < 0 or m < 0 or m > l or abs(x) > 1:
        raise ValueError("invalid associated Legendre arguments")
    pmm = 1.0
    if m > 0:
        pmm = float(np.prod(np.arange(1, 2 * m, 2))) * math.sqrt(1 - x * x) ** m
    if l == m:
        return pmm
    pm1 = (2 * m + 1) * x * pmm
    if l == m + 1:
        return pm1
    pm2 = pm1
    for li in range(m + 2, l + 1):
        pm2 = ((2 * li - 1) * x * pm1 - (li + m - 1) * pmm) / (li - m)
        pmm, pm1 = pm1, pm2
    return pm2


def group_factor(Eb):
    """𝒯_g of every group from the energy boundaries Eb [G + 1] (MeV, descending)."""
    Eb = np.asarray(Eb, np.float64)
    dE = Eb[:-1] - Eb[1:]
    num = np.sqrt(Eb[:-1] + 2 * MC2) + np.sqrt(Eb[:-1])
    den = np.sqrt(Eb[1:] + 2 * MC2) + np.sqrt(Eb[1:])
    return 2.0 / dE * np.log(num / den)


def _phi(eta, xi):
    """The azimuth of electromagnetic_scattering_matrix.jl:59-73 (polar axis x)."""
    if xi != 0:
        if eta > 0:
            return math.copysign(1.0, xi) * math.atan(abs(xi / eta))
        if eta == 0:
            return math.copysign(1.0, xi) * math.pi / 2
        return math.copysign(1.0, xi) * (math.pi - math.atan(abs(xi / eta)))
    return 0.0 if eta >= 0 else math.pi


def derivative_matrix(ang, B, charge):
    """M [Nd, Np] of electromagnetic_scattering_matrix.jl:57-117."""
    mu, eta, xi = (np.asarray(a, float) for a in (ang.mu, ang.eta, ang.xi))
    pl, pm = np.asarray(ang.pl), np.asarray(ang.pm)
    Nd, P = len(mu), len(pl)
    M = np.zeros((Nd, P))
    for n in range(Nd):
        phi = _phi(eta[n], xi[n])
        u = mu[n]
        for p in range(P):
            l, m = int(pl[p]), int(pm[p])
            am = abs(m)
            lf = sum(math.log(k) for k in range(1, l - am + 1)) - sum(math.log(k) for k in range(1, l + am + 1))
            Clm = math.sqrt((2 - (m == 0)) * math.exp(lf))
            Tm = math.cos(m * phi) if m >= 0 else math.sin(am * phi)
            nrm = (2 * l + 1) / (4 * math.pi) * charge * C_LIGHT
            v = 0.0
            if abs(u) != 1:  # derivative in μ
                if l > 0:
                    if l > am:
                        v += ((l + am) * ferrer(l - 1, am, u) - l * u * ferrer(l, am, u)) / (1 - u * u) * Clm * Tm * nrm * \
                            (B[2] * eta[n] - B[1] * xi[n])
                    elif l == m:
                        v += -l * u * ferrer(l, l, u) / (1 - u * u) * Clm * Tm * nrm * (B[2] * eta[n] - B[1] * xi[n])
            elif am == 1:
                v += -l * (l + 1) / 2 * u ** l * Clm * Tm * nrm * (B[2] * math.cos(phi) - B[1] * math.sin(phi))
            Plm = ferrer(l, am, u)  # derivative in φ
            dT = -m * math.sin(m * phi) if m >= 0 else am * math.cos(am * phi)
            if abs(u) != 1:
                fac = B[1] * u * eta[n] + B[2] * u * xi[n] - B[0] * (1 - u * u)
                v += Plm * Clm * dT * (2 * l + 1) / (4 * math.pi) * charge * C_LIGHT / (1 - u * u) * fac
            elif am == 1:
                fac = B[1] * math.cos(phi) + B[2] * math.sin(phi)
                v += l * (l + 1) / 2 * u ** (l + 1) * Clm * dT * (2 * l + 1) / (4 * math.pi) * charge * C_LIGHT * fac
            M[n, p] = v
    return M


def moment_operator(ang, B, charge, Eb, qdims, E=(0.0, 0.0, 0.0)):
    """ℳ_EM [G, Np, Np] (and λ₀_EM = zeros(G)) for a magnetic field B [3] and the particle charge."""
    if any(v != 0 for v in E):
        raise ValueError("Electric-field transport not yet implemented; only magnetic fields are supported.")
    if qdims != 3:
        raise ValueError("Quadrature on the unit sphere is required for electromagnetic transport.")
    G = -(np.asarray(ang.D) @ derivative_matrix(ang, B, charge))
    T = group_factor(Eb)
    return T[:, None, None] * G[None], np.zeros(len(T))
