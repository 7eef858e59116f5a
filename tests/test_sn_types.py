"""Standard-SN and Galerkin-M angular operators (angular_polynomial_basis.jl:90-117 for 1D,
:139-159 and the Galerkin branch for 2D/3D).

CPU: the oracle's operators against their closed forms -- standard SN is
M[n, p] = (2l+1)/(4 pi) R_p(Omega_n), D[p, n] = w_n R_p(Omega_n) (1D: (2l+1)/2 P_l(mu_n) and
w_n P_l(mu_n)), with R from scipy; 1D Galerkin-M is M[n, l] = (2l+1)/2 P_l(mu_n), D = M^-1 --
and the product's host setup against the oracle.  GPU: full solves against the oracle."""
import math
import sys

import numpy as np
import pytest
from scipy import special

from util import rel_err

PI = math.pi


def _scipy_rlm(l, m, mu, phi):
    am = abs(m)
    C = math.sqrt((2 - (m == 0)) * math.factorial(l - am) / math.factorial(l + am))
    P = (-1) ** am * special.lpmv(am, l, mu)
    T = math.cos(m * phi) if m >= 0 else math.sin(am * phi)
    return C * P * T


def test_standard_1d_closed_form(oracle):
    mu, eta, xi, w = oracle.quadrature(16, "gauss-legendre", 1, 1)
    L = 5
    Np, M, D, pl, pm = oracle.basis(mu, eta, xi, w, L, "standard", 1)
    assert Np == L + 1 and list(pl) == list(range(L + 1))
    Mref = np.array([[(2 * l + 1) / 2 * special.eval_legendre(l, m) for l in range(Np)] for m in mu])
    Dref = np.array([[w[n] * special.eval_legendre(l, mu[n]) for n in range(len(mu))] for l in range(Np)])
    assert np.abs(M - Mref).max() < 1e-13 and np.abs(D - Dref).max() < 1e-13
    # discrete orthogonality of GL(16) up to degree 31: D M = I exactly for L = 5
    assert np.abs(D @ M - np.eye(Np)).max() < 1e-13


@pytest.mark.parametrize("qt,N,nd,Q,L", [("gauss-legendre-chebychev", 8, 2, 3, 3), ("carlson", 8, 3, 3, 2),
                                          ("gauss-legendre-chebychev", 4, 2, 2, 2)])
def test_standard_multid_closed_form(oracle, qt, N, nd, Q, L):
    mu, eta, xi, w = oracle.quadrature(N, qt, nd, Q)
    Np, M, D, pl, pm = oracle.basis(mu, eta, xi, w, L, "standard", Q)
    lm = [(l, m) for l in range(L + 1) for m in range(-l, l + 1) if not (Q == 2 and (l + m) % 2 == 1)]
    assert Np == len(lm) and list(zip(pl, pm)) == lm
    for i, (l, m) in enumerate(lm):
        R = np.array([_scipy_rlm(l, m, mu[n], math.atan2(xi[n], eta[n])) for n in range(len(mu))])
        assert np.abs(M[:, i] - (2 * l + 1) / (4 * PI) * R).max() < 1e-12
        assert np.abs(D[i] - w * R).max() < 1e-12


def test_galerkin_m_1d_closed_form(oracle):
    mu, eta, xi, w = oracle.quadrature(8, "gauss-legendre", 1, 1)
    Np, M, D, pl, pm = oracle.basis(mu, eta, xi, w, 7, "galerkin-m", 1)
    assert Np == 8
    Mref = np.array([[(2 * l + 1) / 2 * special.eval_legendre(l, m) for l in range(Np)] for m in mu])
    assert np.abs(M - Mref).max() < 1e-13
    assert np.abs(D @ M - np.eye(Np)).max() < 1e-11
    # for Gauss-Legendre the inverse is the quadrature projection: D[l, n] = w_n P_l(mu_n)
    Dref = np.array([[w[n] * special.eval_legendre(l, mu[n]) for n in range(Np)] for l in range(Np)])
    assert np.abs(D - Dref).max() < 1e-11


@pytest.mark.parametrize("sn_type", ["standard", "galerkin-m"])
@pytest.mark.parametrize("qt,N,nd,Q,L", [("gauss-legendre", 16, 1, 1, 5), ("gauss-legendre-chebychev", 8, 2, 3, 3),
                                          ("carlson", 8, 3, 3, 2)])
def test_host_setup_matches_oracle_sn_types(oracle, sn_type, qt, N, nd, Q, L):
    from radiant_hip import angular
    A = angular.AngularSet(qt, N, nd, Q if nd == 3 or qt == "gauss-legendre" else 0, L, sn_type)
    mu, eta, xi, w = oracle.quadrature(N, qt, nd, A.qdims)
    Np, M, D, pl, pm = oracle.basis(mu, eta, xi, w, L, sn_type, A.qdims)
    assert np.array_equal(pl, A.pl) and np.array_equal(pm, A.pm)
    cond = np.linalg.cond(M) if sn_type == "galerkin-m" else 1.0
    assert rel_err(A.M, M) < 1e-13
    assert rel_err(A.D, D) < 1e-15 * cond * 100 + 1e-13
    for f in range(2 * nd):
        nps, Ms, Ds, n2p = oracle.surface_basis(mu, eta, xi, w, L, sn_type, A.qdims, f)
        assert nps == A.faces[f][0] and np.array_equal(n2p, A.faces[f][3])
        assert rel_err(A.faces[f][1], Ms) < 1e-12 * max(1.0, cond)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("sn_type", ["standard", "galerkin-m"])
@pytest.mark.parametrize("ndims", [1, 2, 3])
def test_sn_types_solve_parity(oracle, sn_type, ndims):
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(ndims, N=6 if ndims == 3 else 8, G=2, L=2, tol=1e-12, seed=41 + ndims)
    prob.sn_type = sn_type
    ref, _ = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("sn type", sn_type, ndims, e, phi.shape)
    assert e < 1e-10 and all(s["converged"] for s in st)
    # and with a reflective face (faithful FP64 state)
    prob.bc = np.array([1] + [0] * (2 * ndims - 1), np.int32)
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("sn type", sn_type, ndims, "reflective x-", e)
    assert e < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("sn_type", ["standard", "galerkin-m"])
def test_sn_types_mixed_precision(oracle, sn_type):
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(2, N=8, G=2, L=2, tol=1e-12, seed=44)
    prob.sn_type = sn_type
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
    e = rel_err(phi, ref)
    _say("sn type mixed", sn_type, e)
    assert e < 1e-6
