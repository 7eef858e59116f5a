"""Gauss-Lobatto (tools/gauss_lobatto.jl) and Lebedev (tools/lebedev.jl: orders 3-11 in closed
form, orders 13-125 from the published Lebedev-Laikov orbit parameters) quadratures.

CPU: nodes and weights against numpy (Lobatto: +-1 and the roots of P'_{N-1}, weights
2/(N(N-1) P_{N-1}^2)), exactness (Lobatto: monomials up to degree 2N-3; Lebedev: every spherical
harmonic up to the rule's order, from scipy, and not the next degree), the host setup equal to the
oracle's, and the Galerkin identities D M = I.  GPU: full solves against the oracle."""
import math
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy import special

from util import rel_err

PI = math.pi


@pytest.mark.parametrize("N", list(range(2, 25)))
def test_gauss_lobatto_vs_numpy(oracle, N):
    from radiant_hip import angular
    x, w = angular.gauss_lobatto(N)
    c = np.zeros(N)
    c[N - 1] = 1.0  # P_{N-1}
    inner = np.sort(npleg.legroots(npleg.legder(c))) if N > 2 else np.zeros(0)
    ref = np.concatenate([[-1.0], inner, [1.0]])
    assert np.abs(np.sort(x) - ref).max() < 1e-14
    wref = 2.0 / (N * (N - 1) * npleg.legval(x, c) ** 2)
    assert np.abs(w - wref).max() < 1e-14
    for k in range(2 * N - 2):  # exact to degree 2N-3
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-13
    mu, eta, xi, wo = oracle.quadrature(N, "gauss-lobatto", 1, 1)
    assert np.abs(mu - x).max() < 1e-15 and np.abs(wo - w).max() < 1e-15


def _ylm(l, m, th, ph):
    if hasattr(special, "sph_harm_y"):
        return special.sph_harm_y(l, m, th, ph)
    return special.sph_harm(m, l, ph, th)


@pytest.mark.parametrize("order", [3, 5, 7, 9, 11])
def test_lebedev_exactness(oracle, order):
    from radiant_hip import angular
    mu, eta, xi, w = angular.lebedev(order)
    assert np.abs(mu ** 2 + eta ** 2 + xi ** 2 - 1).max() < 1e-15
    assert abs(w.sum() - 4 * PI) < 1e-13
    th = np.arccos(np.clip(xi, -1, 1))
    ph = np.arctan2(eta, mu)
    for l in range(1, order + 1):
        for m in range(0, l + 1):
            assert abs((w * _ylm(l, m, th, ph)).sum()) < 1e-13
    worst = max(abs((w * _ylm(order + 1, m, th, ph)).sum()) for m in range(order + 2))
    assert worst > 1e-3  # the next degree is not integrated: the rule is of this order, not higher
    a, b, c, d = oracle.quadrature(order, "lebedev", 3, 3)
    assert np.abs(a - mu).max() < 1e-15 and np.abs(b - eta).max() < 1e-15
    assert np.abs(c - xi).max() < 1e-15 and np.abs(d - w).max() < 1e-15


def _lebedev_orders():
    from radiant_hip.lebedev_rules import ORDER_POINTS
    return sorted(o for o in ORDER_POINTS if o > 11)


@pytest.mark.parametrize("order", _lebedev_orders())
def test_lebedev_table_rules_exact(order):
    """Every table-based rule (13 to 125, 74 to 5294 points): unit directions, weights summing to
    4 pi, the point count of the rule, invariance under the octahedral group (every coordinate
    permutation and sign change maps the set onto itself with equal weights), and exactness for the
    spherical harmonics up to its order (every (l, m) up to order 35; above, a spread subset of
    degrees and orders up to the rule's order)."""
    from radiant_hip import angular
    from radiant_hip.lebedev_rules import ORDER_POINTS
    mu, eta, xi, w = angular.lebedev(order)
    assert len(w) == ORDER_POINTS[order]
    assert np.abs(mu ** 2 + eta ** 2 + xi ** 2 - 1).max() < 1e-15
    assert abs(w.sum() - 4 * PI) < 1e-12
    pts = np.c_[mu, eta, xi]
    key = {tuple(np.round(p, 12)): wi for p, wi in zip(pts, w)}
    for perm in ((1, 0, 2), (2, 1, 0), (0, 2, 1)):
        for sg in ((-1, 1, 1), (1, -1, 1), (1, 1, -1)):
            q = pts[:, perm] * np.array(sg)
            for p, wi in zip(q, w):
                assert abs(key[tuple(np.round(p, 12))] - wi) < 1e-15
    th = np.arccos(np.clip(xi, -1, 1))
    ph = np.arctan2(eta, mu)
    ls = range(1, order + 1) if order <= 35 else sorted(set(range(1, order + 1, 9)) | {order - 1, order})
    for l in ls:
        ms = range(0, l + 1) if order <= 35 else sorted(set(range(0, l + 1, 7)) | {l})
        for m in ms:
            assert abs((w * _ylm(l, m, th, ph)).sum()) < 1e-12, (l, m)
    if order <= 35:  # the rule is of this order, not higher
        worst = max(abs((w * _ylm(order + 1, m, th, ph)).sum()) for m in range(order + 2))
        assert worst > 1e-4


@pytest.mark.parametrize("N", [13, 17, 21])
def test_lebedev_table_galerkin_and_oracle_setup(oracle, N):
    """Galerkin-D on the table-based rules: D M = I, and the oracle (given the rule point by point)
    builds the same basis as the host setup."""
    from radiant_hip import angular
    A = angular.AngularSet("lebedev", N, 3, 3, 3, "galerkin-d")
    Np, M, D, pl, pm = oracle.basis(A.mu, A.eta, A.xi, A.w, 3, "galerkin-d", 3)
    assert np.array_equal(pl, A.pl) and np.array_equal(pm, A.pm)
    cond = np.linalg.cond(D)
    assert cond < 1e4 and np.abs(D @ M - np.eye(Np)).max() < 1e-13 * cond
    assert rel_err(A.D, D) < 1e-13


@pytest.mark.parametrize("qt,N,nd,Q,L", [("gauss-lobatto", 8, 1, 1, 7), ("gauss-lobatto", 7, 1, 1, 6),
                                          ("lebedev", 7, 3, 3, 3), ("lebedev", 11, 3, 3, 3),
                                          ("lebedev", 9, 2, 3, 3)])
def test_new_quadratures_host_matches_oracle(oracle, qt, N, nd, Q, L):
    from radiant_hip import angular
    A = angular.AngularSet(qt, N, nd, Q, L, "galerkin-d")
    mu, eta, xi, w = oracle.quadrature(N, qt, nd, A.qdims)
    assert np.abs(A.mu - mu).max() < 1e-15 and np.abs(A.w - w).max() < 1e-14
    Np, M, D, pl, pm = oracle.basis(mu, eta, xi, w, L, "galerkin-d", A.qdims)
    assert np.array_equal(pl, A.pl) and np.array_equal(pm, A.pm)
    cond = np.linalg.cond(D)
    assert np.abs(D @ M - np.eye(Np)).max() < 1e-13 * cond
    assert rel_err(A.D, D) < 1e-13 and rel_err(A.M, M) < 1e-15 * cond * 100


def test_lobatto_solve_close_to_legendre(oracle):
    """A 1D solve on Gauss-Lobatto(8) is the same physics as on Gauss-Legendre(8): the scalar
    fluxes differ by the angular discretisation only (about 1%), and in 1D standard SN and
    Galerkin-D give the same flux (both are exact Legendre projections on the L+1 moments)."""
    from radiant_hip import problems as P
    out = {}
    for qt in ("gauss-legendre", "gauss-lobatto"):
        for st in ("galerkin-d", "standard"):
            p = P.small_problem(1, N=8, G=2, L=2, quad_type=qt, quad_order=8, tol=1e-12, seed=52)
            p.sn_type = st
            out[qt, st] = oracle.solve(p)[0][:, 0]
    assert rel_err(out["gauss-lobatto", "galerkin-d"], out["gauss-legendre", "galerkin-d"]) < 0.05
    assert rel_err(out["gauss-lobatto", "standard"], out["gauss-lobatto", "galerkin-d"]) < 1e-12
    assert out["gauss-lobatto", "galerkin-d"].min() > 0


@pytest.mark.parametrize("order", [5, 7, 9, 11])
def test_lebedev_upper_hemisphere_refused_alike(oracle, order):
    """Folded onto the upper hemisphere (Qdims = 2), Lebedev rules of order >= 5 keep in-plane points
    (xi = 0), and no face basis of independent half-range harmonics exists for them: the host setup
    and the oracle both refuse, with the same message."""
    from radiant_hip import angular
    with pytest.raises(RuntimeError, match="Surface Galerkin"):
        angular.AngularSet("lebedev", order, 2, 2, 3, "galerkin-d")
    mu, eta, xi, w = oracle.quadrature(order, "lebedev", 2, 2)
    with pytest.raises(RuntimeError, match="Surface Galerkin"):
        for f in range(4):
            oracle.surface_basis(mu, eta, xi, w, 3, "galerkin-d", 2, f)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("ndims,qt,N", [(1, "gauss-lobatto", 8), (1, "gauss-lobatto", 7), (2, "lebedev", 7),
                                        (3, "lebedev", 5), (3, "lebedev", 11), (3, "lebedev", 13),
                                        (3, "lebedev", 17), (2, "lebedev", 21)])
def test_new_quadratures_solve_parity(oracle, ndims, qt, N):
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(ndims, N=6 if ndims == 3 else 8, G=2, L=2, quad_type=qt, quad_order=N, tol=1e-12,
                           seed=51 + ndims)
    if qt == "lebedev":
        prob.qdims = 3  # full sphere (the upper-hemisphere fold has no surface Galerkin basis)
    ref, _ = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("quadrature", qt, N, ndims, e, phi.shape)
    assert e < 1e-10 and all(s["converged"] for s in st)
    mx, _ = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("quadrature", qt, N, ndims, "mixed-fp32", e)
    assert e < 1e-6
