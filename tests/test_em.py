"""External magnetic field (Electromagnetic_Field; sn_flux.jl:116-120, sn_inner_pass.jl:26-32,
electromagnetic_scattering_matrix.jl, ferrer_associated_legendre.jl).

CPU pins of the operator ℳ_EM[g] = -𝒯_g D M:
- the Ferrer associated Legendre functions of host and oracle against scipy (lpmv without its
  Condon-Shortley sign);
- 𝒯_g against the group average of 1/(pc) = 1/sqrt(E (E + 2 mc²)) by numerical quadrature;
- M[n, p] against the chain rule of the real harmonic p along the rotation dΩ/dt = Ω × B (partial
  derivatives by central differences, the rates checked on Rodrigues' rotation about B), times
  charge c (2l+1)/(4π), at every direction off the polar axis; for m = -l the reference's formula
  leaves the ∂Y/∂μ term out, and so does the restatement; at the poles its limits equal the
  rotation's derivative except for the sign of the m < 0 rows at μ = -1, kept as well;
- host (numpy) and oracle (C++) ℳ_EM agree on four quadrature sets; its l = 0 column vanishes;
- the pass adds ℳ_EM Φ to the source of every cell (an oracle pass with the field equals one without
  it whose source carries ℳ_EM z; ℳ_EM is not symmetric, so a transposed operator fails);
- an all-reflective uniform medium keeps the isotropic flux of the closed form: the field only
  turns directions.
GPU: FP64 solves (faithful and reduced state) within 1e-10 of the oracle and mixed FP32 within 1e-6,
for BTE and for BFP with the Fokker-Planck source before the field's; a batch holding such a problem
solves it on its own stream (the lockstep solver does not take it)."""
import math
import sys

import numpy as np
import pytest

from util import rel_err

B_FIELD = [0.3, -0.2, 0.5]
# the solves take a field ten times weaker: with B_FIELD the Lorentz operator outweighs Σt and the
# oracle's GMRES leaves groups unconverged after 1000 iterations
B_SOLVE = [0.03, -0.02, 0.05]
EB = np.array([2.0, 1.0, 0.5, 0.1])


def _field(p, B=B_FIELD, charge=-1.0):
    p.em_B, p.em_charge, p.Eb = list(B), charge, EB[:p.G + 1].copy()
    return p


def _ang(p):
    from radiant_hip import angular
    qd = angular.quadrature_dimension(p.quad_type, p.ndims, p.qdims)
    return angular.angular_set(p.quad_type, p.quad_order, p.ndims, qd, p.legendre_order, p.sn_type), qd


def test_ferrer(oracle):
    from scipy.special import lpmv
    from radiant_hip import electromagnetic as EM
    for l in range(7):
        for m in range(l + 1):
            for x in (-1.0, -0.73, 0.0, 0.31, 0.999, 1.0):
                want = (-1) ** m * lpmv(m, l, x)
                assert abs(EM.ferrer(l, m, x) - want) <= 1e-12 * max(1.0, abs(want))
                assert abs(oracle.ferrer(l, m, x) - want) <= 1e-12 * max(1.0, abs(want))


def test_group_factor():
    from scipy.integrate import quad
    from radiant_hip import electromagnetic as EM
    Eb = np.array([10.0, 2.0, 0.5, 0.01, 1e-4])
    T = EM.group_factor(Eb)
    for g in range(len(Eb) - 1):
        v, _ = quad(lambda E: 1.0 / math.sqrt(E * (E + 2 * EM.MC2)), Eb[g + 1], Eb[g], epsabs=0, epsrel=1e-13)
        assert abs(T[g] - v / (Eb[g] - Eb[g + 1])) <= 1e-11 * T[g]


def _real_sh(l, m, mu, phi):
    """The real harmonic of the basis (polar axis x), from scipy's lpmv."""
    from scipy.special import lpmv
    am = abs(m)
    C = math.sqrt((2 - (m == 0)) * math.factorial(l - am) / math.factorial(l + am))
    return C * (-1) ** am * lpmv(am, l, mu) * (math.cos(m * phi) if m >= 0 else math.sin(am * phi))


@pytest.mark.parametrize("qt,N", [("carlson", 4), ("lebedev", 7), ("gauss-legendre-chebychev", 4)])
def test_derivative_matrix_is_the_lorentz_rotation(qt, N):
    """M[n, p] = charge c (2l+1)/(4π) (∂Y/∂μ dμ/dt + ∂Y/∂φ dφ/dt) with dΩ/dt = Ω × B, the partial
    derivatives by central differences of scipy's harmonics, and dμ/dt, dφ/dt checked against a
    finite difference of Rodrigues' rotation about B.  For m = -l the reference leaves the ∂Y/∂μ
    term out (electromagnetic_scattering_matrix.jl:83-93 tests l == m, not l == |m|), and so does
    the restatement: that term is dropped from the expectation there."""
    from radiant_hip import electromagnetic as EM, problems as P
    p = P.small_problem(3, N=3, G=1, L=2, quad_type=qt, quad_order=N)
    ang, _ = _ang(p)
    charge = -1.0
    M = EM.derivative_matrix(ang, B_FIELD, charge)
    Bv = np.asarray(B_FIELD)
    nB = np.linalg.norm(Bv)
    k = Bv / nB
    t, h = 1e-5, 1e-6
    checked = 0
    for n in range(len(ang.mu)):
        O = np.array([ang.mu[n], ang.eta[n], ang.xi[n]])
        if abs(abs(O[0]) - 1) < 1e-12:
            continue   # on the polar axis φ is undefined; the reference's limits there are not pinned here
        mu, phi = O[0], math.atan2(O[2], O[1])
        dmu = Bv[2] * O[1] - Bv[1] * O[2]
        dphi = (Bv[1] * mu * O[1] + Bv[2] * mu * O[2] - Bv[0] * (1 - mu * mu)) / (1 - mu * mu)
        # the rates are those of the rotation dΩ/dt = Ω × B = -B × Ω (by -|B| t about B)
        rot = lambda a: O * math.cos(a) + np.cross(k, O) * math.sin(a) + k * np.dot(k, O) * (1 - math.cos(a))
        Op, Om = rot(-nB * t), rot(nB * t)
        assert abs((Op[0] - Om[0]) / (2 * t) - dmu) < 1e-8
        dang = math.atan2(Op[2], Op[1]) - math.atan2(Om[2], Om[1])
        dang = (dang + math.pi) % (2 * math.pi) - math.pi
        assert abs(dang / (2 * t) - dphi) < 1e-6 * max(1.0, abs(dphi))
        for q in range(len(ang.pl)):
            l, m = int(ang.pl[q]), int(ang.pm[q])
            dY_dmu = (_real_sh(l, m, mu + h, phi) - _real_sh(l, m, mu - h, phi)) / (2 * h)
            dY_dphi = (_real_sh(l, m, mu, phi + h) - _real_sh(l, m, mu, phi - h)) / (2 * h)
            rate = (0.0 if (l > 0 and m == -l) else dY_dmu * dmu) + dY_dphi * dphi
            want = charge * EM.C_LIGHT * (2 * l + 1) / (4 * math.pi) * rate
            assert abs(M[n, q] - want) <= 1e-6 * max(1.0, abs(want)), (n, l, m, M[n, q], want)
        checked += 1
    assert checked >= len(ang.mu) - 2


def test_derivative_matrix_at_the_poles():
    """The rows at the polar-axis directions (μ = ±1, where φ is undefined and the reference uses
    limit formulas, electromagnetic_scattering_matrix.jl:94-116) against a finite difference of the
    harmonics along Rodrigues' rotation: equal at μ = +1, and at μ = -1 for m >= 0; at μ = -1 the
    reference's limits give the m < 0 rows the opposite sign, and the restatement keeps that."""
    from radiant_hip import electromagnetic as EM, problems as P
    p = P.small_problem(3, N=3, G=1, L=3, quad_type="lebedev", quad_order=7)
    ang, _ = _ang(p)
    M = EM.derivative_matrix(ang, B_FIELD, -1.0)
    Bv = np.asarray(B_FIELD)
    nB = np.linalg.norm(Bv)
    k = Bv / nB
    t = 1e-4
    poles = 0
    for n in range(len(ang.mu)):
        O = np.array([ang.mu[n], ang.eta[n], ang.xi[n]])
        if abs(abs(O[0]) - 1) > 1e-12:
            continue
        poles += 1
        rot = lambda a: O * math.cos(a) + np.cross(k, O) * math.sin(a) + k * np.dot(k, O) * (1 - math.cos(a))
        Op, Om = rot(-nB * t), rot(nB * t)
        for q in range(len(ang.pl)):
            l, m = int(ang.pl[q]), int(ang.pm[q])
            Y = lambda v: _real_sh(l, m, max(-1.0, min(1.0, v[0])), math.atan2(v[2], v[1]))
            want = -1.0 * EM.C_LIGHT * (2 * l + 1) / (4 * math.pi) * (Y(Op) - Y(Om)) / (2 * t)
            if O[0] < 0 and m < 0:
                want = -want
            assert abs(M[n, q] - want) <= 1e-6 * max(1.0, abs(want)), (O[0], l, m, M[n, q], want)
    assert poles == 2


@pytest.mark.parametrize("nd,qt,N", [(3, "carlson", 4), (3, "lebedev", 5), (2, "gauss-legendre-chebychev", 4),
                                     (3, "gauss-legendre-chebychev", 4)])
def test_operator_host_matches_oracle(oracle, nd, qt, N):
    from radiant_hip import electromagnetic as EM, problems as P
    p = _field(P.small_problem(nd, N=3, G=3, L=2, quad_type=qt, quad_order=N))
    ang, qd = _ang(p)
    Mh, lam = EM.moment_operator(ang, p.em_B, p.em_charge, p.Eb, qd)
    Mo = oracle.em_matrix(p)
    assert Mh.shape == Mo.shape and not lam.any()
    assert np.abs(Mh - Mo).max() <= 1e-13 * np.abs(Mh).max()
    assert not Mh[:, :, 0].any()             # the derivative of the constant harmonic
    assert np.abs(Mh[0] - Mh[0].T).max() > 1e-3 * np.abs(Mh[0]).max()   # not symmetric


def test_pass_adds_the_field_source(oracle):
    from radiant_hip import electromagnetic as EM, problems as P
    p = _field(P.small_problem(3, N=3, G=2, L=2, quad_type="carlson", quad_order=4))
    q = P.small_problem(3, N=3, G=2, L=2, quad_type="carlson", quad_order=4)
    nd, np_, nps, S = oracle.sizes(p)
    Mo = oracle.em_matrix(p)
    rng = np.random.default_rng(4)
    z = rng.standard_normal(S)
    qlout = rng.standard_normal(np_ * p.Nc)
    g = 1
    phi = z[:np_ * p.Nc].reshape((np_, p.Nc), order="F")
    extra = (Mo[g] @ phi).reshape(-1, order="F")
    got = oracle.transport_pass(p, g, False, z, qlout)
    want = oracle.transport_pass(q, g, False, z, qlout + extra)
    assert rel_err(got, want) < 1e-13
    wrong = (Mo[g].T @ phi).reshape(-1, order="F")
    assert rel_err(got, oracle.transport_pass(q, g, False, z, qlout + wrong)) > 1e-3


def test_uniform_reflective_medium_is_unchanged(oracle):
    from radiant_hip import problems as P
    for nd in (2, 3):
        kw = dict(N=3, G=2, L=1, Nmat=1, tol=1e-13, bc=[1] * (2 * nd), seed=3, quad_type="gauss-legendre-chebychev",
                  quad_order=4)
        p = P.small_problem(nd, **kw)
        p.qvol[...] = 0.0
        p.qvol[0, :] = 1.0
        ref, _ = oracle.solve(p)
        q = _field(P.small_problem(nd, **kw), B=B_SOLVE)
        q.qvol = p.qvol
        phi, st = oracle.solve(q)
        assert all(s["converged"] for s in st)
        assert rel_err(phi, ref) < 1e-11


# ------------------------------------------------------------------------------------- GPU
GPU_CASES = [(3, "carlson", 4, "BTE"), (2, "gauss-legendre-chebychev", 4, "BTE"), (3, "lebedev", 5, "BFP")]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=[f"{c[0]}d-{c[1]}-{c[3]}" for c in GPU_CASES])
def test_gpu_field_solve(oracle, case):
    from radiant_hip import problems as P, solver as S
    nd, qt, N, stype = case
    p = _field(P.small_problem(nd, N=5, G=3, L=2, tol=1e-12, seed=33, quad_type=qt, quad_order=N),
               B=B_SOLVE)
    if stype == "BFP":
        rng = np.random.default_rng(8)
        p.csd, p.Sb, p.dE = True, 0.7 * rng.uniform(0.5, 2.0, (p.G + 1, p.Nmat)), rng.uniform(0.5, 1.5, p.G)
        p.solver_type, p.T, p.fokker_planck = "BFP", 0.3 * rng.uniform(0.2, 1.0, (p.G, p.Nmat)), "galerkin"
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    for state in ("faithful", "reduced"):
        phi, st = S.compute_flux(p, state=state)
        e = rel_err(phi, ref)
        print("field", case, state, e, flush=True)
        sys.stdout.flush()
        assert e < 1e-10 and all(s["converged"] for s in st)
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    print("field mixed-fp32", case, e, flush=True)
    assert e < 1e-6


@pytest.mark.gpu
def test_gpu_field_in_batch(oracle):
    """The lockstep solver does not take a problem with a field; a batch that holds one solves its
    problems on their own streams, each within 1e-6 of the oracle in mixed FP32."""
    from radiant_hip import problems as P, solver as S
    p = _field(P.small_problem(3, N=4, G=2, L=1, tol=1e-12, quad_type="carlson", quad_order=4), B=B_SOLVE)
    ref, _ = oracle.solve(p)
    phis, _, _ = S.solve_batch([p, p], precision="mixed-fp32")
    for ph in phis:
        e = rel_err(ph, ref)
        print("field batch", e, flush=True)
        assert e < 1e-6


@pytest.mark.gpu
def test_gpu_field_served(oracle):
    """A request with a field through the server (mixed FP32): the lockstep engine does not take it,
    so it runs on its own stream; within 1e-6 of the oracle."""
    from radiant_hip import problems as P, solver as S
    p = _field(P.small_problem(2, N=6, G=2, L=1, tol=1e-12, quad_type="gauss-legendre-chebychev", quad_order=4),
               B=B_SOLVE)
    ref, _ = oracle.solve(p)
    e = rel_err(S.serve_alone(p, "mixed-fp32"), ref)
    print("field served", e, flush=True)
    assert e < 1e-6
