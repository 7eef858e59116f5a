"""Solver types with the angular Fokker-Planck term and the other removal / scattering variants
(SN solver types BFP, FP, CSD, BFP-EF; sn_flux.jl:69-112, sn_inner_pass.jl:22-24,
fokker_planck_source.jl:26-36, fokker_planck_{galerkin,finite_difference,differential_quadrature}.jl).

CPU pins of the Fokker-Planck operator ℳ and of the oracle:
- the host (numpy) and oracle (C++) restatements of ℳ agree for every method and quadrature;
- closed forms: the Galerkin and differential-quadrature operators are diag(λ₀ - l(l+1)) in
  moment space (the second is exact on polynomials of degree < N); the finite-difference
  operators conserve particles (L 1 = 0) and keep the first-moment eigenvalue (L mu = -2 mu,
  and likewise eta and xi for the product set and for the Voronoi operator on the sphere), with
  non-negative couplings (monotone);
- the Voronoi neighbours of the host (convex hull) and of the oracle (the reference's vertex
  triplets) agree with scipy's spherical Voronoi diagram, and satisfy Euler's formula;
- with the Galerkin operator a BFP solve is a BCSD solve whose in-group scattering moments gain
  T (λ₀ - l(l+1)) and whose removal gains T λ₀;
- one BFP pass equals one BCSD pass with T ℳ Φ added to the source (non-symmetric ℳ, so a
  transposed operator or a wrong material/group index of T fails);
- an infinite medium keeps the slowing-down moment recursion of the BCSD test: the operator
  conserves particles;
- FP, CSD and BFP-EF reduce to BFP / BCSD solves with the removal and scattering they use.
GPU: solves against the oracle (FP64 1e-10; mixed FP32 1e-6)."""
import copy
import sys

import numpy as np
import pytest

from util import rel_err


def _mods():
    from radiant_hip import angular, fokker_planck, problems
    return angular, fokker_planck, problems


def _bfp(p, stype="BFP", method="finite-difference", seed=11, tscale=0.5):
    rng = np.random.default_rng(seed)
    p.csd = True
    p.Sb = 0.7 * rng.uniform(0.5, 2.0, (p.G + 1, p.Nmat))
    p.dE = rng.uniform(0.5, 1.5, p.G)
    p.solver_type = stype
    p.T = tscale * rng.uniform(0.2, 1.0, (p.G, p.Nmat))
    p.siga = 0.3 * np.asarray(p.sigt)
    p.fokker_planck = method
    return p


def _host_operator(p):
    angular, fpm, _ = _mods()
    qd = angular.quadrature_dimension(p.quad_type, p.ndims, p.qdims)
    ang = angular.angular_set(p.quad_type, p.quad_order, p.ndims, qd, p.legendre_order, p.sn_type)
    Mp, lam = fpm.moment_operator(ang, p.fokker_planck, p.quad_type, p.quad_order, qd)
    Mn, lam_n = fpm.direction_operator(ang, p.fokker_planck, p.quad_type, p.quad_order, qd)
    assert lam == lam_n
    return ang, Mp, Mn, lam


# (ndims, quadrature, order, qdims, method)
OPS = [(1, "gauss-legendre", 8, 1, "finite-difference"), (1, "gauss-legendre", 8, 1, "galerkin"),
       (1, "gauss-legendre", 8, 1, "differential-quadrature"), (1, "gauss-lobatto", 6, 1, "finite-difference"),
       (2, "gauss-legendre-chebychev", 4, 2, "finite-difference"), (2, "gauss-legendre-chebychev", 4, 3, "finite-difference"),
       (3, "gauss-legendre-chebychev", 4, 3, "finite-difference"), (3, "gauss-legendre-chebychev", 4, 3, "galerkin"),
       (3, "lebedev", 5, 3, "galerkin"),
       # the Voronoi construction (tools/voronoi.jl) of the sets other than the product set
       (3, "lebedev", 5, 3, "finite-difference"), (3, "carlson", 6, 3, "finite-difference"),
       (2, "lebedev", 7, 2, "finite-difference"), (2, "carlson", 4, 2, "finite-difference")]


def _op_problem(nd, qt, N, qd, meth):
    *_, P = _mods()
    p = P.small_problem(nd, N=3, G=2, L=2, quad_type=qt, quad_order=N)
    p.qdims = qd if qt not in ("gauss-legendre", "gauss-lobatto") else 1
    if p.qdims == 2:
        p.sn_type = "standard"  # the folded product set has no half-range Galerkin basis
    return _bfp(p, method=meth)


@pytest.mark.parametrize("case", range(len(OPS)))
def test_operator_host_matches_oracle(oracle, case):
    p = _op_problem(*OPS[case])
    ang, Mp, _, lam = _host_operator(p)
    Mo, lo = oracle.fokker_planck(p)
    assert Mo.shape == Mp.shape
    assert abs(lo - lam) <= 1e-12 * lam
    assert np.abs(Mo - Mp).max() <= 1e-11 * np.abs(Mp).max()


@pytest.mark.parametrize("case", [i for i, c in enumerate(OPS) if c[4] != "finite-difference"])
def test_galerkin_and_dq_are_diagonal_in_moments(case):
    p = _op_problem(*OPS[case])
    ang, Mp, _, lam = _host_operator(p)
    want = lam - ang.pl * (ang.pl + 1.0)
    assert np.allclose(np.diag(Mp), want, rtol=0, atol=1e-9 * lam)
    assert np.abs(Mp - np.diag(np.diag(Mp))).max() <= 1e-9 * lam


@pytest.mark.parametrize("case", [i for i, c in enumerate(OPS) if c[4] == "finite-difference"])
def test_finite_difference_preserves_low_moments(case):
    nd, qt, N, qd, _ = OPS[case]
    p = _op_problem(*OPS[case])
    ang, Mp, Mn, lam = _host_operator(p)
    L = Mn - lam * np.eye(len(Mn))
    one = np.ones(len(Mn))
    assert np.abs(L @ one).max() <= 1e-11 * lam                    # conservation
    cosines = [np.asarray(ang.mu)]
    if qt == "gauss-legendre-chebychev":
        cosines += [np.asarray(ang.eta)] + ([np.asarray(ang.xi)] if qd == 3 else [])
    elif qd == 3:   # Voronoi weights on the sphere: all three first moments
        cosines += [np.asarray(ang.eta), np.asarray(ang.xi)]
    elif qd == 2:   # on the hemisphere the least-squares weights cannot keep them (voronoi.jl:267-353)
        cosines = []
    for v in cosines:                                               # first-moment eigenvalue -2
        assert np.abs(L @ v + 2 * v).max() <= 1e-10 * lam
    off = Mn - np.diag(np.diag(Mn))
    assert off.min() >= -1e-12 * lam and np.diag(Mn).min() >= -1e-12 * lam  # monotone
    # in moments (Galerkin M and D): l = 0 and l = 1 are eigen-moments with λ₀ and λ₀ - 2 where the
    # l = 1 columns of M are the harmonics at the nodes (not for the carlson set, whose M = D^-1
    # mixes them: the direction-space relation above is the one its operator keeps)
    for p_ in np.nonzero(ang.pl <= 1)[0] if p.sn_type == "galerkin-d" and qt != "carlson" else []:
        col = np.zeros(len(Mp))
        col[p_] = lam - 2.0 * ang.pl[p_]
        assert np.abs(Mp[:, p_] - col).max() <= 1e-9 * lam


def test_galerkin_solve_is_modified_scattering(oracle):
    angular, fpm, P = _mods()
    for nd in (1, 3):
        p = P.small_problem(nd, N=6 if nd == 1 else 3, G=3, L=2, tol=1e-13, seed=61,
                            quad_type="gauss-legendre-chebychev" if nd == 3 else None, quad_order=None if nd == 1 else 2)
        p = _bfp(p, method="galerkin", seed=12)
        ang, Mp, _, lam = _host_operator(p)
        got, st = oracle.solve(p)
        q = copy.deepcopy(p)
        q.solver_type, q.T = "BCSD", None
        Ls = int(ang.pl.max())
        sigs = np.zeros((p.Nmat, p.G, p.G, Ls + 1))
        sigs[..., :p.sigs.shape[3]] = p.sigs
        for m in range(p.Nmat):
            for g in range(p.G):
                sigs[m, g, g, :] += p.T[g, m] * (lam - np.arange(Ls + 1) * (np.arange(Ls + 1) + 1.0))
        q.sigs, q.Lxs = sigs, Ls
        q.sigt = p.sigt + p.T * lam
        ref, _ = oracle.solve(q)
        assert all(s["converged"] for s in st)
        assert rel_err(got, ref) < 1e-10, nd


def test_pass_adds_fokker_planck_source(oracle):
    angular, fpm, P = _mods()
    p = _bfp(P.small_problem(2, N=3, G=2, L=2, seed=62), seed=13)
    ang, Mp, _, lam = _host_operator(p)
    assert np.abs(Mp - Mp.T).max() > 1e-3 * np.abs(Mp).max()   # a transposed ℳ would differ
    nd_, Np, Nps, S = oracle.sizes(p)
    rng = np.random.default_rng(5)
    z = rng.standard_normal(S)
    qlout = rng.standard_normal(Np * p.Nc)
    q = copy.deepcopy(p)
    q.solver_type, q.T = "BCSD", None
    q.sigt = p.sigt + p.T * lam
    mat = np.asarray(p.mat).reshape(-1, order="F") - 1
    phi = z[:Np * p.Nc].reshape((Np, p.Nc), order="F")
    for g in range(p.G):
        extra = ((Mp @ phi) * p.T[g, mat][None, :]).reshape(-1, order="F")
        got = oracle.transport_pass(p, g, False, z, qlout)
        want = oracle.transport_pass(q, g, False, z, qlout + extra)
        assert rel_err(got, want) < 1e-11, g
        # A z of BFP = A z of BCSD plus the sweep of T ℳ Φ alone
        got = oracle.transport_pass(p, g, True, z, qlout)
        want = oracle.transport_pass(q, g, True, z, qlout) + oracle.transport_pass(q, g, False, np.zeros(S), extra)
        assert rel_err(got, want) < 1e-11, g


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_infinite_medium_conserves(oracle, nd):
    *_, P = _mods()
    G = 3
    p = P.small_problem(nd, N=3, G=G, L=2, Nmat=1, tol=1e-13, seed=63, bc=[1] * (2 * nd),
                        quad_type="gauss-legendre-chebychev" if nd > 1 else None, quad_order=None if nd == 1 else 2)
    q = np.zeros_like(p.qvol)
    q[0, :] = 1.0
    q[1, :] = 0.4
    p.qvol = q
    _bfp(p, seed=14)
    phi, st = oracle.solve(p)
    St, Ss, Sb, dE = p.sigt[:, 0], p.sigs[0, :, :, 0], p.Sb[:, 0], p.dE
    ref = np.zeros(G)
    E = 0.0
    for g in range(G):
        Sm, Sp = Sb[g] / dE[g], Sb[g + 1] / dE[g]
        src = q[g, 0] + sum(Ss[gi, g] * ref[gi] for gi in range(g)) + (Sp + Sm) * E
        ref[g] = src / (St[g] + 2 * Sp - Ss[g, g])
        E = (2 * ref[g] - E) * (dE[g + 1] / dE[g] if g + 1 < G else 1.0)
    for g in range(G):
        assert np.allclose(phi[g, 0, :], ref[g], rtol=1e-9, atol=0), (g, phi[g, 0, :2], ref[g])
    assert np.abs(phi[:, 1:, :]).max() < 1e-8 * ref.max()
    assert all(s["converged"] for s in st)


def test_types_reduce_to_bfp_and_bcsd(oracle):
    *_, P = _mods()
    base = _bfp(P.small_problem(2, N=4, G=3, L=2, tol=1e-13, seed=64), seed=15)
    # FP = BFP with the absorption as removal and no Boltzmann scattering
    fp = copy.deepcopy(base)
    fp.solver_type = "FP"
    ref = copy.deepcopy(base)
    ref.sigt, ref.sigs = base.siga, np.zeros_like(base.sigs)
    assert rel_err(oracle.solve(fp)[0], oracle.solve(ref)[0]) < 1e-12
    # CSD = BCSD likewise
    csd = copy.deepcopy(base)
    csd.solver_type = "CSD"
    ref2 = copy.deepcopy(ref)
    ref2.solver_type = "BCSD"
    assert rel_err(oracle.solve(csd)[0], oracle.solve(ref2)[0]) < 1e-12
    # BFP-EF = BCSD without in-group scattering and its l = 0 moment taken off the removal
    ef = copy.deepcopy(base)
    ef.solver_type = "BFP-EF"
    ref3 = copy.deepcopy(base)
    ref3.solver_type = "BCSD"
    s = np.array(base.sigs)
    t = np.array(base.sigt)
    for m in range(base.Nmat):
        t[:, m] -= np.diag(s[m, :, :, 0])
        for g in range(base.G):
            s[m, g, g, :] = 0.0
    ref3.sigs, ref3.sigt = s, t
    assert rel_err(oracle.solve(ef)[0], oracle.solve(ref3)[0]) < 1e-12
    # and they all differ from one another
    assert rel_err(oracle.solve(fp)[0], oracle.solve(base)[0]) > 1e-3


def test_inconsistent_types_refused(oracle):
    _, fpm, P = _mods()
    p = P.small_problem(1, N=4, G=2)
    p.solver_type = "BFP"   # without the slowing-down data
    with pytest.raises(RuntimeError):
        oracle.solve(p)
    with pytest.raises(ValueError):
        fpm.prepare(p, None, 1)



@pytest.mark.parametrize("qt,N", [("lebedev", 3), ("lebedev", 5), ("lebedev", 9), ("carlson", 4), ("carlson", 8)])
def test_voronoi_neighbours(oracle, qt, N):
    """The Voronoi neighbours of the host (convex hull, coplanar facets merged) and of the oracle
    (vertex triplets by Newton, as voronoi.jl) agree with scipy's spherical Voronoi diagram: two
    nodes are neighbours when their regions share two vertices (vertices closer than 1e-7 merged;
    on the octahedral sets four nodes meet at one vertex, and the diagonal pairs share only it)."""
    from scipy.spatial import SphericalVoronoi
    angular, fpm, _ = _mods()
    mu, eta, xi, w = angular.quadrature(N, qt, 3, 3)
    Pn = np.stack([mu, eta, xi], 1)
    sv = SphericalVoronoi(Pn, radius=1.0, center=np.zeros(3))
    rep = np.arange(len(sv.vertices))
    for i in range(len(sv.vertices)):
        close = np.nonzero(np.abs(sv.vertices[:i] - sv.vertices[i]).max(axis=1) < 1e-7)[0]
        if len(close):
            rep[i] = rep[close[0]]
    regions = [set(rep[r]) for r in sv.regions]
    want = sorted((n, m) for n in range(len(Pn)) for m in range(n + 1, len(Pn)) if len(regions[n] & regions[m]) >= 2)
    assert fpm.voronoi_edges(Pn) == want
    nb = oracle.voronoi_neighbours(Pn)
    assert sorted({(min(n, m), max(n, m)) for n in range(len(nb)) for m in nb[n]}) == want
    assert all(len(set(v)) == len(v) for v in nb)   # one edge per neighbour
    # Euler on the sphere: V - E + F = 2 with F the nodes; the Voronoi vertices V = 2 + E - F
    assert len(set(rep)) == 2 + len(want) - len(Pn)


# ------------------------------------------------------------------------------------------
# GPU parity
# ------------------------------------------------------------------------------------------

def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (ndims, solver type, method, quadrature, order, qdims, bc)
GPU_CASES = [(1, "BFP", "finite-difference", None, None, 1, None),
             (1, "BFP", "differential-quadrature", None, None, 1, [1, 2]),
             (1, "FP", "galerkin", None, None, 1, None),
             (2, "BFP", "finite-difference", "gauss-legendre-chebychev", 4, 3, None),
             (2, "BFP", "finite-difference", "gauss-legendre-chebychev", 4, 3, [1, 0, 2, 2]),
             (2, "CSD", "finite-difference", None, None, 0, None),
             (2, "BFP-EF", "finite-difference", None, None, 0, None),
             (3, "BFP", "finite-difference", "gauss-legendre-chebychev", 2, 3, None),
             (3, "FP", "galerkin", "lebedev", 5, 3, None),
             (3, "BFP", "finite-difference", "carlson", 4, 3, None),
             (3, "BFP", "finite-difference", "lebedev", 5, 3, [1, 0, 0, 0, 2, 2]),
             (2, "BFP", "finite-difference", "lebedev", 7, 2, None)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_fp_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    *_, P = _mods()
    nd, stype, meth, qt, qo, qd, bc = GPU_CASES[case]
    p = P.small_problem(nd, N=5 if nd == 3 else 7, G=3, L=2, tol=1e-12, seed=70 + case, bc=bc,
                        quad_type=qt, quad_order=qo)
    if qd:
        p.qdims = qd
    if qd == 2 and qt != "gauss-legendre-chebychev":
        p.sn_type = "standard"  # the folded set has no half-range Galerkin basis
    # on the folded sets with the standard operators the solve converges for a small momentum
    # transfer only (with T ~ 0.5, 1000 GMRES iterations in the oracle leave groups unconverged)
    _bfp(p, stype=stype, method=meth, seed=20 + case, tscale=0.05 if qd == 2 else 0.5)
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    phi, st = S.compute_flux(p)
    e = rel_err(phi, ref)
    _say("fp", nd, stype, meth, qt, bc, e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    if bc is None:
        mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
        e = rel_err(mx, ref)
        _say("fp mixed-fp32", nd, stype, e)
        assert e < 1e-6


@pytest.mark.gpu
def test_fp_gpu_other_accelerations_and_batch(oracle):
    from radiant_hip import solver as S
    *_, P = _mods()
    p = _bfp(P.small_problem(2, N=6, G=3, L=2, tol=1e-12, seed=90), seed=31)
    ref, _ = oracle.solve(p)
    for acc in ("anderson", "bicgstab"):
        q = copy.deepcopy(p)
        q.accel = acc
        r, _ = oracle.solve(q)
        phi, _ = S.compute_flux(q)
        e = rel_err(phi, r)
        _say("fp", acc, e)
        assert e < 1e-10, acc
    phis, _, _ = S.solve_batch([p, p], precision="mixed-fp32")
    for ph in phis:
        assert rel_err(ph, ref) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_fp_in_lockstep_batch(oracle, precision):
    """Problems with the angular Fokker-Planck term in the lockstep solver (their source
    qlout + T ℳ Φ formed at the start of each round), mixed with each other, with BTE problems and
    with reflective/periodic faces: within 1e-6 of the oracle, bit-identical alone and in the mix."""
    import os
    from radiant_hip import solver as S
    *_, P = _mods()
    probs = []
    for k, (nd, stype, meth, qt, qo, qd, bc) in enumerate(GPU_CASES[:7]):
        p = P.small_problem(nd, N=5 if nd == 3 else 7, G=3, L=2, tol=1e-12, seed=170 + k, bc=bc,
                            quad_type=qt, quad_order=qo)
        if qd:
            p.qdims = qd
        probs.append(_bfp(p, stype=stype, method=meth, seed=120 + k))
    probs.append(P.small_problem(2, N=8, G=3, L=2, tol=1e-12, seed=180))
    phis, stats, _ = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = oracle.solve(p)
        e = rel_err(phis[i], ref)
        _say("fp batch", precision, i, e)
        assert e < 1e-6 and all(s["converged"] for s in stats[i])
    alone, _, _ = S.solve_batch([probs[3]], precision=precision)
    assert np.array_equal(alone[0], phis[3])
    os.environ["RADIANT_HIP_PROFILE"] = "1"
    try:  # kernel times are recorded by lockstep solvers only
        rb = S.ResidentBatch(probs[:3], precision=precision)
        rb.solve()
        _, n, _, _, _ = rb.kernel_times()
        rb.close()
    finally:
        os.environ.pop("RADIANT_HIP_PROFILE", None)
    assert n[:, 1].sum() > 0
