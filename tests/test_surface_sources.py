"""Surface sources (particle_transport/surface_source.jl:22-408, structures/Source.jl:152-190, the
sweeps' boundary blocks, e.g. sn_sweep_2D.jl:70-99).

CPU pins of the oracle and of the host moments:
- a face source is the same inflow as a reflective face whose boundary state holds the same
  moments: T(0) with the source equals A z with that state (the reflective branch is pinned in
  test_oracle_bc.py);
- the moments are I * Y_p(|mu_s|, phi_s) (times the cell's face area in 2D/3D) on the cells whose
  centres lie in the source's bounds: 1D against scipy's Legendre polynomials, 2D/3D against the
  oracle's half-range harmonics (pinned by orthonormality);
- superposition and scaling; a group with only a surface source is solved (not skipped).
GPU: passes and solves against the oracle."""
import copy
import math
import sys

import numpy as np
import pytest
from scipy import special

from util import rel_err


def _P():
    from radiant_hip import problems
    return problems


def _SRC():
    from radiant_hip import sources
    return sources


def _beam(nd):
    return (0.8, 0.5, math.sqrt(1 - 0.64 - 0.25)) if nd > 1 else (1.0, 0.0, 0.0)


def _prob(nd, faces=("X-",), bc=None, seed=71, G=2, N=None, group=0, qvol=False):
    P, SRC = _P(), _SRC()
    N = N or (6 if nd == 3 else 8)
    p = P.small_problem(nd, N=N, G=G, L=2, tol=1e-12, seed=seed, bc=bc)
    if not qvol:
        p.qvol = np.zeros_like(p.qvol)
    mu, eta, xi = _beam(nd)
    for f in faces:
        # a beam entering through this face: flip the normal component as needed
        d = [mu, eta, xi]
        ax = "XYZ".index(f[0])
        d[ax] = abs(d[ax]) if f[1] == "-" else -abs(d[ax])
        SRC.add_surface_source(p, f, d, intensity=1.0 + 0.5 * ax, group=group,
                               bounds={"x": (1.0, N - 2.0), "y": (1.0, N - 3.0), "z": (2.0, N - 1.0)})
    return p


@pytest.mark.parametrize("nd,face", [(1, "X-"), (1, "X+"), (2, "X-"), (2, "Y+"), (3, "Z-"), (3, "X+")])
def test_source_equals_reflective_feedback(oracle, nd, face):
    p = _prob(nd, faces=(face,))
    f = "X-X+Y-Y+Z-Z+".index(face) // 2
    ax = f // 2
    nd_, np_, nps, S = oracle.sizes(p)
    nc = p.Nc
    npf = [p.Ny * p.Nz, p.Nx * p.Nz, p.Nx * p.Ny]
    g = 0
    # T(0) with the surface source, void faces
    t0 = oracle.transport_pass(p, g, 0, np.zeros(S), np.zeros(np_ * nc))
    # A z with the same moments as the boundary state of a reflective face, no surface source
    r = copy.deepcopy(p)
    r.ssrc, r.Np_source = None, 0
    r.bc = np.zeros(2 * nd, np.int32)
    r.bc[f] = 1
    z = np.zeros(S)
    off = np_ * nc + sum(2 * nps * npf[a] for a in range(ax))
    Q = p.ssrc[f][g]  # [Np_source, npos]
    side = f % 2
    for pos in range(npf[ax]):
        for q in range(p.Np_source):
            z[off + q + nps * (side + 2 * pos)] = Q[q, pos]
    az = oracle.transport_pass(r, g, 1, z, np.zeros(np_ * nc))
    assert np.abs(t0).max() > 0
    assert rel_err(t0[:np_ * nc], az[:np_ * nc]) < 1e-14
    assert rel_err(t0[np_ * nc:], az[np_ * nc:]) < 1e-14  # exit moments too


@pytest.mark.parametrize("nd,face", [(1, "X-"), (1, "X+"), (2, "X-"), (2, "Y-"), (3, "Y+"), (3, "Z-")])
def test_moments_closed_form(oracle, nd, face):
    from radiant_hip.angular import angular_set, quadrature_dimension
    SRC, P = _SRC(), _P()
    p = P.small_problem(nd, N=6, G=3, L=2, tol=1e-12, seed=5)
    f = "X-X+Y-Y+Z-Z+".index(face) // 2
    ax = f // 2
    mu, eta, xi = _beam(nd)
    d = [mu, eta, xi]
    d[ax] = abs(d[ax]) if face[1] == "-" else -abs(d[ax])
    bounds = {"x": (1.0, 4.0), "y": (0.0, 2.5), "z": (2.0, 6.0)}
    Q, norm = SRC.surface_source_moments(p, face, d, intensity=2.5, group=1, bounds=bounds)
    qd = quadrature_dimension(p.quad_type, nd, p.qdims)
    ang = angular_set(p.quad_type, p.quad_order, nd, qd, p.legendre_order, p.sn_type)
    pl, pm = ang.face_lm[f]
    if qd == 1:
        ref = np.array([2.5 * math.sqrt(2 * l + 1) * special.eval_legendre(l, 2 * abs(d[0]) - 1) for l in pl])
    else:
        L = int(max(pl))
        R = oracle.real_sh(L, abs(d[0]), math.atan2(d[2], d[1]), half_range=True)
        ref = np.array([2.5 * R[l * l + l + m] for l, m in zip(pl, pm)])
    assert Q.shape[0] == 3 and not Q[0].any() and not Q[2].any()
    if nd == 1:
        assert np.allclose(Q[1, :, 0], ref, rtol=1e-13, atol=1e-15)
        return
    # cells whose centres lie in the bounds, with their face areas (all widths are 1 here)
    cen = np.arange(6) + 0.5
    trans = [a for a in "xyz"[:nd] if a != "xyz"[ax]]
    ins = [(cen >= bounds[a][0]) & (cen <= bounds[a][1]) for a in trans]
    mask = ins[0] if len(ins) == 1 else (ins[0][:, None] & ins[1][None, :]).reshape(-1, order="F")
    for pos in range(Q.shape[2]):
        want = ref if mask[pos] else 0 * ref
        assert np.allclose(Q[1, :, pos], want, rtol=1e-13, atol=1e-15)
    assert abs(norm - 2.5 * mask.sum()) < 1e-12


def test_superposition_and_scaling(oracle):
    a = _prob(2, faces=("X-",))
    b = _prob(2, faces=("Y+",))
    ab = _prob(2, faces=("X-", "Y+"))
    pa, _ = oracle.solve(a)
    pb, _ = oracle.solve(b)
    pab, _ = oracle.solve(ab)
    assert rel_err(pab, pa + pb) < 1e-9
    s = copy.deepcopy(a)
    s.ssrc = {k: 3.0 * v for k, v in a.ssrc.items()}
    ps, _ = oracle.solve(s)
    assert rel_err(ps, 3.0 * pa) < 1e-9


def test_group_with_only_a_surface_source_is_solved(oracle):
    p = _prob(2, faces=("X-",), group=1)
    phi, st = oracle.solve(p)
    assert not phi[0].any() and st[0]["passes"] == 0       # no source in group 0: skipped
    assert np.abs(phi[1]).max() > 0 and st[1]["passes"] > 0  # the surface source alone is solved


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(1, ("X-",), None), (1, ("X-", "X+"), [0, 1]), (2, ("X-",), None), (2, ("X-", "Y+"), [0, 0, 1, 0]),
         (3, ("Z-",), None), (3, ("X+", "Y-"), [2, 2, 0, 0, 1, 0])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_surface_source_gpu_parity(oracle, case):
    nd, faces, bc = CASES[case]
    from radiant_hip import solver as S
    p = _prob(nd, faces=faces, bc=bc, qvol=(case % 2 == 1))
    ref, rst = oracle.solve(p)
    phi, st = S.compute_flux(p)
    e = rel_err(phi, ref)
    _say("surface source", nd, faces, bc, e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    # pass level: T(z) on a random state, A z and z - A z
    h = S.ResidentProblem(p)
    try:
        nd_, np_, nps, Sl = oracle.sizes(p)
        rng = np.random.default_rng(case)
        z = rng.standard_normal(Sl)
        q = np.zeros(np_ * p.Nc)
        for mode in (0, 1, 2):
            r = oracle.transport_pass(p, 0, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                r = z - r
            got = h.transport_pass(0, mode, q, z)
            assert rel_err(got, r) < 1e-12, mode
    finally:
        h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_surface_source_reduced_and_mixed(oracle, case):
    nd, faces, bc = CASES[case]
    from radiant_hip import solver as S
    p = _prob(nd, faces=faces, bc=bc, qvol=(case % 2 == 1))
    ref, _ = oracle.solve(p)
    phi, st = S.compute_flux(p, state="reduced")
    e = rel_err(phi, ref)
    _say("surface source reduced fp64", nd, faces, bc, e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    mx, st = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("surface source mixed fp32", nd, faces, bc, e)
    assert e < 1e-6 and all(s["converged"] for s in st)
