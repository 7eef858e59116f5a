"""Pins of the oracle's reflective / periodic boundary branches (sn_sweep_{1,2,3}D.jl boundary
blocks, e.g. sn_sweep_2D.jl:70-99; codes from Geometry.jl:604-632: 0 void, 1 reflective,
2 periodic), by what the mathematics fixes:

- a uniform source in one material with every face reflective, or every face periodic, is an
  infinite medium: the scalar flux is the closed form phi_g = (q_g + sum_{g'<g} Ss0(g'->g) phi_g')
  / (St_g - Ss0(g->g)) in every cell and every other moment vanishes (D M = I for Galerkin-D);
- a reflective x- face is a mirror: the problem equals the right half of the void problem on the
  doubled, mirrored grid (quadratures symmetric in mu);
- periodic x faces make the solution invariant under a cyclic shift of the source and materials.
"""
import copy

import numpy as np
import pytest

from util import rel_err


def _P():
    from radiant_hip import problems as P
    return P


def _uniform(ndims, bc, G=3, L=2, N=4, seed=5):
    P = _P()
    p = P.small_problem(ndims, N=N, G=G, L=L, Nmat=1, tol=1e-13, seed=seed, bc=[bc] * (2 * ndims))
    q = np.zeros_like(p.qvol)
    q[0, :] = 1.0
    q[1, :] = 0.25
    p.qvol = q
    return p


def _infinite_medium(p):
    G = p.G
    st = p.sigt[:, 0]
    ss = p.sigs[0, :, :, 0]  # [from, to]
    q = p.qvol[:, 0]
    phi = np.zeros(G)
    for g in range(G):
        src = q[g] + sum(ss[gi, g] * phi[gi] for gi in range(g))
        phi[g] = src / (st[g] - ss[g, g])
    return phi


@pytest.mark.parametrize("ndims", [1, 2, 3])
@pytest.mark.parametrize("bc", [1, 2])
def test_infinite_medium(oracle, ndims, bc):
    p = _uniform(ndims, bc)
    phi, st = oracle.solve(p)
    ref = _infinite_medium(p)
    for g in range(p.G):
        assert np.allclose(phi[g, 0, :], ref[g], rtol=1e-9, atol=0), (g, phi[g, 0, :3], ref[g])
    assert np.abs(phi[:, 1:, :]).max() < 1e-9 * np.abs(ref).max()
    assert all(s["converged"] for s in st)


def _mirror_pair(ndims, N=6, seed=9):
    """(reflective-at-x- problem on N cells in x, void problem on 2N cells mirrored about x=0)."""
    P = _P()
    dims = (N, 3, 3)[:ndims]
    p = P.small_problem(ndims, dims=dims, G=2, L=2, tol=1e-13, seed=seed,
                        bc=[1] + [0] * (2 * ndims - 1))
    q = np.zeros_like(p.qvol)
    rng = np.random.default_rng(seed)
    q[0, :] = rng.uniform(0.0, 1.0, p.Nc)
    p.qvol = q
    d = copy.deepcopy(p)
    d.Nx = 2 * N
    d.dx = np.concatenate([p.dx[::-1], p.dx])
    shp = (p.Nx, p.Ny, p.Nz)
    m = p.mat.reshape(shp, order="F")
    d.mat = np.concatenate([m[::-1], m], axis=0)
    qq = q.reshape((p.G,) + shp, order="F")
    d.qvol = np.concatenate([qq[:, ::-1], qq], axis=1).reshape(p.G, -1, order="F")
    d.bc = np.zeros(2 * ndims, np.int32)
    return p, d


@pytest.mark.parametrize("ndims", [1, 2, 3])
def test_reflective_is_a_mirror(oracle, ndims):
    p, d = _mirror_pair(ndims)
    phi, _ = oracle.solve(p)
    phd, _ = oracle.solve(d)
    G, Np = phi.shape[:2]
    half = phd.reshape((G, Np, 2 * p.Nx, p.Ny, p.Nz), order="F")[:, :, p.Nx:].reshape(G, Np, -1, order="F")
    assert rel_err(phi, half) < 1e-10
    v = copy.deepcopy(p)
    v.bc = np.zeros(2 * ndims, np.int32)
    phv, _ = oracle.solve(v)
    assert rel_err(phv, phi) > 1e-3  # the reflective face is active


@pytest.mark.parametrize("ndims", [1, 2])
def test_periodic_shift_invariance(oracle, ndims):
    P = _P()
    N, k = 8, 3
    p = P.small_problem(ndims, N=N, G=2, L=2, tol=1e-13, seed=13, bc=[2, 2] + [0] * (2 * ndims - 2))
    shp = (p.Nx, p.Ny, p.Nz)
    s = copy.deepcopy(p)
    s.mat = np.roll(p.mat.reshape(shp, order="F"), k, axis=0)
    s.qvol = np.roll(p.qvol.reshape((p.G,) + shp, order="F"), k, axis=1).reshape(p.G, -1, order="F")
    phi, _ = oracle.solve(p)
    phs, _ = oracle.solve(s)
    G, Np = phi.shape[:2]
    back = np.roll(phs.reshape((G, Np) + shp, order="F"), -k, axis=2).reshape(G, Np, -1, order="F")
    assert rel_err(phi, back) < 1e-10
    # and periodic differs from reflective (the faces are not interchangeable)
    r = copy.deepcopy(p)
    r.bc = np.array([1, 1] + [0] * (2 * ndims - 2), np.int32)
    phr, _ = oracle.solve(r)
    assert rel_err(phr, phi) > 1e-3
