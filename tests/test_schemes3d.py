"""Spatial closures of the 3D cell: DD and DG of orders up to 3 per axis, with and without full
coupling (flux_3D_BTE of sn_3D_bte.jl:34-152, sn_sweep_3D.jl:57-240); the GPU sweep holds cells of
at most 12 moments (3 x 2 x 2 with full coupling, any orders up to 3 without).

CPU pins of the oracle's cell, as for 2D (psi = sum C_a C_b C_c psi_abc P_a P_b P_c on the cell):
- polynomial transport solutions are reproduced exactly, cell and face moments, by DD up to degrees
  (Ox, Oy, Oz) and by DG up to (Ox - 1, Oy - 1, Oz - 1) (cross terms with full coupling, single-axis
  terms without), and not by DG beyond;
- the zeroth moments obey the cell balance.
Solve-level: an all-reflective infinite medium has no spatial slopes.
GPU: passes (1e-12) and solves (1e-10 FP64, 1e-6 mixed FP32) against the oracle."""
import copy
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as Lg

from util import rel_err


def _face(c, Cs, O3, ax, side, fc):
    v = Lg.legval(side, np.moveaxis(c, ax, 0))
    o = [k for k in range(3) if k != ax]
    v = v / (Cs[o[0]][:, None] * Cs[o[1]][None, :])
    n1, n2 = O3[o[0]], O3[o[1]]
    if fc:
        return np.array([v[p, q] for q in range(n2) for p in range(n1)])
    return np.array([v[p, 0] for p in range(n1)] + [v[0, q] for q in range(1, n2)])


def _cell_case(oracle, schemes, O3, fc, deg, om, rng):
    St, h = 0.9, np.array([0.37, 0.51, 0.43])
    A = [max(O3[a], deg[a] + 1) for a in range(3)]
    c = np.zeros(A)
    for a in range(deg[0] + 1):
        for b in range(deg[1] + 1):
            for d in range(deg[2] + 1):
                if fc or (a > 0) + (b > 0) + (d > 0) <= 1:
                    c[a, b, d] = rng.standard_normal()
    Qc = St * c.copy()
    for ax in range(3):
        dc = Lg.legder(c, axis=ax)
        sl = [slice(None)] * 3
        sl[ax] = slice(0, dc.shape[ax])
        Qc[tuple(sl)] += om[ax] * 2 / h[ax] * dc
    Cs = [np.sqrt(2 * np.arange(A[a]) + 1) for a in range(3)]
    norm = Cs[0][:, None, None] * Cs[1][None, :, None] * Cs[2][None, None, :]
    Qm, cm = Qc / norm, c / norm
    ox, oy, oz = O3
    nm = ox * oy * oz if fc else ox + oy + oz - 2
    Qn, want = np.zeros(nm), np.zeros(nm)
    for a in range(ox):
        for b in range(oy):
            for d in range(oz):
                if fc:
                    j = ox * oy * d + ox * b + a
                elif (a > 0) + (b > 0) + (d > 0) <= 1:
                    j = a + b + d + (ox - 1 if b > 0 else 0) + (ox - 1 + oy - 1 if d > 0 else 0)
                else:
                    continue
                Qn[j], want[j] = Qm[a, b, d], cm[a, b, d]
    s = np.sign(om)
    fin = [_face(c, Cs, O3, a, -s[a], fc) for a in range(3)]
    fout = [_face(c, Cs, O3, a, s[a], fc) for a in range(3)]
    psi, fx, fy, fz = oracle.cell_3d(schemes, O3, fc, om, h, St, Qn, *fin)
    err = max(np.abs(psi - want).max(), *(np.abs(f - g).max() for f, g in zip((fx, fy, fz), fout)))
    bal = St * psi[0] + sum(abs(om[a]) / h[a] * ((fx, fy, fz)[a][0] - fin[a][0]) for a in range(3)) - Qn[0]
    return err, abs(bal)


@pytest.mark.parametrize("scheme", ["DD", "DG"])
@pytest.mark.parametrize("O3", [(1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1), (1, 2, 2), (2, 2, 2), (3, 1, 1), (1, 3, 2),
                                (3, 2, 2), (3, 3, 3)])
@pytest.mark.parametrize("fc", [True, False])
def test_cell_polynomial_exactness_and_balance(oracle, scheme, O3, fc):
    rng = np.random.default_rng(sum(O3))
    top = O3 if scheme == "DD" else tuple(o - 1 for o in O3)
    for om in [(0.6, 0.3, 0.5), (-0.5, 0.7, -0.2), (0.2, -0.8, 0.4)]:
        for deg in [tuple(o - 1 for o in O3), O3]:
            err, bal = _cell_case(oracle, (scheme,) * 3, O3, fc, deg, np.array(om), rng)
            assert bal < 1e-13
            if all(dg <= t for dg, t in zip(deg, top)):
                assert err < 1e-13, (scheme, O3, fc, deg, err)
            else:
                assert err > 1e-6, (scheme, O3, fc, deg, err)


def _set(p, sx, sy, sz, ox, oy, oz, fc):
    p.scheme, p.scheme_y, p.scheme_z = sx, sy, sz
    p.scheme_order, p.scheme_order_y, p.scheme_order_z, p.full_coupling = ox, oy, oz, fc
    return p


CONFIGS = [("DD", "DD", "DD", 2, 2, 2, True), ("DG", "DG", "DG", 2, 2, 2, False), ("DD", "DG", "DD", 2, 1, 2, True),
           ("DG", "DD", "DG", 1, 2, 1, True)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_infinite_medium_has_no_slopes(oracle, cfg):
    from radiant_hip import problems as P
    base = P.small_problem(3, N=3, G=2, L=2, Nmat=1, tol=1e-13, seed=42, bc=[1] * 6)
    base.qvol = np.ones_like(base.qvol)
    ref, _ = oracle.solve(base)
    p = _set(copy.deepcopy(base), *cfg)
    phi, st = oracle.solve(p)
    K = oracle.spatial_moments(p)
    assert all(s["converged"] for s in st)
    assert rel_err(phi[:, :, 0::K], ref) < 1e-10
    assert np.abs(np.delete(phi, np.arange(0, phi.shape[2], K), axis=2)).max() < 1e-10 * np.abs(ref).max()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


GPU_CASES = [(cfg, None) for cfg in CONFIGS] + [(("DD", "DD", "DD", 2, 2, 2, True), [1, 2, 0, 1, 2, 0]),
                                                (("DG", "DG", "DG", 2, 2, 1, False), [1] * 6),
                                                # order 3: 7 moments without coupling, 12 and 6 with it
                                                (("DD", "DD", "DD", 3, 3, 3, False), None),
                                                (("DG", "DD", "DG", 3, 2, 2, True), [1, 2, 0, 1, 2, 0]),
                                                (("DD", "DG", "DD", 1, 3, 2, True), [2] * 6),
                                                (("DG", "DG", "DD", 3, 3, 2, False), [1] * 6)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_scheme3d_gpu_parity(oracle, case):
    from radiant_hip import problems as P, solver as S
    cfg, bc = GPU_CASES[case]
    p = _set(P.small_problem(3, dims=(4, 3, 5), G=2, L=2, tol=1e-12, seed=110 + case, bc=bc), *cfg)
    _say("scheme3d", cfg, bc)
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(500 + case)
        z = rng.standard_normal(Sl)
        q = rng.standard_normal(np_ * p.Nc * K)
        q.reshape(-1, np_)[:, [i for i in range(np_) if h.d.ang.pl[i] > p.Lxs]] = 0.0
        for mode in (0, 1, 2):
            want = oracle.transport_pass(p, 1, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                want = z - want
            e = rel_err(h.transport_pass(1, mode, q, z), want)
            _say("  pass mode", mode, e)
            assert e < 1e-12
    finally:
        h.close()
    phi, st = S.compute_flux(p)
    e = rel_err(phi, ref)
    _say("  solve", e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("  mixed-fp32", e)
    assert e < 1e-6


@pytest.mark.gpu
def test_scheme3d_moment_limit_refused():
    """Full coupling of orders 4 x 4 x 5 gives 80 moments per cell, above the 64 of the general
    walker (orders 3 x 3 x 2, 18 moments, run there): refused at set-up with the limits named."""
    from radiant_hip import problems as P, solver as S
    from radiant_hip._lib import RTError
    p = _set(P.small_problem(3, dims=(4, 3, 5), G=2, L=2, seed=140), "DD", "DD", "DD", 4, 4, 5, True)
    with pytest.raises(RTError, match="at most 64 moments per cell"):
        S.ResidentProblem(p)


@pytest.mark.gpu
def test_scheme3d_faces_in_device_memory(oracle):
    """A grid whose face moments exceed the 160 KiB LDS of a workgroup (72 x 72 x 2 at 2x2x2 with
    full coupling: 175 KiB in FP64) keeps them in device memory; passes, FP64 and mixed solves
    match the oracle as in LDS."""
    import os
    from radiant_hip import problems as P, solver as S
    p = _set(P.small_problem(3, dims=(72, 72, 2), G=2, L=1, tol=1e-12, seed=160, bc=[1, 0, 0, 2, 0, 0]),
             "DD", "DG", "DD", 2, 2, 2, True)
    p.qvol = np.zeros_like(p.qvol)
    p.qvol[0, : p.Nx * 30] = 1.0   # a source in the first 30 rows of both layers
    faces = 8 * (72 * 2 * 4 + 72 * 2 * 4 + 72 * 72 * 4)
    assert faces > 160 * 1024
    nth = max(1, min(16, os.cpu_count() or 1))
    ref, rst = oracle.solve(p, nthreads=nth)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        rng = np.random.default_rng(161)
        z = rng.standard_normal(Sl)
        q = rng.standard_normal(np_ * p.Nc * K)
        q.reshape(-1, np_)[:, [i for i in range(np_) if h.d.ang.pl[i] > p.Lxs]] = 0.0
        want = oracle.transport_pass(p, 1, 0, z, q, nthreads=nth)
        e = rel_err(h.transport_pass(1, 0, q, z), want)
        _say("faces in HBM: pass", e)
        assert e < 1e-12
    finally:
        h.close()
    phi, st = S.compute_flux(p)
    e = rel_err(phi, ref)
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    em = rel_err(mx, ref)
    _say("faces in HBM: solve", e, "mixed-fp32", em)
    assert e < 1e-10 and em < 1e-6 and all(s["converged"] for s in st)
