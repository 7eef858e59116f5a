"""Spatial closures of the 2D cell: DD and DG of orders up to 3 per axis, with and without full
coupling of the high-order moments (SN.set_scheme, set_is_full_coupling; flux_2D_BTE of
sn_2D_bte.jl:27-116, sn_sweep_2D.jl:54-200, face moments of sn_one_speed.jl:92-106).

CPU pins of the oracle's cell (closed forms, with psi(u, v) = sum_ab C_a C_b psi_ab P_a(u) P_b(v)
on the cell, C_k = sqrt(2k + 1), u, v in [-1, 1]):
- a transport solution that is a polynomial of degrees (dx, dy) in the cell is reproduced exactly
  (cell moments and both outgoing face moments) by DD for (dx, dy) <= (Ox, Oy), by DG and by mixed
  DD/DG for (dx, dy) <= (Ox - 1, Oy - 1), with cross terms under full coupling and without them
  under partial coupling; DG is not exact beyond;
- the zeroth moments obey the cell balance St psi_00 + |mu|/dx (fx_out - fx_in)_0 +
  |eta|/dy (fy_out - fy_in)_0 = Q_00;
- DD of orders (1, 1) is the classic diamond cell.
Solve-level: an all-reflective infinite medium has no spatial slopes, so every closure gives the
first-order flux with vanishing higher moments.
GPU: passes (1e-12) and solves (1e-10 FP64, 1e-6 mixed FP32) against the oracle."""
import copy
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as Lg

from util import rel_err


def _basis(ox, oy, fc):
    """Basis order of the cell moments: (ix, iy) at ox*iy + ix, or ix for iy = 0 and ox-1+iy."""
    if fc:
        return [(ix, iy) for iy in range(oy) for ix in range(ox)]
    return [(ix, 0) for ix in range(ox)] + [(0, iy) for iy in range(1, oy)]


def _cell_case(oracle, schemes, ox, oy, fc, dxd, dyd, mu, eta, rng):
    St, dx, dy = 0.9, 0.37, 0.51
    A, B = max(ox, dxd + 1), max(oy, dyd + 1)
    c = np.zeros((A, B))
    for a in range(dxd + 1):
        for b in range(dyd + 1):
            if fc or a == 0 or b == 0:
                c[a, b] = rng.standard_normal()
    Qc = St * c.copy()
    du, dv = Lg.legder(c, axis=0), Lg.legder(c, axis=1)
    Qc[:du.shape[0], :] += mu * 2 / dx * du
    Qc[:, :dv.shape[1]] += eta * 2 / dy * dv
    Ca, Cb = np.sqrt(2 * np.arange(A) + 1), np.sqrt(2 * np.arange(B) + 1)
    order = _basis(ox, oy, fc)
    Qm, cm = Qc / np.outer(Ca, Cb), c / np.outer(Ca, Cb)
    Qn = np.array([Qm[a, b] for a, b in order])
    sx, sy = np.sign(mu), np.sign(eta)
    fx_in = np.array([Lg.legval(-sx, c[:, b]) for b in range(B)]) / Cb
    fy_in = np.array([Lg.legval(-sy, c[a, :]) for a in range(A)]) / Ca
    psi, fx, fy = oracle.cell_2d(schemes, (ox, oy), fc, mu, eta, St, dx, dy, Qn, fx_in[:oy], fy_in[:ox])
    want = np.array([cm[a, b] for a, b in order])
    fx_out = np.array([Lg.legval(sx, c[:, b]) for b in range(B)]) / Cb
    fy_out = np.array([Lg.legval(sy, c[a, :]) for a in range(A)]) / Ca
    err = max(np.abs(psi - want).max(), np.abs(fx - fx_out[:oy]).max(), np.abs(fy - fy_out[:ox]).max())
    bal = St * psi[0] + abs(mu) / dx * (fx[0] - fx_in[0]) + abs(eta) / dy * (fy[0] - fy_in[0]) - Qn[0]
    return err, abs(bal)


ORDERS = [(1, 1), (2, 1), (1, 2), (2, 2), (3, 2), (2, 3), (3, 3), (4, 1), (1, 4), (4, 3), (4, 4)]


@pytest.mark.parametrize("schemes", [("DD", "DD"), ("DG", "DG"), ("DD", "DG")])
@pytest.mark.parametrize("orders", ORDERS)
@pytest.mark.parametrize("fc", [True, False])
def test_cell_polynomial_exactness_and_balance(oracle, schemes, orders, fc):
    ox, oy = orders
    rng = np.random.default_rng(ox * 10 + oy)
    top = (ox, oy) if schemes == ("DD", "DD") else (ox - 1, oy - 1)
    for mu, eta in [(0.6, 0.3), (-0.5, 0.7), (0.2, -0.8)]:
        for dxd, dyd in [(ox - 1, oy - 1), (ox, oy)]:
            err, bal = _cell_case(oracle, schemes, ox, oy, fc, dxd, dyd, mu, eta, rng)
            assert bal < 1e-13
            if dxd <= top[0] and dyd <= top[1]:
                assert err < 1e-13, (schemes, orders, fc, dxd, dyd, err)
            elif schemes == ("DG", "DG"):
                assert err > 1e-6, (schemes, orders, fc, dxd, dyd, err)


def test_cell_dd11_is_diamond(oracle):
    for mu, eta, q, fxi, fyi in [(0.6, 0.3, 0.7, 0.4, 0.1), (-0.25, 0.8, 2.0, 0.0, 0.5)]:
        psi, fx, fy = oracle.cell_2d(("DD", "DD"), (1, 1), True, mu, eta, 1.3, 0.2, 0.3, [q], [fxi], [fyi])
        hx, hy = abs(mu) / 0.2, abs(eta) / 0.3
        want = (q + 2 * hx * fxi + 2 * hy * fyi) / (1.3 + 2 * hx + 2 * hy)
        assert abs(psi[0] - want) < 1e-15
        assert abs(fx[0] - (2 * want - fxi)) < 1e-15 and abs(fy[0] - (2 * want - fyi)) < 1e-15


def _set(p, sx, sy, ox, oy, fc):
    p.scheme, p.scheme_y, p.scheme_order, p.scheme_order_y, p.full_coupling = sx, sy, ox, oy, fc
    return p


CONFIGS = [("DD", "DD", 2, 2, True), ("DG", "DG", 2, 2, False), ("DD", "DG", 3, 2, True), ("DG", "DD", 1, 3, True),
           ("DG", "DG", 3, 3, False)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_infinite_medium_has_no_slopes(oracle, cfg):
    from radiant_hip import problems as P
    base = P.small_problem(2, N=4, G=2, L=2, Nmat=1, tol=1e-13, seed=42, bc=[1, 1, 1, 1])
    base.qvol = np.ones_like(base.qvol)
    ref, _ = oracle.solve(base)
    p = _set(copy.deepcopy(base), *cfg)
    phi, st = oracle.solve(p)
    K = oracle.spatial_moments(p)
    assert phi.shape[2] == 16 * K and all(s["converged"] for s in st)
    assert rel_err(phi[:, :, 0::K], ref) < 1e-10
    assert np.abs(np.delete(phi, np.arange(0, phi.shape[2], K), axis=2)).max() < 1e-10 * np.abs(ref).max()


# ------------------------------------------------------------------------------------------
# GPU parity
# ------------------------------------------------------------------------------------------

def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


GPU_CASES = [(cfg, None) for cfg in CONFIGS] + [(("DD", "DD", 3, 3, True), None),
                                                (("DD", "DD", 2, 2, True), [1, 2, 0, 1]),
                                                (("DG", "DG", 2, 3, False), [1, 1, 1, 1]),
                                                # order 4: 12 moments with full coupling, 7 without
                                                (("DD", "DG", 4, 3, True), [2, 2, 0, 1]),
                                                (("DG", "DD", 4, 4, False), None),
                                                (("DD", "DD", 1, 4, True), [1, 1, 1, 1])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_scheme2d_gpu_parity(oracle, case):
    from radiant_hip import problems as P, solver as S
    cfg, bc = GPU_CASES[case]
    p = _set(P.small_problem(2, dims=(6, 5), G=2, L=2, tol=1e-12, seed=80 + case, bc=bc), *cfg)
    _say("scheme2d", cfg, bc)
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(400 + case)
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
def test_scheme2d_gpu_accelerations_and_batch(oracle):
    from radiant_hip import problems as P, solver as S
    p = _set(P.small_problem(2, N=6, G=2, L=2, tol=1e-12, seed=95), "DG", "DD", 2, 2, True)
    ref, _ = oracle.solve(p)
    for acc in ("anderson", "bicgstab"):
        q = copy.deepcopy(p)
        q.accel = acc
        r, _ = oracle.solve(q)
        phi, _ = S.compute_flux(q)
        e = rel_err(phi, r)
        _say("scheme2d", acc, e)
        assert e < 1e-10, acc
    phis, _, _ = S.solve_batch([p, p], precision="mixed-fp32")
    for ph in phis:
        assert rel_err(ph, ref) < 1e-6


@pytest.mark.gpu
def test_scheme2d_moment_limit_refused():
    """Orders above 8 per axis are refused at set-up with the limit named (orders 4 x 4 with full
    coupling, 16 moments per cell, run on the general walker)."""
    from radiant_hip import problems as P, solver as S
    from radiant_hip._lib import RTError
    p = _set(P.small_problem(2, dims=(6, 5), G=2, L=2, seed=99), "DD", "DD", 9, 9, False)
    with pytest.raises(RTError, match=r"orders must lie in 1\.\.8 in 2D"):
        S.ResidentProblem(p)
