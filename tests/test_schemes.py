"""Spatial closures of the 1D cell: high-order diamond difference and the DG family
(SN.set_scheme; scheme_weights.jl:25-90, flux_1D_BTE of sn_1D_bte.jl:27-64, sn_sweep_1D.jl:93-110).

CPU pins of the oracle's cell (closed forms, with psi(u) = sum_j C_j psi_j P_(j-1)(u) on the cell,
u in [-1, 1], C_j = sqrt(2j - 1)):
- a transport solution that is a polynomial of degree d in the cell is reproduced exactly (moments
  and outgoing face flux) by DD of order K for d <= K, by DG for d <= K - 1 and by DG-/DG+ for
  d <= K - 2, and not beyond;
- the zeroth moment obeys the cell balance St psi_1 + |mu|/dx (f_out - f_in) = Q_1 for every
  closure and order;
- DD of order 1 is the classic diamond cell;
- the host's closure weights match the closed form of each family.
Solve-level: an infinite medium (reflective slab, uniform source) has no spatial slopes, so every
closure and order gives the first-order flux with vanishing higher spatial moments.
GPU: passes (1e-12) and solves (1e-10 FP64, 1e-6 mixed FP32) against the oracle."""
import copy
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as Lg

from util import rel_err


def _S():
    from radiant_hip import problems, schemes
    return problems, schemes


def _cell_case(oracle, scheme, K, deg, mu, seed):
    rng = np.random.default_rng(seed)
    St, dx = 0.9, 0.37
    c = np.zeros(max(K, deg + 1))
    c[:deg + 1] = rng.standard_normal(deg + 1)
    dc = Lg.legder(c)
    Qc = St * c.copy()
    Qc[:len(dc)] += mu * 2 / dx * dc          # mu dpsi/dx + St psi, dx/du = dx/2
    Cj = np.sqrt(2 * np.arange(1, len(c) + 1) - 1)
    Qn = (Qc / Cj)[:K]
    sx = 1 if mu > 0 else -1
    fin = Lg.legval(-sx, c)
    psi, fout = oracle.cell_1d(scheme, K, mu, St, dx, Qn, fin)
    err = max(np.abs(psi - (c / Cj)[:K]).max(), abs(fout - Lg.legval(sx, c)))
    bal = St * psi[0] + abs(mu) / dx * (fout - fin) - Qn[0]
    return err, abs(bal)


EXACT = {"DD": 0, "DG": 1, "DG-": 2, "DG+": 2}   # exact up to degree K - EXACT


@pytest.mark.parametrize("scheme", ["DD", "DG", "DG-", "DG+"])
@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 6, 7, 8])
def test_cell_polynomial_exactness_and_balance(oracle, scheme, K):
    if scheme in ("DG-", "DG+") and K < 2:
        pytest.skip("the modified DG closures start at order 2")
    top = K - EXACT[scheme]
    for mu in (0.7, -0.4):
        for deg in range(0, K + 1):
            err, bal = _cell_case(oracle, scheme, K, deg, mu, seed=10 * K + deg)
            assert bal < 1e-13, (scheme, K, deg, mu)
            if deg <= top:
                assert err < 1e-13, (scheme, K, deg, mu, err)
            else:
                assert err > 1e-6, (scheme, K, deg, mu, err)   # not exact beyond its degree


def test_cell_dd1_is_diamond(oracle):
    for mu, St, dx, q, f in [(0.6, 1.3, 0.2, 0.7, 0.4), (-0.25, 0.4, 1.1, 2.0, 0.0), (0.0, 0.8, 0.5, 1.0, 0.3)]:
        psi, fo = oracle.cell_1d("DD", 1, mu, St, dx, [q], f)
        h = abs(mu) / dx
        assert abs(psi[0] - (q + 2 * h * f) / (St + 2 * h)) < 1e-15
        assert abs(fo - ((2 * psi[0] - f) if mu != 0 else f)) < 1e-15


def test_host_weights():
    _, sch = _S()
    for K in (1, 2, 3, 4):
        dd = sch.weights_1d("DD", K)
        assert dd[0] == (-1) ** K and all(dd[1 + n] == 1 - (-1) ** (K - n) for n in range(K))
        dg = sch.weights_1d("DG", K)
        assert dg[0] == 0 and np.all(dg[1:] == 1)
        if K > 1:
            assert sch.weights_1d("DG-", K)[K] == (K - 1) / (2 * K - 1)
            assert sch.weights_1d("DG+", K)[K] == 1001
    with pytest.raises(ValueError):
        sch.weights_1d("DG-", 1)
    # the adaptive closure's base weights: those of DD at order 1 and DG at order 2 (scheme_weights.jl:47,52)
    assert np.array_equal(sch.weights_1d("AWD", 1), sch.weights_1d("DD", 1))
    assert np.array_equal(sch.weights_1d("AWD", 2), sch.weights_1d("DG", 2))
    with pytest.raises(ValueError):
        sch.weights_1d("AWD", 3)


def _slab(scheme, K, bc=None, seed=41, G=3, N=9):
    P, _ = _S()
    p = P.small_problem(1, N=N, G=G, L=2, tol=1e-13, seed=seed, bc=bc)
    p.scheme, p.scheme_order = scheme, K
    return p


@pytest.mark.parametrize("scheme,K", [("DD", 2), ("DG", 1), ("DG", 3), ("DG-", 2), ("DG+", 3), ("DD", 4)])
def test_infinite_medium_has_no_slopes(oracle, scheme, K):
    P, _ = _S()
    base = P.small_problem(1, N=6, G=3, L=2, Nmat=1, tol=1e-13, seed=42, bc=[1, 1])
    base.qvol = np.ones_like(base.qvol)
    ref, _ = oracle.solve(base)
    p = copy.deepcopy(base)
    p.scheme, p.scheme_order = scheme, K
    phi, st = oracle.solve(p)
    assert phi.shape == (3, ref.shape[1], 6 * K)
    assert all(s["converged"] for s in st)
    for j in range(K):
        mom = phi[:, :, j::K]
        if j == 0:
            assert rel_err(mom, ref) < 1e-10
        else:
            assert np.abs(mom).max() < 1e-10 * np.abs(ref).max()


def test_higher_order_differs_and_refusals(oracle):
    P, _ = _S()
    a, _ = oracle.solve(_slab("DD", 1))
    b, _ = oracle.solve(_slab("DG", 2))
    assert rel_err(b[:, :, 0::2], a) > 1e-4
    for nd in (2, 3):  # the modified DG closures exist in 1D only
        r = P.small_problem(nd, N=3, G=2)
        r.scheme, r.scheme_order = "DG-", 2
        with pytest.raises(RuntimeError, match="not implemented"):
            oracle.solve(r)
    q = _slab("DG-", 2)   # with slowing down the closures are DD or DG (scheme_weights.jl, isCSD)
    q.csd, q.Sb, q.dE = True, np.ones((q.G + 1, q.Nmat)), np.ones(q.G)
    with pytest.raises(RuntimeError, match="not implemented"):
        oracle.solve(q)


# ------------------------------------------------------------------------------------------
# GPU parity
# ------------------------------------------------------------------------------------------

def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


GPU_CASES = [("DG", 1, None), ("DD", 2, None), ("DG", 3, None), ("DD", 4, None), ("DG-", 2, None),
             ("DG+", 3, None), ("DD", 3, [1, 2]), ("DG", 2, [2, 2]), ("DD", 2, None, 300),
             ("DD", 6, None), ("DG", 8, [1, 2]), ("DG+", 5, None), ("DG-", 7, [2, 2])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_scheme_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    scheme, K, bc, *n = GPU_CASES[case]
    p = _slab(scheme, K, bc=bc, seed=50 + case, N=n[0] if n else 9)
    p.tol = 1e-12
    _say("scheme", scheme, K, bc, p.Nx)
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(200 + case)
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
    if bc is None:
        mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
        e = rel_err(mx, ref)
        _say("  mixed-fp32", e)
        assert e < 1e-6


@pytest.mark.gpu
def test_scheme_gpu_accelerations_and_batch(oracle):
    from radiant_hip import solver as S
    p = _slab("DG", 2, seed=61)
    p.tol = 1e-12
    ref, _ = oracle.solve(p)
    for acc in ("anderson", "bicgstab", "livolant"):
        q = copy.deepcopy(p)
        q.accel = acc
        r, _ = oracle.solve(q)
        phi, _ = S.compute_flux(q)
        e = rel_err(phi, r)
        _say("scheme", acc, e)
        assert e < 1e-10, acc
    phis, _, _ = S.solve_batch([p, p], precision="mixed-fp32")
    for ph in phis:
        assert rel_err(ph, ref) < 1e-6
