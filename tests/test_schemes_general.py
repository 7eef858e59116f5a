"""2D/3D closures beyond 12 moments per cell: the general walker (sweep_hog.hip), which solves the
cell system of any order up to 8 per axis and 64 moments (sn_2D_bte.jl:35-116, sn_3D_bte.jl:34-152
take any order; the register walkers hold at most 12 moments).

CPU: the oracle's solves of these closures converge, and an infinite medium keeps its first-order
flux with vanishing slopes at these orders too (the cell systems are those the lower-order tests pin
by polynomial exactness).
GPU: passes (1e-12) and solves (1e-10 FP64, 1e-6 mixed FP32) against the oracle; a batch or the
server holding such a problem solves it on its own stream; orders above 8 or more than 64 moments
are refused.  The slowing-down cells (E, x, y, z) beyond 8 moments take the same checks."""
import copy
import sys

import numpy as np
import pytest

from util import rel_err


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _set(p, orders, fc, scheme="DD"):
    p.scheme = p.scheme_y = scheme
    p.scheme_order, p.scheme_order_y = orders[0], orders[1]
    if len(orders) == 3:
        p.scheme_z, p.scheme_order_z = scheme, orders[2]
    p.full_coupling = fc
    return p


# (dims, orders, full coupling, scheme, boundary conditions): 15-64 moments per cell, or orders
# above the register walkers' 4 (2D) and 3 (3D)
CASES = [((6, 5), (4, 4), True, "DD", None),                 # 16 moments
         ((5, 4), (5, 3), True, "DG", [1, 2, 0, 1]),         # 15
         ((4, 4), (8, 8), False, "DD", None),                # order 8, 15 moments
         ((3, 3), (8, 8), True, "DG", [1, 1, 1, 1]),         # 64
         ((4, 3, 3), (3, 3, 2), True, "DD", None),           # 18
         ((3, 3, 3), (3, 3, 3), True, "DG", [1, 2, 0, 1, 1, 0]),  # 27
         ((3, 4, 3), (4, 5, 4), False, "DD", None)]          # orders 4-5, 11 moments


def _problem(case, tol=1e-12):
    from radiant_hip import problems as P
    dims, orders, fc, scheme, bc = CASES[case]
    p = P.small_problem(len(dims), dims=dims, G=2, L=1, tol=tol, seed=170 + case, bc=bc)
    return _set(p, orders, fc, scheme)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_converges(oracle, case):
    p = _problem(case, tol=1e-10)
    phi, st = oracle.solve(p)
    assert all(s["converged"] for s in st)
    assert np.isfinite(phi).all() and np.abs(phi).max() > 0


@pytest.mark.parametrize("nd,orders,fc", [(2, (5, 4), True), (3, (3, 3, 3), True), (3, (4, 2, 3), False)])
def test_infinite_medium_has_no_slopes(oracle, nd, orders, fc):
    from radiant_hip import problems as P
    dims = (3, 3) if nd == 2 else (2, 3, 2)
    base = P.small_problem(nd, dims=dims, G=2, L=1, Nmat=1, tol=1e-12, seed=190, bc=[1] * (2 * nd))
    base.qvol = np.ones_like(base.qvol)
    ref, _ = oracle.solve(base)
    p = _set(copy.deepcopy(base), orders, fc)
    phi, st = oracle.solve(p)
    assert all(s["converged"] for s in st)
    K = oracle.spatial_moments(p)
    assert phi.shape[-1] == ref.shape[-1] * K
    assert rel_err(phi[:, :, 0::K], ref) < 1e-9
    for j in range(1, K):
        assert np.abs(phi[:, :, j::K]).max() < 1e-9 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_general_closure_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    p = _problem(case)
    _say("general closure", CASES[case])
    ref, rst = oracle.solve(p)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(600 + case)
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
def test_general_closure_in_batch(oracle):
    """The lockstep solver refuses the general walker; the batch entry point then solves that
    problem on its own stream, beside a register-walker problem."""
    from radiant_hip import problems as P, solver as S
    a = _problem(0)
    b = _set(P.small_problem(2, dims=(6, 5), G=2, L=1, tol=1e-12, seed=199), (2, 2), True)
    phis, _, _ = S.solve_batch([a, b], precision="mixed-fp32")
    for p, ph in zip((a, b), phis):
        ref, _ = oracle.solve(p)
        e = rel_err(ph, ref)
        _say("batch", e)
        assert e < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nd,orders,msg", [(2, (9, 2), r"orders must lie in 1\.\.8 in 2D"),
                                            (3, (5, 5, 3), "orders up to 8 per axis and at most 64 moments per cell")])
def test_general_closure_limits_refused(nd, orders, msg):
    """Orders above 8 per axis, or more than 64 moments per cell (5 x 5 x 3 = 75 with full
    coupling), are refused at set-up with the limits named."""
    from radiant_hip import problems as P, solver as S
    from radiant_hip._lib import RTError
    dims = (4, 4) if nd == 2 else (3, 3, 3)
    p = _set(P.small_problem(nd, dims=dims, G=2, L=1, seed=198), orders, True)
    with pytest.raises(RTError, match=msg):
        S.ResidentProblem(p)


# Slowing-down cells (E, x, y, z) beyond the register walker of 8 moments and orders 4 (1D) / 2
# (2D, 3D): (ndims, orders (E, x, y, z), schemes, full coupling, boundary conditions)
BFP_CASES = [(1, (2, 6, 1, 1), ("DG", "DD", "DD", "DD"), True, None),        # 12 moments
             (1, (1, 8, 1, 1), ("DD", "DG", "DD", "DD"), True, [1, 2]),      # order 8
             (2, (2, 3, 3, 1), ("DG", "DD", "DD", "DD"), True, None),        # 18
             (2, (2, 4, 2, 1), ("DD", "DG", "DD", "DD"), False, [1, 1, 1, 1]),
             (3, (2, 2, 2, 2), ("DG", "DD", "DD", "DD"), True, None),        # 16
             (3, (2, 3, 2, 2), ("DD", "DD", "DG", "DD"), False, [1, 1, 0, 0, 2, 2])]


def _bfp_problem(case, tol=1e-12):
    from radiant_hip import problems as P
    nd, orders, schemes, fc, bc = BFP_CASES[case]
    dims = {1: (9,), 2: (5, 4), 3: (3, 4, 3)}[nd]
    p = P.small_problem(nd, dims=dims if nd > 1 else None, N=dims[0], G=3, L=2, tol=tol, seed=220 + case, bc=bc)
    rng = np.random.default_rng(230 + case)
    p.csd = True
    p.Sb = 0.8 * rng.uniform(0.5, 2.0, (p.G + 1, p.Nmat))
    p.dE = rng.uniform(0.5, 1.5, p.G)
    p.scheme_e, p.scheme, p.scheme_y, p.scheme_z = schemes
    p.scheme_order_e, p.scheme_order, p.scheme_order_y, p.scheme_order_z = orders
    p.full_coupling = fc
    return p


@pytest.mark.parametrize("case", range(len(BFP_CASES)))
def test_oracle_converges_slowing_down(oracle, case):
    phi, cut, st = oracle.solve_cutoff(_bfp_problem(case, tol=1e-10))
    assert all(s["converged"] for s in st)
    assert np.isfinite(phi).all() and np.abs(phi).max() > 0 and np.abs(cut).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(BFP_CASES)))
def test_general_slowing_down_closure_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    p = _bfp_problem(case)
    _say("general slowing-down closure", BFP_CASES[case])
    ref, rcut, rst = oracle.solve_cutoff(p)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(700 + case)
        z = rng.standard_normal(Sl)
        q = rng.standard_normal(np_ * p.Nc * K)
        q.reshape(-1, np_)[:, [i for i in range(np_) if h.d.ang.pl[i] > p.Lxs]] = 0.0
        for mode in (0, 1, 2):
            want = oracle.transport_pass(p, 0, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                want = z - want
            e = rel_err(h.transport_pass(0, mode, q, z), want)
            _say("  pass mode", mode, e)
            assert e < 1e-12
    finally:
        h.close()
    phi, cut, st = S.compute_flux_cutoff(p)
    e, ec = rel_err(phi, ref), rel_err(cut, rcut)
    _say("  solve", e, "cutoff", ec)
    assert e < 1e-10 and ec < 1e-10 and all(s["converged"] for s in st)
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("  mixed-fp32", e)
    assert e < 1e-6


@pytest.mark.gpu
def test_general_slowing_down_limits_refused():
    """Energy orders above 2 and slowing-down cells above 64 moments are refused at set-up."""
    from radiant_hip import solver as S
    from radiant_hip._lib import RTError
    p = _bfp_problem(4)
    p.scheme_order, p.scheme_order_y, p.scheme_order_z = 5, 4, 2   # 2 * 5 * 4 * 2 = 80 moments
    with pytest.raises(RTError, match="at most 64 moments per cell"):
        S.ResidentProblem(p)


@pytest.mark.gpu
def test_general_closures_served(oracle):
    """Requests whose cells need the general walkers (a 2D 4 x 4 closure and a 3D 2 x 2 x 2 x 2
    slowing-down cell) through the server in mixed FP32: the lockstep engine does not take them, the
    server solves each on its own stream, and the flux matches the oracle."""
    from radiant_hip import solver as S
    for p in (_problem(0), _bfp_problem(4)):
        ref = oracle.solve(p)[0]
        e = rel_err(S.serve_alone(p, "mixed-fp32"), ref)
        _say("general closure served", e)
        assert e < 1e-6
