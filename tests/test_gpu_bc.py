"""Reflective and periodic boundaries on the GPU: single passes and full solves against the oracle,
whose boundary branches are pinned in test_oracle_bc.py.  Tolerances as for the void parity tests:
1e-12 per pass, 1e-10 per FP64 solve (faithful or reduced state), 1e-6 for the mixed modes."""
import sys

import numpy as np
import pytest

from util import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _P():
    from radiant_hip import problems
    return problems


def _S():
    from radiant_hip import solver
    return solver


# (ndims, bc per face x-, x+, y-, y+, z-, z+, grid)
CASES = [
    (1, [1, 1], None), (1, [2, 2], None), (1, [1, 0], None), (1, [0, 2], None),
    (2, [1, 1, 1, 1], None), (2, [2, 2, 0, 1], None), (2, [0, 1, 2, 2], None), (2, [1, 0, 0, 0], (9, 5)),
    (3, [1] * 6, None), (3, [2, 2, 1, 0, 0, 1], None), (3, [0, 0, 2, 2, 1, 1], (7, 4, 5)),
    (3, [1, 2, 0, 1, 2, 0], (70, 3, 3)),
]


def _prob(ndims, bc, dims, tol=1e-12):
    P = _P()
    N = {1: 9, 2: 7, 3: 5}[ndims]
    return P.small_problem(ndims, N=N, dims=dims, G=2, L=2, tol=tol, seed=31 + ndims, bc=bc)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_bc_pass_parity(oracle, case):
    ndims, bc, dims = CASES[case]
    S = _S()
    prob = _prob(ndims, bc, dims)
    _say("bc pass parity", ndims, bc, dims)
    h = S.ResidentProblem(prob)
    try:
        nd, np_, nps, Sl = oracle.sizes(prob)
        assert Sl == h.state_length()
        rng = np.random.default_rng(100 + case)
        z = rng.standard_normal(Sl)
        q = np.zeros(np_ * prob.Nc)
        actp = [p for p in range(np_) if h.d.ang.pl[p] <= prob.Lxs]
        for p in actp:
            q[p::np_] = rng.standard_normal(prob.Nc)
        for mode in (0, 1, 2):
            ref = oracle.transport_pass(prob, 1, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                ref = z - ref
            got = h.transport_pass(1, mode, q, z)
            e = rel_err(got, ref)
            _say("  mode", mode, "err", e)
            assert e < 1e-12
    finally:
        h.close()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_bc_solve_parity(oracle, case):
    ndims, bc, dims = CASES[case]
    S = _S()
    prob = _prob(ndims, bc, dims)
    ref, rst = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("bc solve parity", ndims, bc, dims, e, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10
    assert all(s["converged"] for s in st)


@pytest.mark.parametrize("accel", ["anderson", "livolant", "bicgstab"])
def test_bc_other_accelerations(oracle, accel):
    S = _S()
    prob = _prob(2, [1, 2, 2, 0], None, tol=1e-12)
    prob.accel = accel
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("bc", accel, e)
    assert e < 1e-10


def test_bc_infinite_medium_closed_form():
    """All faces reflective, one material, uniform source: the closed-form infinite-medium flux."""
    S = _S()
    P = _P()
    p = P.small_problem(3, N=6, G=3, L=2, Nmat=1, tol=1e-13, seed=5, bc=[1] * 6)
    q = np.zeros_like(p.qvol)
    q[0, :] = 1.0
    p.qvol = q
    phi, _ = S.compute_flux(p)
    st, ss = p.sigt[:, 0], p.sigs[0, :, :, 0]
    ref = np.zeros(p.G)
    for g in range(p.G):
        ref[g] = (q[g, 0] + sum(ss[gi, g] * ref[gi] for gi in range(g))) / (st[g] - ss[g, g])
    for g in range(p.G):
        assert np.allclose(phi[g, 0, :], ref[g], rtol=1e-9, atol=0)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_bc_reduced_and_mixed_parity(oracle, case):
    """The reduced state carries the boundary moments: plain FP64 to 1e-10, mixed FP32 to 1e-6."""
    ndims, bc, dims = CASES[case]
    S = _S()
    prob = _prob(ndims, bc, dims)
    ref, _ = oracle.solve(prob)
    phi, st = S.compute_flux(prob, state="reduced")
    e = rel_err(phi, ref)
    _say("bc reduced fp64", ndims, bc, dims, e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    mx, st = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("bc mixed fp32", ndims, bc, dims, e, [s["refinements"] for s in st] if "refinements" in st[0] else "")
    assert e < 1e-6 and all(s["converged"] for s in st)


def test_bc_bf16_and_batch(oracle):
    S = _S()
    prob = _prob(2, [1, 2, 2, 0], None)
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob, precision="mixed-bf16", state="reduced")
    e = rel_err(phi, ref)
    _say("bc mixed bf16", e)
    assert e < 1e-6
    # in a lockstep batch (boundary moments in the reduced state)
    phis, _, _ = S.solve_batch([prob, prob], precision="mixed-fp32")
    for ph in phis:
        assert rel_err(ph, ref) < 1e-6
