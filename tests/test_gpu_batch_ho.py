"""Spatial closures above order 1 (1D DD/DG/DG-/DG+, 2D and 3D DD/DG per axis, full or partial
coupling) in lockstep batches and the server's lockstep engine: each such problem's sweep runs its
own closure kernel in its pass's precision inside the round, with the moment columns of the
contractions and the Krylov state as in its single-problem solve.  Checked against the oracle's
FP64 solve (1e-6 in the mixed modes), mixed with an order-1 problem and reflective/periodic
faces, bit-identical alone and in the mix."""
import os
import sys

import numpy as np
import pytest

from util import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mix():
    from radiant_hip import problems as P
    out = []
    p = P.small_problem(1, N=9, G=3, L=2, tol=1e-12, seed=51)
    p.scheme, p.scheme_order = "DG", 2
    out.append(p)
    p = P.small_problem(1, N=9, G=2, L=2, tol=1e-12, seed=52, bc=[1, 2])
    p.scheme, p.scheme_order = "DG+", 3
    out.append(p)
    p = P.small_problem(2, dims=(6, 5), G=2, L=2, tol=1e-12, seed=53)
    p.scheme, p.scheme_y, p.scheme_order, p.scheme_order_y, p.full_coupling = "DD", "DD", 2, 2, True
    out.append(p)
    p = P.small_problem(2, dims=(6, 5), G=2, L=2, tol=1e-12, seed=54, bc=[1, 2, 0, 1])
    p.scheme, p.scheme_y, p.scheme_order, p.scheme_order_y, p.full_coupling = "DG", "DG", 2, 3, False
    out.append(p)
    p = P.small_problem(3, dims=(4, 3, 5), G=2, L=2, tol=1e-12, seed=55)
    p.scheme, p.scheme_y, p.scheme_z = "DG", "DG", "DG"
    p.scheme_order, p.scheme_order_y, p.scheme_order_z, p.full_coupling = 2, 2, 1, False
    out.append(p)
    out.append(P.small_problem(2, N=8, G=3, L=2, tol=1e-12, seed=56))  # order 1
    p = P.small_problem(3, dims=(3, 4, 3), G=2, L=2, tol=1e-12, seed=57, bc=[0, 1, 2, 2, 1, 0])
    p.scheme, p.scheme_y, p.scheme_z = "DD", "DG", "DD"
    p.scheme_order, p.scheme_order_y, p.scheme_order_z, p.full_coupling = 3, 2, 2, True  # 12 moments
    out.append(p)
    p = P.small_problem(2, dims=(5, 7), G=2, L=2, tol=1e-12, seed=58, bc=[2, 2, 1, 0])
    p.scheme, p.scheme_y, p.scheme_order, p.scheme_order_y, p.full_coupling = "DG", "DD", 4, 2, True
    out.append(p)
    p = P.small_problem(1, N=7, G=2, L=2, tol=1e-12, seed=59)
    p.scheme, p.scheme_order = "DD", 7
    out.append(p)
    return out


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_batch_with_higher_order_closures(oracle, precision):
    from radiant_hip import solver as S
    probs = _mix()
    phis, stats, _ = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = oracle.solve(p)
        e = rel_err(phis[i], ref)
        _say("ho batch", precision, i, e)
        assert e < 1e-6 and all(s["converged"] for s in stats[i])
    for i in (2, 4):
        alone, _, _ = S.solve_batch([probs[i]], precision=precision)
        assert np.array_equal(alone[0], phis[i]), i
    os.environ["RADIANT_HIP_PROFILE"] = "1"
    try:  # kernel times are recorded by lockstep solvers only
        rb = S.ResidentBatch(probs[:3], precision=precision)
        rb.solve()
        _, n, _, _, _ = rb.kernel_times()
        rb.close()
    finally:
        os.environ.pop("RADIANT_HIP_PROFILE", None)
    assert n[:, 1].sum() > 0
