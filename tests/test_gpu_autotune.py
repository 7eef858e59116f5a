"""Launch-shape autotuner of the sweeps (rt_autotune): every candidate schedule computes the same
bits, so solves with the default shapes (mode 0), with a new candidate on every launch (mode 2)
and with online tuning (mode 1) must agree exactly, for the 1D/2D line kernel at 1, 2 and 4 cells
per lane, the 3D scan kernel, inflow-taking passes, mixed precision and a lockstep batch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cases():
    from radiant_hip import problems as P
    out = []
    out.append(("2D Nx 12", P.small_problem(2, N=12, G=2, L=2, seed=301), "fp64"))
    out.append(("2D 100x6", P.small_problem(2, dims=(100, 6), G=2, L=2, seed=302), "mixed-fp32"))
    out.append(("2D 200x5", P.small_problem(2, dims=(200, 5), G=2, L=2, seed=303), "fp64"))
    out.append(("2D reflective", P.small_problem(2, N=10, G=2, L=2, seed=304, bc=[1, 2, 0, 1]), "mixed-fp32"))
    out.append(("3D 6^3", P.small_problem(3, N=6, G=2, L=2, seed=305), "fp64"))
    out.append(("3D 70x4x5", P.small_problem(3, dims=(70, 4, 5), G=2, L=2, seed=306), "mixed-fp32"))
    return out


def _solve_all(S):
    res = []
    for name, p, prec in _cases():
        phi, _ = S.compute_flux(p, precision=prec, state="reduced" if prec != "fp64" else "faithful")
        res.append(phi)
    probs = [c[1] for c in _cases()[:2]]
    phis, _, _ = S.solve_batch(probs, precision="mixed-fp32")
    return res + phis


def test_autotune_modes_agree_bitwise():
    from radiant_hip import solver as S
    try:
        S.autotune(0)
        ref = _solve_all(S)
        S.autotune(2)
        cyc = _solve_all(S)
        for i, (a, b) in enumerate(zip(ref, cyc)):
            d = float(np.abs(a - b).max())
            _say("mode 2 vs 0", i, d)
            assert np.array_equal(a, b), i
        S.autotune(1)
        _solve_all(S)
        tuned = _solve_all(S)   # the second round runs the chosen variants
        for i, (a, b) in enumerate(zip(ref, tuned)):
            assert np.array_equal(a, b), i
        rep = S.autotune_report()
        _say(rep)
        assert "line" in rep and "scan3d" in rep and "variant" in rep
    finally:
        S.autotune(1)


def test_autotune_concurrent_sub_batches():
    """Eight problems run as four sub-batches on their own host threads while the tuner times
    their first launches: the results equal those of the default shapes."""
    from radiant_hip import problems as P, solver as S
    probs = [P.small_problem(2, dims=(40 + 8 * i, 6), G=2, L=2, seed=320 + i) for i in range(8)]
    try:
        S.autotune(0)
        ref, _, _ = S.solve_batch(probs, precision="mixed-fp32")
        S.autotune(1)
        for _ in range(2):
            got, _, _ = S.solve_batch(probs, precision="mixed-fp32")
            for a, b in zip(ref, got):
                assert np.array_equal(a, b)
        _say(S.autotune_report())
    finally:
        S.autotune(1)

