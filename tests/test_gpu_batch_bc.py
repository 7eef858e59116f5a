"""Reflective/periodic faces and surface sources in lockstep batches and the server's lockstep
engine: such a problem's reduced Krylov state carries its boundary moments, its incoming face
fluxes are formed before the sweeps of each round (inflow-taking sweep launches per precision)
and its exit-face moments after the last D contraction.

Checked against the oracle's FP64 solve (1e-6 in the mixed modes) for an uneven mix of void,
reflective, periodic and surface-source problems in 1D/2D/3D, bit-identical alone and in the mix,
and through the server."""
import sys
import time

import numpy as np
import pytest

from util import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mix():
    from radiant_hip import problems as P
    from radiant_hip import sources as SRC
    out = [P.small_problem(2, N=7, G=2, L=2, tol=1e-12, seed=33, bc=[1, 2, 2, 0]),
           P.small_problem(1, N=9, G=2, L=2, tol=1e-12, seed=32, bc=[1, 2]),
           P.small_problem(3, N=5, G=2, L=2, tol=1e-12, seed=34, bc=[1, 2, 0, 1, 2, 0]),
           P.small_problem(2, N=8, G=3, L=2, tol=1e-12, seed=35),                       # void
           P.small_problem(2, dims=(20, 6), G=2, L=2, tol=1e-12, seed=36, bc=[2, 2, 1, 1])]
    s = P.small_problem(2, N=8, G=2, L=2, tol=1e-12, seed=71, bc=[0, 0, 1, 0])        # beam on X-
    s.qvol = np.zeros_like(s.qvol)
    SRC.add_surface_source(s, "X-", [0.8, 0.5, 0.11 ** 0.5], intensity=1.0, group=0,
                           bounds={"x": (1.0, 6.0), "y": (1.0, 5.0), "z": (2.0, 7.0)})
    out.append(s)
    return out


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_batch_with_boundaries(oracle, precision):
    from radiant_hip import solver as S
    probs = _mix()
    phis, stats, _ = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = oracle.solve(p)
        e = rel_err(phis[i], ref)
        _say("batch", precision, "problem", i, "bc", p.bc, "err vs oracle", e)
        assert e < 1e-6
        assert all(s["converged"] for s in stats[i])
    for i in (0, 5):
        alone, _, _ = S.solve_batch([probs[i]], precision=precision)
        d = float(np.abs(alone[0] - phis[i]).max())
        _say("  alone vs mix", i, d)
        assert d == 0.0


def test_batch_with_boundaries_kernel_times():
    """The boundary problems run in the lockstep solver itself (kernel times are recorded by
    lockstep batches only)."""
    from radiant_hip import solver as S
    import os
    os.environ["RADIANT_HIP_PROFILE"] = "1"
    try:
        rb = S.ResidentBatch(_mix()[:3], precision="mixed-fp32")
        rb.solve()
        ms, n, upd, fl, by = rb.kernel_times()
        rb.close()
    finally:
        os.environ.pop("RADIANT_HIP_PROFILE", None)
    _say("sweep launches", n[:, 1])
    assert n[:, 1].sum() > 0


def test_server_with_boundaries(oracle):
    from radiant_hip import solver as S
    probs = _mix()
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=3, precision="mixed-fp32")
    try:
        ids = []
        for p in probs:
            ids.append(srv.submit(p))
            time.sleep(0.01)
        res = {}
        t0 = time.time()
        while len(res) < len(ids) and time.time() - t0 < 90:
            for rid in ids:
                if rid not in res:
                    r = srv.poll(rid)
                    if r is not None:
                        res[rid] = r
            time.sleep(0.005)
        assert len(res) == len(ids)
        for k, rid in enumerate(ids):
            phi, _, lat = res[rid]
            ref, _ = oracle.solve(probs[k])
            alone = S.serve_alone(probs[k])  # the same request alone in a server
            e, d = rel_err(phi, ref), float(np.abs(alone - phi).max())
            _say("  served", k, "bc", probs[k].bc, "err", e, "vs alone", d)
            assert e < 1e-6 and d == 0.0
    finally:
        srv.close()
