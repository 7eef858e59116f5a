"""Batched (lockstep) and served solves: every result within 1e-6 of the FP64 solve, and the same
alone and in the mix (the batched kernels never combine numbers of different problems)."""
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
    return [P.config_2d(N=24, G=3, L=2, quad_order=4, seed=11),
            P.config_2d(N=40, G=4, L=3, quad_order=6, block=5, seed=12),
            P.small_problem(3, N=6, G=3, L=2, tol=1e-10),
            P.small_problem(1, N=30, G=2, L=3, tol=1e-10),
            P.config_3d(N=10, width=1.0, G=3, L=3, quad_order=6, seed=13)]


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_batch_matches_fp64_and_alone(precision):
    from radiant_hip import solver as S
    probs = _mix()
    _say("batch", precision)
    phis, stats, info = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = S.compute_flux(p)
        e = rel_err(phis[i], ref)
        _say("  problem", i, "err vs fp64", e, [s["refinements"] for s in stats[i]])
        assert e < 1e-6
        assert all(s["converged"] for s in stats[i])
    # alone vs in the mix: identical
    for i in (0, 2):
        alone, _, _ = S.solve_batch([probs[i]], precision=precision)
        d = float(np.abs(alone[0] - phis[i]).max())
        _say("  alone vs mix", i, d)
        assert d == 0.0


def test_server_continuous_batching():
    from radiant_hip import solver as S
    from radiant_hip import problems as P
    probs = _mix() + [P.config_2d(N=32, G=3, L=2, quad_order=4, seed=14)]
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=3, precision="mixed-fp32")
    try:
        ids = []
        for k, p in enumerate(probs):
            ids.append(srv.submit(p))
            time.sleep(0.01)  # requests arrive while earlier ones run
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
        for rid, p in zip(ids, probs):
            phi, st, lat = res[rid]
            ref, _ = S.compute_flux(p)
            e = rel_err(phi, ref)
            alone = S.serve_alone(p)  # the same request alone in a server
            d = float(np.abs(alone - phi).max())
            _say("  request", rid, "latency ms", round(lat, 2), "err", e, "alone vs served", d)
            assert e < 1e-6 and lat > 0
            assert d == 0.0  # the same alone and in the served mix
    finally:
        srv.close()


def test_server_cancel_running_then_serve():
    """A request cancelled while it runs leaves its slot; the engine goes on serving."""
    from radiant_hip import solver as S, _lib
    from radiant_hip import problems as P
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=2, precision="mixed-fp32")
    try:
        big = srv.submit(P.config_2d(N=64, G=8, L=3, quad_order=8, seed=21))
        time.sleep(0.05)
        srv.cancel(big)
        t0 = time.time()
        while True:
            try:
                r = srv.poll(big)
                if r is not None:
                    _say("  cancelled request finished first")
                    break
            except _lib.RTError as e:
                assert e.code == _lib.RT_ECANCELLED
                _say("  cancelled")
                break
            assert time.time() - t0 < 60
            time.sleep(0.005)
        p = P.config_2d(N=24, G=3, L=2, quad_order=4, seed=22)
        rid = srv.submit(p)
        t0 = time.time()
        r = None
        while r is None:
            assert time.time() - t0 < 60
            r = srv.poll(rid)
            time.sleep(0.005)
        ref, _ = S.compute_flux(p)
        e = rel_err(r[0], ref)
        _say("  after cancel: err", e)
        assert e < 1e-6
    finally:
        srv.close()


def test_server_fp64_workers():
    """FP64 servers solve each request on its own stream: results equal the one-shot solve."""
    from radiant_hip import solver as S
    from radiant_hip import problems as P
    probs = [P.config_2d(N=24, G=3, L=2, quad_order=4, seed=31), P.small_problem(1, N=30, G=2, L=3, tol=1e-10)]
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=2, precision="fp64")
    try:
        ids = [srv.submit(p) for p in probs]
        res = {}
        t0 = time.time()
        while len(res) < len(ids):
            assert time.time() - t0 < 60
            for rid in ids:
                if rid not in res:
                    r = srv.poll(rid)
                    if r is not None:
                        res[rid] = r
            time.sleep(0.005)
        for rid, p in zip(ids, probs):
            ref, _ = S.compute_flux(p)
            e = rel_err(res[rid][0], ref)
            _say("  fp64 request", rid, "err", e)
            assert e < 1e-10
    finally:
        srv.close()


def test_server_budget_refusal_and_cancel():
    from radiant_hip import solver as S, _lib
    from radiant_hip import problems as P
    srv = S.Server(hbm_budget_bytes=1024, max_active=2, precision="mixed-fp32")
    try:
        rid = srv.submit(P.small_problem(2, N=6, G=2))
        t0 = time.time()
        while True:
            try:
                r = srv.poll(rid)
            except _lib.RTError as e:
                assert e.code == _lib.RT_ENOMEM
                break
            assert r is None and time.time() - t0 < 30
            time.sleep(0.01)
    finally:
        srv.close()


def test_server_budget_gated_wait():
    """The HBM budget, not the slot count, makes a request wait: with room for one request's
    bytes, the second stays queued while the first one's finished flux is resident and
    uncollected, and is admitted and solved once a poll collects the first (or cancel drops it)."""
    from radiant_hip import solver as S
    from radiant_hip import problems as P
    p1 = P.config_2d(N=48, G=4, L=3, quad_order=8, seed=41)
    p2 = P.config_2d(N=40, G=3, L=2, quad_order=8, seed=42)
    p3 = P.config_2d(N=32, G=3, L=2, quad_order=4, seed=43)
    b = [S.Desc(p, "mixed-fp32", "reduced").device_bytes() for p in (p1, p2, p3)]
    budget = b[0] + b[2] - 1  # room for p1 alone, p2 + p3, not p1 + p2 nor p1 + p3
    assert budget >= max(b) and b[0] + b[1] > budget and b[0] + b[2] > budget
    srv = S.Server(hbm_budget_bytes=budget, max_active=8, precision="mixed-fp32")
    try:
        r1 = srv.submit(p1)
        r2 = srv.submit(p2)
        t0 = time.time()
        while not srv.finished(r1):
            assert time.time() - t0 < 60
            time.sleep(0.005)
        time.sleep(0.5)
        assert not srv.finished(r2), "admitted although the budget holds the first result"
        phi1 = srv.poll(r1)[0]  # collecting frees the budget
        t0 = time.time()
        r = None
        while r is None:
            assert time.time() - t0 < 60
            r = srv.poll(r2)
            time.sleep(0.005)
        _say("  budget-gated request latency ms", round(r[2], 1))
        for phi, p in ((phi1, p1), (r[0], p2)):
            ref, _ = S.compute_flux(p)
            assert rel_err(phi, ref) < 1e-6
        # a finished result dropped with cancel also returns its bytes to the budget
        r3 = srv.submit(p3)
        r4 = srv.submit(p1)
        t0 = time.time()
        while not srv.finished(r3):
            assert time.time() - t0 < 60
            time.sleep(0.005)
        time.sleep(0.3)
        assert not srv.finished(r4)
        lat3, st3 = srv.discard(r3)
        assert lat3 > 0 and all(s["converged"] for s in st3)
        t0 = time.time()
        r = None
        while r is None:
            assert time.time() - t0 < 60
            r = srv.poll(r4)
            time.sleep(0.005)
        assert rel_err(r[0], S.compute_flux(p1)[0]) < 1e-6
    finally:
        srv.close()


def test_server_destroy_with_uncollected_results():
    """Destroying a server that still holds finished, uncollected results releases them before
    the engine's stream goes; the library keeps working afterwards."""
    from radiant_hip import solver as S
    from radiant_hip import problems as P
    probs = [P.config_2d(N=24, G=3, L=2, quad_order=4, seed=51 + k) for k in range(3)]
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=3, precision="mixed-fp32")
    ids = [srv.submit(p) for p in probs]
    t0 = time.time()
    while not all(srv.finished(r) for r in ids):
        assert time.time() - t0 < 60
        time.sleep(0.005)
    srv.close()
    phi, _ = S.compute_flux(probs[0])
    assert np.isfinite(phi).all()


def test_worklist_sweep_equals_per_task_launch():
    """The lockstep line sweeps as a persistent work list (RADIANT_HIP_WORKLIST=1: a fixed grid of
    workgroups taking tasks from a counter) give the same bits as one workgroup per task, for an
    uneven mix of 1D/2D problems in mixed FP32 and BF16, and with 7 workgroups for all the tasks."""
    import os
    from radiant_hip import problems as P, solver as S
    probs = [P.small_problem(2, dims=(20 + 9 * i, 6 + 5 * i), G=2 + i % 3, L=2, seed=700 + i) for i in range(6)]
    probs += [P.small_problem(1, N=30 + 7 * i, G=3, L=3, seed=720 + i) for i in range(3)]
    for precision, grid in (("mixed-fp32", None), ("mixed-bf16", None), ("mixed-fp32", "7")):
        ref, _, _ = S.solve_batch(probs, precision=precision)
        os.environ["RADIANT_HIP_WORKLIST"] = "1"
        if grid:
            os.environ["RADIANT_HIP_WORKLIST_GRID"] = grid  # 7 workgroups: many tasks each
        try:
            got, st, _ = S.solve_batch(probs, precision=precision)
        finally:
            os.environ.pop("RADIANT_HIP_WORKLIST", None)
            os.environ.pop("RADIANT_HIP_WORKLIST_GRID", None)
        for i, (a, b) in enumerate(zip(ref, got)):
            assert np.array_equal(a, b), (precision, i)
            assert all(s["converged"] for s in st[i])
