"""Lockstep rounds run as captured HIP graphs (RADIANT_HIP_GRAPHS=1): every round's arena upload,
passes and Krylov launches are captured from the stream and launched as one graph; rounds of one
shape reuse a cached executable graph, updated in place with the round's parameters.

The graph runs the same kernels with the same arguments as the eager rounds, so the fluxes must
be bit-identical to the eager ones: checked for an uneven mix of 1D/2D/3D problems on every one
of several graph solves in one process, with eager solves and synchronisations between them, and
for served requests (continuous batching) against the same problems solved alone, eagerly."""
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
    return [P.config_2d(N=24, G=3, L=2, quad_order=4, seed=21),
            P.config_2d(N=40, G=4, L=3, quad_order=6, block=5, seed=22),
            P.small_problem(3, N=6, G=3, L=2, tol=1e-10),
            P.small_problem(1, N=30, G=2, L=3, tol=1e-10),
            P.small_problem(2, dims=(70, 9), G=3, L=2, seed=23)]


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_graph_rounds_equal_eager(monkeypatch, precision):
    from radiant_hip import solver as S
    probs = _mix()
    monkeypatch.delenv("RADIANT_HIP_GRAPHS", raising=False)
    ref, rst, rinfo = S.solve_batch(probs, precision=precision)
    assert rinfo.graph_launches == 0
    for rep in range(3):
        monkeypatch.setenv("RADIANT_HIP_GRAPHS", "1")
        got, st, info = S.solve_batch(probs, precision=precision)
        _say("graphs", precision, "replay", rep, "launches", info.graph_launches, "instantiated",
             info.graph_instantiations)
        assert info.graph_launches > 0
        assert info.graph_instantiations < info.graph_launches  # most rounds update a cached graph
        for i, (a, b) in enumerate(zip(ref, got)):
            assert np.array_equal(a, b), (rep, i, float(np.abs(a - b).max()))
        assert [[s["niter"] for s in x] for x in st] == [[s["niter"] for s in x] for x in rst]
        # eager work between the graph solves
        monkeypatch.delenv("RADIANT_HIP_GRAPHS")
        again, _, _ = S.solve_batch(probs[:2], precision=precision)
        for a, b in zip(ref[:2], again):
            assert np.array_equal(a, b)


def test_graph_rounds_served(monkeypatch):
    from radiant_hip import solver as S
    probs = _mix()
    monkeypatch.delenv("RADIANT_HIP_GRAPHS", raising=False)
    alone = [S.serve_alone(p) for p in probs]  # each request alone in a server, eager rounds
    monkeypatch.setenv("RADIANT_HIP_GRAPHS", "1")
    srv = S.Server(hbm_budget_bytes=2 * 1024 ** 3, max_active=3, precision="mixed-fp32")
    try:
        ids = []
        for p in probs:
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
        for k, rid in enumerate(ids):
            phi, _, lat = res[rid]
            d = float(np.abs(phi - alone[k]).max())
            _say("  served with graphs", rid, "latency ms", round(lat, 2), "vs alone (eager)", d)
            assert d == 0.0
            ref, _ = S.compute_flux(probs[k])
            assert rel_err(phi, ref) < 1e-6
    finally:
        srv.close()
