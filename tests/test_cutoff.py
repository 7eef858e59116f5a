"""Flux at the cutoff energy of the slowing-down solver types (sn_flux.jl:151,213-218, returned by
flux.add_flux_cutoff at sn_flux.jl:229): 𝚽cutoff[p, c] = sum_n Dn[p, n] 𝚽E12[n, c], with 𝚽E12 the
energy-face angular flux leaving the last group (not rescaled).

CPU pins of the oracle:
- infinite medium (one material, uniform isotropic source, every face reflective): the scalar
  moment equals 2 phi_{G-1} - E_{G-1} of the moment recursion of the slowing-down pins (E_g the
  energy inflow moment of group g) and the higher moments vanish;
- on a general problem, truncating the groups after g leaves groups 0..g unchanged (downscatter
  Gauss-Seidel and energy flux carried downwards), and the diamond energy closure
  ψE_out = 2ψ − ψE_in of every direction gives, on all moments,
      cutoff(G) = 2 𝚽l[G-1] − (ΔE_{G-1} / ΔE_{G-2}) cutoff(G-1);
  a cutoff taken from the rescaled inflow, from another group, or through M instead of D fails;
- without slowing down the cutoff is zero.
GPU: rt_solve_cutoff, rt_problem_get_flux_cutoff, rt_batch_get_flux_cutoff and
rt_server_poll_cutoff against the oracle (FP64 1e-10, mixed 1e-6)."""
import copy
import sys

import numpy as np
import pytest

from util import rel_err


def _P():
    from radiant_hip import problems
    return problems


def _csd(p, seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    p.csd = True
    p.Sb = scale * rng.uniform(0.5, 2.0, (p.G + 1, p.Nmat))
    p.dE = rng.uniform(0.5, 1.5, p.G)
    return p


def _truncate(p, G):
    q = copy.deepcopy(p)
    q.G = G
    q.sigt = p.sigt[:G].copy()
    q.sigs = p.sigs[:, :G, :G, :].copy()
    q.qvol = p.qvol[:G].copy()
    q.Sb = p.Sb[:G + 1].copy()
    q.dE = p.dE[:G].copy()
    return q


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_cutoff_infinite_medium(oracle, nd):
    P = _P()
    G = 4
    p = P.small_problem(nd, N=4, G=G, L=2, Nmat=1, tol=1e-13, seed=182, bc=[1] * (2 * nd))
    q = np.zeros_like(p.qvol)
    q[0, :] = 1.0
    q[2, :] = 0.3
    p.qvol = q
    _csd(p, seed=4, scale=0.7)
    phi, cut, st = oracle.solve_cutoff(p)
    St, Ss, Sb, dE = p.sigt[:, 0], p.sigs[0, :, :, 0], p.Sb[:, 0], p.dE
    ref = np.zeros(G)
    E = 0.0
    for g in range(G):
        Sm, Sp = Sb[g] / dE[g], Sb[g + 1] / dE[g]
        src = q[g, 0] + sum(Ss[gi, g] * ref[gi] for gi in range(g)) + (Sp + Sm) * E
        ref[g] = src / (St[g] + 2 * Sp - Ss[g, g])
        if g + 1 < G:
            E = (2 * ref[g] - E) * dE[g + 1] / dE[g]
    expect = 2 * ref[G - 1] - E
    assert np.allclose(cut[0, :], expect, rtol=1e-9, atol=0), (cut[0, :2], expect)
    assert np.abs(cut[1:, :]).max() < 1e-9 * abs(expect)
    assert all(s["converged"] for s in st)


@pytest.mark.parametrize("nd,bc", [(1, None), (2, None), (3, None), (2, [1, 0, 2, 2])])
def test_cutoff_group_truncation_identity(oracle, nd, bc):
    P = _P()
    G = 4
    p = _csd(P.small_problem(nd, N=5 if nd == 3 else 7, G=G, L=2, tol=1e-13, seed=190 + nd, bc=bc), seed=7)
    phi, cut, _ = oracle.solve_cutoff(p)
    phi3, cut3, _ = oracle.solve_cutoff(_truncate(p, G - 1))
    assert rel_err(phi3, phi[:G - 1]) < 1e-13
    expect = 2 * phi[G - 1] - (p.dE[G - 1] / p.dE[G - 2]) * cut3
    assert rel_err(cut, expect) < 1e-11
    assert np.abs(cut).max() > 0
    # a cutoff built from the rescaled inflow of the last group, or from the group before, differs
    assert rel_err(cut3 * p.dE[G - 1] / p.dE[G - 2], cut) > 1e-3


def test_cutoff_zero_without_slowing_down(oracle):
    P = _P()
    p = P.small_problem(2, N=6, G=3, tol=1e-12, seed=195)
    _, cut, _ = oracle.solve_cutoff(p)
    assert cut.shape[1] == p.Nc and not cut.any()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(1, None), (2, None), (3, None), (2, [1, 0, 2, 2]), (3, [1, 1, 0, 0, 2, 2])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_cutoff_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    nd, bc = CASES[case]
    P = _P()
    p = _csd(P.small_problem(nd, N=5 if nd == 3 else 7, G=3, L=2, tol=1e-12, seed=284 + case, bc=bc), seed=case)
    _say("cutoff", nd, bc)
    ref, rcut, _ = oracle.solve_cutoff(p)
    phi, cut, st = S.compute_flux_cutoff(p)
    e, ec = rel_err(phi, ref), rel_err(cut, rcut)
    _say("cutoff fp64", nd, bc, e, ec)
    assert e < 1e-10 and ec < 1e-10 and all(s["converged"] for s in st)
    rp = S.ResidentProblem(p, precision="mixed-fp32", state="reduced")
    for _ in range(2):  # a second solve on the same handle gives the same cutoff
        rp.solve()
        ec = rel_err(rp.flux_cutoff(), rcut)
        _say("cutoff mixed-fp32", nd, bc, ec)
        assert ec < 1e-6
    rp.close()


@pytest.mark.gpu
def test_cutoff_zero_without_slowing_down_gpu():
    from radiant_hip import solver as S
    P = _P()
    p = P.small_problem(2, N=6, G=3, tol=1e-12, seed=195)
    _, cut, _ = S.compute_flux_cutoff(p)
    assert not cut.any()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_cutoff_batch_and_server(oracle, precision):
    """Lockstep batch (solved twice on one handle) and server: the cutoff of every slowing-down
    problem against the oracle, including one whose first group has no source (its energy outflow
    is zero, and the next group's inflow must not take a previous solve's)."""
    from radiant_hip import solver as S
    P = _P()
    probs = [_csd(P.small_problem(2, N=7, G=3, L=2, tol=1e-12, seed=291), seed=21),
             _csd(P.small_problem(1, N=12, G=4, L=3, tol=1e-12, seed=292, source_group=1), seed=22),
             _csd(P.small_problem(3, N=5, G=3, L=2, tol=1e-12, seed=293, bc=[1, 2, 0, 1, 2, 0]), seed=23),
             P.small_problem(2, N=8, G=3, L=2, tol=1e-12, seed=294)]
    refs = [oracle.solve_cutoff(p) for p in probs]
    rb = S.ResidentBatch(probs, precision=precision)
    for rep in range(2):
        stats, _ = rb.solve()
        for i, p in enumerate(probs):
            e, ec = rel_err(rb.flux(i), refs[i][0]), rel_err(rb.flux_cutoff(i), refs[i][1])
            _say("cutoff batch", precision, rep, i, e, ec)
            assert e < 1e-6 and all(s["converged"] for s in stats[i])
            if p.csd:
                assert ec < 1e-6
            else:
                assert not rb.flux_cutoff(i).any()
    rb.close()
    srv = S.Server(hbm_budget_bytes=4 << 30, max_active=4, precision=precision)
    try:
        ids = [srv.submit(p) for p in probs[:3]]
        done = {}
        import time
        t0 = time.time()
        while len(done) < len(ids) and time.time() - t0 < 60:
            for k, rid in enumerate(ids):
                if k not in done:
                    r = srv.poll(rid, cutoff=True)
                    if r is not None:
                        done[k] = r
            time.sleep(0.01)
        assert len(done) == len(ids)
        for k, r in done.items():
            ec = rel_err(r[3], refs[k][1])
            _say("cutoff server", precision, k, ec)
            assert rel_err(r[0], refs[k][0]) < 1e-6 and ec < 1e-6
    finally:
        srv.close()


@pytest.mark.gpu
def test_zero_source_first_group_in_batch(oracle):
    """A first group without source in the lockstep batch and through solve_batch, solved twice on
    one handle: its flux fill reaches the device before the next group's out-of-group source."""
    from radiant_hip import solver as S
    P = _P()
    probs = [P.small_problem(2, N=6, G=3, tol=1e-12, source_group=1, seed=301),
             P.small_problem(1, N=10, G=4, tol=1e-12, source_group=2, seed=302),
             P.small_problem(3, N=4, G=3, tol=1e-12, source_group=1, seed=303)]
    refs = [oracle.solve(p)[0] for p in probs]
    phis, stats, _ = S.solve_batch(probs, precision="mixed-fp32")
    for i in range(len(probs)):
        assert rel_err(phis[i], refs[i]) < 1e-6 and np.all(phis[i][0] == 0.0)
        # the lockstep solver learns that the group has no source from its c = T(0) pass
        assert stats[i][0]["niter"] == 0 and stats[i][0]["converged"]
    rb = S.ResidentBatch(probs, precision="mixed-fp32")
    for rep in range(2):
        rb.solve()
        for i in range(len(probs)):
            e = rel_err(rb.flux(i), refs[i])
            _say("zero-source batch", rep, i, e)
            assert e < 1e-6
    rb.close()
