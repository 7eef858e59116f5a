"""Continuous slowing down (SN solver "BCSD", the DD scheme of order 1 in space and energy:
flux_{1,2,3}D_BFP of sn_{1,2,3}D_bfp.jl with O = 1, the energy flux 𝚽E12 carried between groups by
sn_flux.jl:166,183 and zeroed in homogeneous passes by sn_inner_pass.jl:38).

CPU pins of the oracle:
- with zero stopping powers the solve is the plain BTE solve;
- an infinite medium (one material, uniform isotropic source, every face reflective) follows a
  moment recursion derived independently here:
    phi_g (St_g + 2 S+_g - Ss_gg) = q_g + sum_{g'<g} Ss_g'g phi_g' + (S+_g + S-_g) E_g,
    E_0 = 0,  E_{g+1} = (2 phi_g - E_g) dE_{g+1} / dE_g,
  with S-_g = Sb_g / dE_g and S+_g = Sb_{g+1} / dE_g; higher moments vanish.
GPU: solves against the oracle (FP64 1e-10; mixed FP32 1e-6)."""
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


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_zero_stopping_powers_is_bte(oracle, nd):
    P = _P()
    p = P.small_problem(nd, N=5 if nd == 3 else 7, G=3, L=2, tol=1e-12, seed=81)
    ref, _ = oracle.solve(p)
    q = _csd(copy.deepcopy(p), scale=0.0)
    got, _ = oracle.solve(q)
    assert rel_err(got, ref) < 1e-14


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_infinite_medium_recursion(oracle, nd):
    P = _P()
    G = 4
    p = P.small_problem(nd, N=4, G=G, L=2, Nmat=1, tol=1e-13, seed=82, bc=[1] * (2 * nd))
    q = np.zeros_like(p.qvol)
    q[0, :] = 1.0
    q[2, :] = 0.3
    p.qvol = q
    _csd(p, seed=4, scale=0.7)
    phi, st = oracle.solve(p)
    St, Ss, Sb, dE = p.sigt[:, 0], p.sigs[0, :, :, 0], p.Sb[:, 0], p.dE
    ref = np.zeros(G)
    E = 0.0
    for g in range(G):
        Sm, Sp = Sb[g] / dE[g], Sb[g + 1] / dE[g]
        src = q[g, 0] + sum(Ss[gi, g] * ref[gi] for gi in range(g)) + (Sp + Sm) * E
        ref[g] = src / (St[g] + 2 * Sp - Ss[g, g])
        E = (2 * ref[g] - E) * (dE[g + 1] / dE[g] if g + 1 < G else 1.0)
    for g in range(G):
        assert np.allclose(phi[g, 0, :], ref[g], rtol=1e-9, atol=0), (g, phi[g, 0, :2], ref[g])
    assert np.abs(phi[:, 1:, :]).max() < 1e-9 * ref.max()
    assert all(s["converged"] for s in st)


def test_energy_flux_reaches_sourceless_groups(oracle):
    """Without scattering to carry particles down, only the slowing-down term feeds groups > 0."""
    P = _P()
    p = P.small_problem(2, N=7, G=3, L=1, tol=1e-12, seed=83)
    p.sigs = np.zeros_like(p.sigs)
    q = np.zeros_like(p.qvol)
    q[0] = p.qvol[0]
    p.qvol = q
    plain, _ = oracle.solve(p)
    assert not plain[1:].any()
    _csd(p, seed=5)
    phi, st = oracle.solve(p)
    assert np.abs(phi[1:, 0]).max() > 0 and st[1]["passes"] > 0


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(1, None), (2, None), (3, None), (2, [1, 0, 2, 2]), (3, [1, 1, 0, 0, 2, 2])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_csd_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    nd, bc = CASES[case]
    P = _P()
    p = _csd(P.small_problem(nd, N=5 if nd == 3 else 7, G=3, L=2, tol=1e-12, seed=84 + case, bc=bc), seed=case)
    ref, _ = oracle.solve(p)
    phi, st = S.compute_flux(p)
    e = rel_err(phi, ref)
    _say("csd", nd, bc, e)
    assert e < 1e-10 and all(s["converged"] for s in st)
    if bc is None:
        mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
        e = rel_err(mx, ref)
        _say("csd mixed-fp32", nd, e)
        assert e < 1e-6


@pytest.mark.gpu
def test_csd_other_accelerations_and_batch(oracle):
    from radiant_hip import solver as S
    P = _P()
    p = _csd(P.small_problem(2, N=7, G=3, L=2, tol=1e-12, seed=90), seed=9)
    ref, _ = oracle.solve(p)
    for acc in ("anderson", "bicgstab"):
        q = copy.deepcopy(p)
        q.accel = acc
        r, _ = oracle.solve(q)
        phi, _ = S.compute_flux(q)
        assert rel_err(phi, r) < 1e-10, acc
    phis, _, _ = S.solve_batch([p, p], precision="mixed-fp32")
    for ph in phis:
        assert rel_err(ph, ref) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_csd_in_lockstep_batch(oracle, precision):
    """Slowing-down problems run in the lockstep solver (the energy inflow joins the angular
    source of each c = T(0) pass, the energy outflow of T(z) is formed at the reconstruction and
    carried to the next group), mixed with BTE problems and reflective/periodic faces: within 1e-6
    of the oracle, bit-identical alone and in the mix."""
    import os
    from radiant_hip import solver as S
    P = _P()
    probs = [_csd(P.small_problem(2, N=7, G=3, L=2, tol=1e-12, seed=91), seed=11),
             _csd(P.small_problem(1, N=12, G=4, L=3, tol=1e-12, seed=92), seed=12),
             _csd(P.small_problem(3, N=5, G=3, L=2, tol=1e-12, seed=93, bc=[1, 2, 0, 1, 2, 0]), seed=13),
             P.small_problem(2, N=8, G=3, L=2, tol=1e-12, seed=94),
             _csd(P.small_problem(2, dims=(20, 6), G=3, L=2, tol=1e-12, seed=95, bc=[1, 1, 2, 2]), seed=14)]
    phis, stats, _ = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = oracle.solve(p)
        e = rel_err(phis[i], ref)
        _say("csd batch", precision, i, e)
        assert e < 1e-6 and all(s["converged"] for s in stats[i])
    alone, _, _ = S.solve_batch([probs[0]], precision=precision)
    assert np.array_equal(alone[0], phis[0])
    os.environ["RADIANT_HIP_PROFILE"] = "1"
    try:  # kernel times are recorded by lockstep solvers only
        rb = S.ResidentBatch(probs[:3], precision=precision)
        rb.solve()
        _, n, _, _, _ = rb.kernel_times()
        rb.close()
    finally:
        os.environ.pop("RADIANT_HIP_PROFILE", None)
    assert n[:, 1].sum() > 0
