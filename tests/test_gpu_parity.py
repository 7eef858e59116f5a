"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same seeded inputs.

Tolerances: pass-level FP64 1e-12 relative (the M/D operators of the two setups differ by
~1e-13 · cond(D)); full FP64 solves 1e-10 relative on the flux moments (GMRES tol 1e-12 on
both sides so the iterate error is far below the bar); mixed precision 1e-6 against FP64.
"""
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


def _random_state(rng, S, Np, Nc):
    z = rng.standard_normal(S)
    return z


@pytest.mark.parametrize("ndims", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_pass_parity(oracle, ndims, mode):
    P, S = _P(), _S()
    prob = P.small_problem(ndims, N=5 if ndims == 3 else 7)
    _say("pass parity", ndims, mode)
    h = S.ResidentProblem(prob)
    try:
        nd, np_, nps, Sl = oracle.sizes(prob)
        assert Sl == h.state_length()
        rng = np.random.default_rng(11 + ndims + 10 * mode)
        z = rng.standard_normal(Sl)
        nc = prob.Nc
        # out-of-group source: a seeded field on the moments that carry scattering
        q = np.zeros(np_ * nc)
        ang = h.d.ang
        actp = [p for p in range(np_) if ang.pl[p] <= prob.Lxs]
        for p in actp:
            q[p::np_] = rng.standard_normal(nc)
        for g in range(prob.G):
            ref = oracle.transport_pass(prob, g, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                ref = z - ref
            got = h.transport_pass(g, mode, q, z)
            e = rel_err(got, ref)
            _say("  g", g, "err", e)
            assert e < 1e-12
    finally:
        h.close()


@pytest.mark.parametrize("ndims", [1, 2, 3])
def test_solve_parity_small(oracle, ndims):
    P, S = _P(), _S()
    prob = P.small_problem(ndims, N=6 if ndims == 3 else 8, tol=1e-12)
    _say("solve parity small", ndims)
    ref, rst = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("  err", e, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10
    assert [s["niter"] for s in st] == [s["niter"] for s in rst]


@pytest.mark.parametrize("accel", ["none", "livolant", "anderson", "bicgstab"])
def test_solve_parity_accels(oracle, accel):
    P, S = _P(), _S()
    prob = P.small_problem(2, N=6, tol=1e-12, accel=accel)
    _say("accel parity", accel)
    ref, rst = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("  err", e, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10


def test_c0_parity(oracle):
    P, S = _P(), _S()
    prob = P.config0()
    prob.tol = 1e-12
    _say("C0 parity")
    ref, rst = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("  err", e)
    assert e < 1e-10
    for g in range(prob.G):
        assert rel_err(phi[g], ref[g]) < 1e-10


def test_c1_geometry_parity(oracle):
    """C1 at full grid and angular size, 3 groups (the oracle's cost scales with G)."""
    P, S = _P(), _S()
    prob = P.config_2d(G=3, seed=20260102)
    prob.tol = 1e-12
    _say("C1 (3 groups) parity")
    ref, rst = oracle.solve(prob, nthreads=8)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("  err", e, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10


def test_c2_geometry_parity(oracle):
    """C2 angular set (Carlson S12, 168 directions) on a 12^3 grid, 3 groups."""
    P, S = _P(), _S()
    prob = P.config_3d(N=12, width=1.2, G=3, seed=20260103)
    prob.tol = 1e-12
    _say("C2-angles parity")
    ref, rst = oracle.solve(prob, nthreads=8)
    phi, st = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("  err", e)
    assert e < 1e-10


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_mixed_precision_error(oracle, precision):
    P, S = _P(), _S()
    prob = P.config_2d(N=32, G=4, L=3, seed=5)
    prob.tol = 1e-10
    _say("mixed", precision)
    ref, _ = S.compute_flux(prob)
    phi, st = S.compute_flux(prob, precision=precision, state="reduced")
    e = rel_err(phi, ref)
    _say("  err", e, [(s["niter"], s["refinements"], s["passes_lo"]) for s in st])
    assert e < 1e-6


def test_reduced_state_fp64(oracle):
    P, S = _P(), _S()
    prob = P.small_problem(3, N=6, tol=1e-12)
    ref, _ = oracle.solve(prob)
    phi, st = S.compute_flux(prob, state="reduced")
    e = rel_err(phi, ref)
    _say("reduced fp64 err", e)
    assert e < 1e-10


def test_zero_source_group(oracle):
    """A group with no source anywhere returns zero flux without passes (sn_one_speed.jl:116-120)."""
    P, S = _P(), _S()
    prob = P.small_problem(2, N=6, G=3, tol=1e-12, source_group=1)
    ref, rst = oracle.solve(prob)
    phi, st = S.compute_flux(prob)
    assert np.all(phi[0] == 0.0) and st[0]["passes"] == 0 and st[0]["converged"]
    assert rel_err(phi, ref) < 1e-10


@pytest.mark.parametrize("kind", ["g", "r"])
@pytest.mark.parametrize("ndims", [1, 2, 3])
def test_alternate_sweep_kernel_parity(oracle, ndims, kind, monkeypatch):
    """The other sweep kernels (the row-scan kernel is the default): RADIANT_HIP_SWEEP=r selects the
    wave-per-direction rows/planes kernels, g the pencil kernel (fallback for grids the others do
    not cover)."""
    monkeypatch.setenv("RADIANT_HIP_SWEEP", kind)
    P, S = _P(), _S()
    prob = P.small_problem(ndims, N=5 if ndims == 3 else 9, tol=1e-12)
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("generic sweep", ndims, e)
    assert e < 1e-10


@pytest.mark.parametrize("ndims", [1, 2, 3])
def test_scan_shuffle_variant_parity(oracle, ndims, monkeypatch):
    """The ds_bpermute scan of the row-scan sweep (the DPP scan is the default), selected by
    RADIANT_HIP_SCAN=shfl; both must match the oracle."""
    monkeypatch.setenv("RADIANT_HIP_SCAN", "shfl")
    P, S = _P(), _S()
    prob = P.small_problem(ndims, N=5 if ndims == 3 else 70, tol=1e-12)
    ref, _ = oracle.solve(prob)
    phi, _ = S.compute_flux(prob)
    e = rel_err(phi, ref)
    _say("shuffle scan", ndims, e)
    assert e < 1e-10


@pytest.mark.parametrize("shape", [0, 1])
def test_scan3d_shapes_parity(oracle, shape, monkeypatch):
    """Both 3D row-scan shapes (pencils per item, prefetch depth), selected by RADIANT_HIP_SCAN3D,
    match the oracle: on 5^3, where some waves own no pencil of most diagonals, on 12^3, where a
    wave's items wrap over several diagonals, and on long-x grids with 2, 4 and 8 cells per lane;
    the mixed FP32 solve stays within 1e-6 of the oracle."""
    monkeypatch.setenv("RADIANT_HIP_SCAN3D", str(shape))
    P, S = _P(), _S()
    probs = [P.small_problem(3, N=5, tol=1e-12), P.config_3d(N=12, width=1.2, G=2, quad_order=4, seed=17)]
    # 2, 4 and 8 cells per lane along x
    probs += [P.small_problem(3, dims=d, G=2, tol=1e-12, seed=s) for d, s in (((70, 4, 3), 21), ((130, 3, 4), 22),
                                                                               ((300, 3, 3), 23))]
    for prob in probs:
        ref, _ = oracle.solve(prob)
        phi, _ = S.compute_flux(prob)
        e = rel_err(phi, ref)
        _say("scan3d shape", shape, prob.Nx, e)
        assert e < 1e-10
        mx, _ = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
        e = rel_err(mx, ref)
        _say("scan3d shape", shape, prob.Nx, "mixed-fp32", e)
        assert e < 1e-6
