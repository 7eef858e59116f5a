"""The benchmark workloads at their own sizes on the GPU, through the C ABI.

C2 -- 3D 64^3, 40 groups, Carlson 12 (168 directions), P3, 3 random materials, void faces:
  * the plain FP64 solve converges in every group; for a set of groups the converged flux is a
    fixed point of the FP64 transport pass T (rt_problem_pass on the faithful state, the pass the
    oracle pins at <= 1e-12), and the particle balance of the group holds exactly:
        sum_c V_c (Σt φ0 - Σs_gg,0 φ0 - sum_{g'<g} Σs_g'g,0 φ0_g' - q_g) + leakage_g = 0,
    the leakage taken from the exit-face half-range moments of T (ψ_exit = Mn_surf · moments,
    leakage = sum over faces and exit directions of w_n |Ω_n . normal| ψ_n times the face area);
    this is the DD balance summed over cells and directions (interior faces cancel, and
    sum_n w_n M[n, p] = δ_p0 because D[0, :] = w and D M = I);
  * mixed FP32 and mixed BF16 (FP32 accumulation, FP64 Arnoldi and refinement) are within 1e-6 of
    the FP64 flux moments;
  * FP64 parity with the oracle on the same 3D path (Carlson 12, P3, 40 groups, 3 random
    materials) at 20^3 cells, the largest grid the oracle solves within the test's time on the
    host cores (OpenMP over directions): <= 1e-10 and the same GMRES iteration counts.
C3 -- the 256-problem uneven 2D batch (N 32..160, G 8..32, P1..P5, GLC 8/10) in one lockstep
  mixed-precision batch: every group of every problem converged, a seeded sample within 1e-6 of
  its FP64 solve, and each sampled problem bit-identical solved alone and in the mix.
C4 -- 32 requests of the 3D Carlson stream (N 32..96, G 10..40, Carlson 8/10/12, P3) submitted at
  Poisson 50/s arrival times to rt_server: every request served, a sample within 1e-6 of its FP64
  solve and bit-identical to the same request solved alone.
Reference path: sn_flux.jl:150-224, sn_one_speed.jl:81-196, sn_sweep_3D.jl:57-240,
sn_3D_bte.jl:42-153."""
import os
import sys
import time

import numpy as np
import pytest

from util import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mods():
    from radiant_hip import problems, solver
    return problems, solver


@pytest.fixture(scope="module")
def c2_fp64():
    problems, solver = _mods()
    p = problems.config2()
    _say("C2 FP64 faithful solve: 64^3, G=40, Carlson 12, P3")
    rp = solver.ResidentProblem(p, precision="fp64", state="faithful")
    t0 = time.time()
    stats, info = rp.solve()
    _say(f"C2 FP64 solve {time.time() - t0:.2f} s, passes {info.passes}")
    phi = rp.flux()
    yield p, rp, phi, stats
    rp.close()


def _qlout(p, ang, phi, g):
    """Out-of-group scattering (scattering_source.jl:59-68) + volume source (sn_flux.jl:179) of
    group g on the moments with a scattering source, [Np, Nc]; zero on the other moments."""
    Np, Nc = phi.shape[1], phi.shape[2]
    q = np.zeros((Np, Nc))
    mat = p.mat.reshape(-1, order="F") - 1
    for k in range(Np):
        l = int(ang.pl[k])
        if l > p.Lxs:
            continue
        for gi in range(g):
            s = p.sigs[:, gi, g, l][mat]
            if s.any():
                q[k] += s * phi[gi, k]
    q[0] += p.qvol[g]
    return q


def test_c2_fp64_converged_fixed_point_and_balance(c2_fp64):
    p, rp, phi, stats = c2_fp64
    assert all(s["converged"] for s in stats), [s["resid"] for s in stats]
    ang = rp.d.ang
    Np, Nc, Nps = ang.Np, p.Nc, ang.Nps
    S = rp.state_length()
    V = p.dx[0] * p.dy[0] * p.dz[0]
    mat = p.mat.reshape(-1, order="F") - 1
    cosines = (ang.mu, ang.eta, ang.xi)
    areas = (p.dy[0] * p.dz[0], p.dx[0] * p.dz[0], p.dx[0] * p.dy[0])
    npos = (p.Ny * p.Nz, p.Nx * p.Nz, p.Nx * p.Ny)
    # leakage weights per axis and exit slot: slot 1 (+ face) holds directions with a positive
    # cosine, in the basis of the - face they enter by (face 2a), slot 0 the negative ones (2a+1)
    cw = {}
    for a in range(3):
        for slot, f in ((1, 2 * a), (0, 2 * a + 1)):
            nps, Ms, Ds, n2p = ang.faces[f]
            idx = np.nonzero(n2p)[0]
            rows = n2p[idx] - 1
            cw[a, slot] = (ang.w[idx] * np.abs(cosines[a][idx])) @ Ms[rows, :]
    for g in (0, 1, 2, 20, p.G - 1):
        q = _qlout(p, ang, phi, g)
        z = np.zeros(S)
        z[:Np * Nc] = np.asfortranarray(phi[g]).ravel(order="F")
        tz = rp.transport_pass(g, 0, q.ravel(order="F"), z)
        fp = rel_err(tz[:Np * Nc], z[:Np * Nc])
        off = Np * Nc
        leak = 0.0
        for a in range(3):
            b = tz[off:off + Nps * 2 * npos[a]].reshape((Nps, 2, npos[a]), order="F")
            off += Nps * 2 * npos[a]
            for slot in (0, 1):
                leak += areas[a] * float(cw[a, slot] @ b[:, slot, :].sum(axis=1))
        assert off == S
        st = p.sigt[g][mat]
        sgg = p.sigs[:, g, g, 0][mat]
        src = p.qvol[g] + sum(p.sigs[:, gi, g, 0][mat] * phi[gi, 0] for gi in range(g))
        removal = V * float(np.sum((st - sgg) * phi[g, 0]))
        total = V * float(np.sum(src))
        bal = abs(removal + leak - total) / total
        _say(f"C2 group {g}: fixed point {fp:.2e}, balance {bal:.2e} (leakage {leak / total:.3f} of the source)")
        assert fp < 1e-8 and bal < 1e-8 and leak > 0


@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_c2_mixed_precision_within_1e6(c2_fp64, precision):
    problems, solver = _mods()
    p, _, phi, _ = c2_fp64
    rp = solver.ResidentProblem(p, precision=precision, state="reduced")
    t0 = time.time()
    stats, info = rp.solve()
    dt = time.time() - t0
    mx = rp.flux()
    rp.close()
    e = rel_err(mx, phi)
    eg = max(rel_err(mx[g], phi[g]) for g in range(p.G) if np.any(phi[g]))
    _say(f"C2 {precision}: {dt:.2f} s, passes {info.passes} ({info.passes_lo} low), err {e:.2e}, worst group {eg:.2e}")
    assert all(s["converged"] for s in stats)
    assert e < 1e-6 and eg < 1e-6


def test_c2_path_oracle_parity_20cubed(oracle):
    """The C2 path (3D, Carlson 12, P3, 40 groups, 3 random materials) against the oracle at
    20^3 cells, FP64 faithful, GMRES(30) tol 1e-10."""
    problems, solver = _mods()
    p = problems.config_3d(N=20, width=2.0, G=40, L=3, quad_order=12, Nmat=3, seed=20260103)
    nth = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    ref, rst = oracle.solve(p, nthreads=nth)
    _say(f"oracle 20^3 x 40 groups, Carlson 12: {time.time() - t0:.1f} s on {nth} threads")
    phi, st = solver.compute_flux(p)
    e = rel_err(phi, ref)
    _say(f"C2 path at 20^3: rel err vs oracle {e:.2e}")
    assert e < 1e-10
    assert [s["niter"] for s in st] == [s["niter"] for s in rst]


C3_SAMPLE = (0, 77, 191)


def test_c3_256_uneven_batch():
    problems, solver = _mods()
    probs = problems.config3(256, quad_orders=(8, 10))
    _say("C3: 256 problems, lockstep mixed-fp32")
    t0 = time.time()
    rb = solver.ResidentBatch(probs, precision="mixed-fp32")
    _say(f"C3 upload {time.time() - t0:.1f} s")
    t0 = time.time()
    stats, info = rb.solve()
    _say(f"C3 solve {time.time() - t0:.2f} s, passes {info.passes}")
    bad = [(i, g) for i, st in enumerate(stats) for g, s in enumerate(st) if not s["converged"]]
    assert not bad, bad[:10]
    mix = {i: rb.flux(i) for i in C3_SAMPLE}
    rb.close()
    for i in C3_SAMPLE:
        ref, _ = solver.compute_flux(probs[i])
        e = rel_err(mix[i], ref)
        alone, _, _ = solver.solve_batch([probs[i]], precision="mixed-fp32")
        same = bool(np.array_equal(alone[0], mix[i]))
        _say(f"C3[{i}] N={probs[i].Nx} G={probs[i].G} GLC-{probs[i].quad_order}: err vs FP64 {e:.2e}, "
             f"alone == in the mix: {same}")
        assert e < 1e-6 and same


C4_SAMPLE = (2, 5, 22)


def test_c4_served_stream_32():
    problems, solver = _mods()
    n = 32
    t_arr = problems.poisson_arrivals(n, rate=50.0)
    srv = solver.Server(hbm_budget_bytes=200 << 30, max_active=8, precision="mixed-fp32")
    ids, got, lat = [], {}, {}
    try:
        t0 = time.perf_counter()
        for i in range(n):
            p = problems.config4_request(i)
            wait = t_arr[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(p))
        _say(f"C4: {n} requests submitted in {time.perf_counter() - t0:.2f} s")
        while len(lat) < n and time.perf_counter() - t0 < 100:
            for k, rid in enumerate(ids):
                if k in lat:
                    continue
                r = srv.poll(rid)
                if r is not None:
                    lat[k] = r[2]
                    assert all(s["converged"] for s in r[1]), k
                    if k in C4_SAMPLE:
                        got[k] = r[0]
                    del r
            time.sleep(0.005)
        _say(f"C4: {len(lat)} served in {time.perf_counter() - t0:.1f} s, latency p50 "
             f"{np.percentile(list(lat.values()), 50):.0f} ms p99 {np.percentile(list(lat.values()), 99):.0f} ms")
    finally:
        srv.close()
    assert len(lat) == n
    for k in C4_SAMPLE:
        p = problems.config4_request(k)
        ref, _ = solver.compute_flux(p)
        e = rel_err(got[k], ref)
        alone = solver.serve_alone(p)  # the same request alone in a server
        same = bool(np.array_equal(alone, got[k]))
        _say(f"C4[{k}] N={p.Nx} G={p.G}: err vs FP64 {e:.2e}, served == alone: {same}")
        assert e < 1e-6 and same
