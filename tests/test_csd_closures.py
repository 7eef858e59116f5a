"""Slowing-down closures of any order in space and energy (SN.set_scheme("E", ...) with the
slowing-down solver types; flux_{1,2,3}D_BFP of sn_1D_bfp.jl:44-125, sn_2D_bfp.jl:43-154 and
sn_3D_bfp.jl:51-191; the in-group stopping-power moments S = [(S⁻+S⁺)/2, (S⁻-S⁺)/(2√3)] / ΔE of
sn_flux.jl:92-96; 𝒲 = ½∫P_a P_b P_c of scheme_weights.jl:411-412 and integrals.jl:254).

CPU pins of the oracle's cell (ψ = Σ C_e C_a C_b C_c ψ_eabc P_e(u) P_a(x) P_b(y) P_c(z) on the
cell and the group, C_k = √(2k+1); u = +1 at the group's upper energy, where particles enter):
- the slowing-down cell reproduces a polynomial solution of
      St ψ + Σ_axes Ω_a (2/Δ_a) ∂ψ/∂x_a − 2 ∂(S(u) ψ)/∂u = Q,   S(u) = (S⁻+S⁺)/2 + u (S⁻−S⁺)/2
  (S⁻, S⁺ per unit group width) exactly, cell moments and outgoing face moments on the x, y, z
  and energy faces: DD up to the closure order on every axis, DG one degree less, with cross
  terms under full coupling and single-axis terms without; not beyond, for DG.  A constant and a
  linear stopping power are both covered, so a swapped S⁻/S⁺, a wrong slope moment or a wrong 𝒲
  index fails;
- the zeroth moments obey the balance St ψ_0 + Σ_a h_a (f_out − f_in)_0 + S⁺ E_out,0 − S⁻ E_in,0
  = Q_0;
- 𝒲 against Gauss–Legendre quadrature of P_a P_b P_c;
- DD of order 1 everywhere is the order-1 slowing-down cell.
Energy closures of order 3 and above are refused: with them the reference's in-group
stopping-power term is not consistent with the continuous operator (its moment weights differ by
C_k from the weak form for k ≥ 1), so no exactness property pins them.
Solve-level: with zero stopping powers every closure gives the BTE solve of the same spatial
closure, on the energy moment 0, with vanishing energy slopes; an all-reflective infinite medium
has no spatial slopes.
GPU: passes (1e-12) and solves (1e-10) against the oracle."""
import copy
import itertools
import sys

import numpy as np
import pytest
from numpy.polynomial import legendre as Lg

from util import rel_err


def _P():
    from radiant_hip import problems
    return problems


def test_legendre_triple_against_quadrature(oracle):
    x, w = Lg.leggauss(12)
    P = [Lg.legval(x, np.eye(8)[k]) for k in range(8)]
    for a, b, c in itertools.product(range(5), repeat=3):
        want = 0.5 * float(np.sum(w * P[a] * P[b] * P[c]))
        assert abs(oracle.legendre_triple(a, b, c) - want) < 1e-15, (a, b, c)


def _index(m, O, fc, skip=-1):
    """Cell (skip = -1) or face index of the multi-index m over axes (E, x, y, z)."""
    r, stride, off = 0, 1, 0
    for k in range(4):
        if k == skip:
            continue
        if fc:
            r += m[k] * stride
            stride *= O[k]
        else:
            if m[k] > 0:
                r = m[k] + off
            off += O[k] - 1
    return r


def _nface(O, fc, skip):
    o = [O[k] for k in range(4) if k != skip]
    return int(np.prod(o)) if fc else 1 + sum(v - 1 for v in o)


def _cell_case(oracle, schemes, O, fc, deg, nd, om, Sm, Sp, rng):
    """Manufactured polynomial ψ(u, x, y, z) of degrees deg; returns (max error, balance)."""
    St = 0.9
    h = np.array([0.37, 0.51, 0.43])
    h[nd:] = 0.0
    om = np.array(om, float)
    om[nd:] = 0.0
    A = [max(O[k], deg[k] + 1) for k in range(4)]
    c = np.zeros(A)
    for m in itertools.product(*(range(d + 1) for d in deg)):
        if fc or sum(v > 0 for v in m) <= 1:
            c[m] = rng.standard_normal()
    sig0, sig1 = (Sm + Sp) / 2, (Sm - Sp) / 2          # S(u) = sig0 + sig1 u
    Qc = St * c
    for ax in range(1, 1 + nd):
        dc = Lg.legder(c, axis=ax)
        sl = [slice(None)] * 4
        sl[ax] = slice(0, dc.shape[ax])
        Qc[tuple(sl)] += om[ax - 1] * 2 / h[ax - 1] * dc
    cp = np.concatenate([c, np.zeros((1,) + c.shape[1:])], axis=0)      # room for u * ψ
    uc = np.zeros_like(cp)                               # u P_k = ((k+1) P_(k+1) + k P_(k-1)) / (2k+1)
    for k in range(c.shape[0]):
        uc[k + 1] += (k + 1) / (2 * k + 1) * cp[k]
        if k > 0:
            uc[k - 1] += k / (2 * k + 1) * cp[k]
    Sc = sig0 * cp + sig1 * uc
    Qc = Qc - 2 * Lg.legder(Sc, axis=0)                  # d/du lowers the degree back to that of c
    Cs = [np.sqrt(2 * np.arange(A[k]) + 1) for k in range(4)]
    norm = np.einsum("e,a,b,d->eabd", *Cs)
    Qm, cm = Qc / norm, c / norm
    nm = int(np.prod(O)) if fc else sum(O) - 3
    Qn, want = np.zeros(nm), np.zeros(nm)
    for m in itertools.product(*(range(o) for o in O)):
        if fc or sum(v > 0 for v in m) <= 1:
            j = _index(m, O, fc)
            Qn[j], want[j] = Qm[m], cm[m]

    def face(ax, side):
        v = Lg.legval(side, np.moveaxis(c, ax, 0))      # polynomial over the other axes
        others = [k for k in range(4) if k != ax]
        v = v / np.einsum("i,j,k->ijk", *(Cs[k] for k in others))
        f = np.zeros(_nface(O, fc, ax))
        for m in itertools.product(*(range(O[k]) for k in others)):
            if fc or sum(t > 0 for t in m) <= 1:
                full = [0] * 4
                for k, t in zip(others, m):
                    full[k] = t
                f[_index(full, O, fc, ax)] = v[m]
        return f

    s = np.sign(om)
    s_in = [1.0] + [-s[a] if a < nd else -1.0 for a in range(3)]    # energy enters at u = +1
    s_out = [-1.0] + [s[a] if a < nd else 1.0 for a in range(3)]
    fin = [face(k, s_in[k]) for k in range(4)]
    fout = [face(k, s_out[k]) for k in range(4)]
    psi, fe, fx, fy, fz = oracle.cell_bfp(schemes, O, fc, nd, om, h, St, Sm, Sp, Qn, *fin)
    got = (fe, fx, fy, fz)
    err = np.abs(psi - want).max()
    for k in range(1 + nd):
        err = max(err, np.abs(got[k] - fout[k]).max())
    bal = St * psi[0] + Sp * fe[0] - Sm * fin[0][0] - Qn[0]
    for a in range(nd):
        bal += abs(om[a]) / h[a] * (got[a + 1][0] - fin[a + 1][0])
    return err, abs(bal)


CELLS = [  # (ndims, orders (E, x, y, z))
    (1, (2, 1, 1, 1)), (1, (1, 2, 1, 1)), (1, (2, 2, 1, 1)), (1, (2, 3, 1, 1)), (1, (1, 3, 1, 1)),
    (2, (2, 1, 1, 1)), (2, (2, 2, 1, 1)), (2, (1, 2, 2, 1)), (2, (2, 2, 2, 1)),
    (3, (2, 1, 1, 1)), (3, (1, 2, 1, 2)), (3, (2, 2, 2, 1)), (3, (2, 2, 2, 2)),
]


@pytest.mark.parametrize("scheme", ["DD", "DG"])
@pytest.mark.parametrize("cell", CELLS)
@pytest.mark.parametrize("fc", [True, False])
def test_cell_polynomial_exactness_and_balance(oracle, scheme, cell, fc):
    nd, O = cell
    rng = np.random.default_rng(sum(O) + 10 * nd)
    top = O if scheme == "DD" else tuple(o - 1 for o in O)
    for om, (Sm, Sp) in [((0.6, 0.3, 0.5), (0.8, 0.8)), ((-0.5, 0.7, -0.2), (1.3, 0.4)),
                         ((0.2, -0.8, -0.4), (0.35, 0.9))]:
        for deg in [tuple(o - 1 for o in O), O]:
            deg = tuple(d if k <= nd else 0 for k, d in enumerate(deg))
            err, bal = _cell_case(oracle, (scheme,) * 4, O, fc, deg, nd, om, Sm, Sp, rng)
            assert bal < 1e-12, (scheme, cell, fc, deg, bal)
            if all(dg <= t for dg, t in zip(deg, top)):
                assert err < 1e-12, (scheme, cell, fc, deg, om, err)
            else:
                assert err > 1e-6, (scheme, cell, fc, deg, err)


def test_cell_mixed_closures(oracle):
    """DG in energy with DD in space and the reverse: exact to the degree of each axis' closure."""
    rng = np.random.default_rng(5)
    for schemes, O, deg in [(("DG", "DD", "DD", "DD"), (2, 2, 1, 1), (1, 2, 0, 0)),
                            (("DD", "DG", "DD", "DD"), (2, 2, 2, 1), (2, 1, 2, 0))]:
        nd = 1 if O[2] == 1 else 2
        err, bal = _cell_case(oracle, schemes, O, True, deg, nd, (0.5, -0.6, 0.0), 1.1, 0.6, rng)
        assert err < 1e-12 and bal < 1e-12, (schemes, err)


def test_cell_dd1_is_the_order1_slowing_down_cell(oracle):
    for nd in (1, 2, 3):
        om = np.array([0.6, -0.3, 0.45])[:nd].tolist() + [0.0] * (3 - nd)
        h = [0.3, 0.7, 0.5][:nd] + [0.0] * (3 - nd)
        St, Sm, Sp, q = 1.1, 0.7, 0.4, 0.9
        fin = [0.25, 0.4, 0.1, 0.3]
        psi, fe, fx, fy, fz = oracle.cell_bfp(("DD",) * 4, (1, 1, 1, 1), True, nd, om, h, St, Sm, Sp, [q],
                                              *([v] for v in fin))
        hh = [abs(om[a]) / h[a] for a in range(nd)]
        want = (q + sum(2 * hh[a] * fin[a + 1] for a in range(nd)) + (Sp + Sm) * fin[0]) / (St + 2 * sum(hh) + 2 * Sp)
        assert abs(psi[0] - want) < 1e-15
        assert abs(fe[0] - (2 * want - fin[0])) < 1e-15
        for a, f in enumerate((fx, fy, fz)[:nd]):
            assert abs(f[0] - (2 * want - fin[a + 1])) < 1e-15


def _bfp(p, orders, schemes=("DD", "DD", "DD", "DD"), fc=True, seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    p.csd = True
    p.Sb = scale * rng.uniform(0.5, 2.0, (p.G + 1, p.Nmat))
    p.dE = rng.uniform(0.5, 1.5, p.G)
    p.scheme_e, p.scheme, p.scheme_y, p.scheme_z = schemes
    p.scheme_order_e, p.scheme_order, p.scheme_order_y, p.scheme_order_z = orders
    p.full_coupling = fc
    return p


@pytest.mark.parametrize("nd,orders,fc", [(1, (2, 2, 1, 1), True), (1, (2, 1, 1, 1), True), (2, (2, 2, 1, 1), False),
                                          (2, (2, 1, 2, 1), True), (3, (2, 2, 1, 1), True)])
def test_zero_stopping_powers_is_bte_of_the_spatial_closure(oracle, nd, orders, fc):
    """S = 0: every energy moment solves the BTE cell of the same spatial closure; energy moment 0
    carries the BTE flux (its source is the volume source), the energy slopes vanish."""
    P = _P()
    p = P.small_problem(nd, N=4 if nd == 3 else 6, G=3, L=1, tol=1e-12, seed=84)
    bte = copy.deepcopy(p)
    bte.scheme, bte.scheme_order = "DD", orders[1]
    bte.scheme_y, bte.scheme_order_y = "DD", orders[2]
    bte.scheme_z, bte.scheme_order_z = "DD", orders[3]
    bte.full_coupling = fc
    ref, _ = oracle.solve(bte)
    q = _bfp(copy.deepcopy(p), orders, fc=fc, scale=0.0)
    got, st = oracle.solve(q)
    assert all(s["converged"] for s in st)
    Ks, K = oracle.spatial_moments(bte), oracle.spatial_moments(q)
    Nc = p.Nx * p.Ny * p.Nz
    gv = got.reshape(got.shape[0], got.shape[1], K, Nc, order="F")
    rv = ref.reshape(ref.shape[0], ref.shape[1], Ks, Nc, order="F")
    O = orders
    for m in itertools.product(*(range(o) for o in O)):
        if not (fc or sum(v > 0 for v in m) <= 1):
            continue
        j = _index(m, O, fc)
        if m[0] == 0:
            js = _index((0,) + m[1:], (1,) + O[1:], fc)
            assert rel_err(gv[:, :, j], rv[:, :, js]) < 1e-12, m
        else:
            assert np.abs(gv[:, :, j]).max() < 1e-12 * np.abs(rv).max(), m


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_infinite_medium_has_no_spatial_slopes(oracle, nd):
    P = _P()
    base = P.small_problem(nd, N=3, G=3, L=1, Nmat=1, tol=1e-13, seed=85, bc=[1] * (2 * nd))
    base.qvol = np.ones_like(base.qvol)
    orders = (2,) + tuple(2 if a < nd else 1 for a in range(3))
    p = _bfp(copy.deepcopy(base), orders, schemes=("DG", "DD", "DG", "DD"), seed=6, scale=0.5)
    phi, st = oracle.solve(p)
    assert all(s["converged"] for s in st)
    K = oracle.spatial_moments(p)
    for m in itertools.product(*(range(o) for o in orders)):
        j = _index(m, orders, True)
        v = phi[:, :, j::K]
        assert np.abs(v - v[:, :, :1]).max() < 1e-10 * np.abs(phi).max()   # uniform over cells
        if any(m[1:]):
            assert np.abs(v).max() < 1e-10 * np.abs(phi).max(), m
    assert np.abs(phi[:, :, 1::K]).max() > 1e-6 * np.abs(phi).max()   # the energy slope is not zero


def test_refusals(oracle):
    P = _P()
    p = P.small_problem(1, N=4, G=2, L=1, seed=86)
    q = copy.deepcopy(p)
    q.scheme_order_e = 2
    with pytest.raises(RuntimeError, match="slowing-down"):
        oracle.solve(q)
    q = _bfp(copy.deepcopy(p), (3, 1, 1, 1))
    with pytest.raises(RuntimeError, match="above 2"):
        oracle.solve(q)
    q = _bfp(copy.deepcopy(p), (2, 2, 1, 1), schemes=("DD", "DG-", "DD", "DD"))
    with pytest.raises(RuntimeError, match="Closure relation"):
        oracle.solve(q)


# ------------------------------------------------------------------------------------------
# GPU parity (sweep_bfp_kernel, bfp_out_kernel, the cutoff flux of the general cell)
# ------------------------------------------------------------------------------------------

def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (ndims, orders (E, x, y, z), schemes (E, x, y, z), full coupling, boundary conditions)
GPU_CASES = [
    (1, (2, 2, 1, 1), ("DG", "DG", "DD", "DD"), True, None),
    (1, (2, 1, 1, 1), ("DG", "DD", "DD", "DD"), True, [1, 2]),
    (1, (1, 3, 1, 1), ("DD", "DD", "DD", "DD"), True, None),
    (1, (2, 3, 1, 1), ("DD", "DG", "DD", "DD"), False, [2, 2]),
    (1, (2, 4, 1, 1), ("DG", "DD", "DD", "DD"), True, None),
    (2, (2, 2, 1, 1), ("DG", "DD", "DD", "DD"), True, [1, 0, 2, 2]),
    (2, (2, 2, 2, 1), ("DG", "DG", "DG", "DD"), True, None),
    (2, (2, 2, 2, 1), ("DD", "DD", "DD", "DD"), False, [1, 1, 1, 1]),
    (3, (2, 2, 1, 1), ("DG", "DD", "DD", "DD"), True, None),
    (3, (1, 2, 2, 2), ("DD", "DG", "DG", "DG"), True, [1, 1, 0, 0, 2, 2]),
    (3, (2, 2, 2, 2), ("DG", "DD", "DD", "DD"), False, None),
]


def _gpu_problem(case):
    nd, orders, schemes, fc, bc = GPU_CASES[case]
    P = _P()
    dims = {1: (9,), 2: (5, 4), 3: (3, 4, 3)}[nd]
    p = P.small_problem(nd, dims=dims if nd > 1 else None, N=dims[0], G=3, L=2, tol=1e-12, seed=120 + case, bc=bc)
    return _bfp(p, orders, schemes=schemes, fc=fc, seed=130 + case, scale=0.8)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(GPU_CASES)))
def test_bfp_closure_gpu_parity(oracle, case):
    from radiant_hip import solver as S
    p = _gpu_problem(case)
    _say("bfp closure", GPU_CASES[case])
    ref, rcut, rst = oracle.solve_cutoff(p)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd_, np_, nps, Sl = oracle.sizes(p)
        assert Sl == h.state_length()
        rng = np.random.default_rng(600 + case)
        z = rng.standard_normal(Sl)
        q = rng.standard_normal(np_ * p.Nc * K)
        q.reshape(-1, np_)[:, [i for i in range(np_) if h.d.ang.pl[i] > p.Lxs]] = 0.0
        for mode in (0, 1, 2):
            want = oracle.transport_pass(p, 0, 0 if mode == 0 else 1, z, q)
            if mode == 2:
                want = z - want
            e = rel_err(h.transport_pass(0, mode, q, z), want)
            _say("  pass mode", mode, e)
            assert e < 1e-12
    finally:
        h.close()
    phi, cut, st = S.compute_flux_cutoff(p)
    e, ec = rel_err(phi, ref), rel_err(cut, rcut)
    _say("  solve", e, "cutoff", ec, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10 and ec < 1e-10 and all(s["converged"] for s in st)
    assert np.abs(rcut).max() > 0
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    e = rel_err(mx, ref)
    _say("  mixed-fp32", e)
    assert e < 1e-6


@pytest.mark.gpu
def test_bfp_closure_fokker_planck_and_accelerations(oracle):
    """BFP (Galerkin ℳ) with DG 2 in energy and space, through GMRES, Anderson and BiCGStab."""
    from radiant_hip import solver as S
    P = _P()
    p = _bfp(P.small_problem(1, N=10, G=3, L=3, tol=1e-12, seed=140), (2, 2, 1, 1), schemes=("DG", "DG", "DD", "DD"),
             seed=141)
    rng = np.random.default_rng(142)
    p.solver_type = "BFP"
    p.T = rng.uniform(0.2, 1.0, (p.G, p.Nmat))
    p.fokker_planck = "galerkin"
    for acc in ("gmres", "anderson", "bicgstab"):
        q = copy.deepcopy(p)
        q.accel = acc
        ref, _ = oracle.solve(q)
        phi, st = S.compute_flux(q)
        e = rel_err(phi, ref)
        _say("bfp fp", acc, e)
        assert e < 1e-10 and all(s["converged"] for s in st)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed-fp32", "mixed-bf16"])
def test_bfp_closure_in_lockstep_batch(oracle, precision):
    """General slowing-down cells in the lockstep solver (their own sweep launch per round, the
    energy outflow of the sweep-free reconstruction from the closure), mixed with an order-1
    slowing-down problem and a BTE problem: within 1e-6 of the oracle, bit-identical alone and in
    the mix, and through the server."""
    from radiant_hip import solver as S
    P = _P()
    probs = [_gpu_problem(0), _gpu_problem(5), _gpu_problem(8),
             _bfp(P.small_problem(2, N=6, G=3, L=2, tol=1e-12, seed=150), (1, 1, 1, 1), seed=151),
             P.small_problem(2, N=7, G=3, L=2, tol=1e-12, seed=152)]
    phis, stats, _ = S.solve_batch(probs, precision=precision)
    for i, p in enumerate(probs):
        ref, _ = oracle.solve(p)
        e = rel_err(phis[i], ref)
        _say("bfp batch", precision, i, e)
        assert e < 1e-6 and all(s["converged"] for s in stats[i])
    alone, _, _ = S.solve_batch([probs[1]], precision=precision)
    assert np.array_equal(alone[0], phis[1])
    served = S.serve_alone(probs[0], precision=precision)
    assert rel_err(served, phis[0]) < 1e-6


@pytest.mark.gpu
def test_bfp_faces_in_device_memory(oracle):
    """3D slowing-down cells whose face moments exceed the LDS (72 x 72 x 2, energy DG 2 and
    spatial 2 x 2 x 1 with full coupling: 341 KiB in FP64) keep them in device memory: pass, FP64
    solve with its cutoff flux, and mixed FP32 against the oracle."""
    import os
    from radiant_hip import solver as S
    P = _P()
    p = _bfp(P.small_problem(3, dims=(72, 72, 2), G=2, L=1, tol=1e-12, seed=170, bc=[1, 0, 2, 2, 0, 1]),
             (2, 2, 2, 1), schemes=("DG", "DD", "DG", "DD"), seed=171, scale=0.6)
    p.qvol = np.zeros_like(p.qvol)
    p.qvol[0, : p.Nx * 30] = 1.0   # a source in the first 30 rows of both layers
    assert 8 * (72 * 2 * 4 + 72 * 2 * 4 + 72 * 72 * 8) > 160 * 1024
    nth = max(1, min(16, os.cpu_count() or 1))
    ref, rcut, rst = oracle.solve_cutoff(p, nthreads=nth)
    assert all(s["converged"] for s in rst)
    K = oracle.spatial_moments(p)
    h = S.ResidentProblem(p)
    try:
        nd, np_, nps, Sl = oracle.sizes(p)
        rng = np.random.default_rng(172)
        z = rng.standard_normal(Sl)
        q = rng.standard_normal(np_ * p.Nc * K)
        q.reshape(-1, np_)[:, [i for i in range(np_) if h.d.ang.pl[i] > p.Lxs]] = 0.0
        want = oracle.transport_pass(p, 0, 1, z, q, nthreads=nth)
        e = rel_err(h.transport_pass(0, 1, q, z), want)
        _say("bfp faces in HBM: pass", e)
        assert e < 1e-12
    finally:
        h.close()
    phi, cut, st = S.compute_flux_cutoff(p)
    e, ec = rel_err(phi, ref), rel_err(cut, rcut)
    mx, _ = S.compute_flux(p, precision="mixed-fp32", state="reduced")
    em = rel_err(mx, ref)
    _say("bfp faces in HBM: solve", e, "cutoff", ec, "mixed-fp32", em)
    assert e < 1e-10 and ec < 1e-10 and em < 1e-6 and all(s["converged"] for s in st)
