"""ctypes wrapper of the CPU oracle (oracle/radiant_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by the tests, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, never by the product path.  The oracle restates
Radiant.jl's ``compute_flux(::SN)`` (src/particle_transport/sn_flux.jl:22-234) in FP64 C++.

A problem is any object with the attributes of ``radiant_hip.problems.Problem`` (Radiant
layouts: Σtot[G,Nmat], Σs[Nmat,G,G,L+1], mat[Nx,Ny,Nz] 1-based, column-major).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libradiant_oracle.so")
_lib = None


class OracleProblem(C.Structure):
    _fields_ = [
        ("ndims", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("dx", C.POINTER(C.c_double)), ("dy", C.POINTER(C.c_double)), ("dz", C.POINTER(C.c_double)),
        ("mat", C.POINTER(C.c_int64)), ("Nmat", C.c_int32), ("G", C.c_int32),
        ("sigt", C.POINTER(C.c_double)), ("sigs", C.POINTER(C.c_double)), ("Lxs", C.c_int32),
        ("qvol", C.POINTER(C.c_double)), ("bc", C.POINTER(C.c_int32)),
        ("quad_type", C.c_char_p), ("quad_order", C.c_int32), ("qdims", C.c_int32),
        ("legendre_order", C.c_int32), ("sn_type", C.c_char_p), ("accel", C.c_char_p),
        ("restart", C.c_int32), ("depth", C.c_int32), ("maxit", C.c_int32), ("tol", C.c_double),
        ("nthreads", C.c_int32), ("verbose", C.c_int32),
        ("ssrc", C.POINTER(C.c_double) * 6), ("Np_source", C.c_int32),
        ("csd", C.c_int32), ("Sb", C.POINTER(C.c_double)), ("dE", C.POINTER(C.c_double)),
        ("solver_type", C.c_int32), ("T", C.POINTER(C.c_double)), ("siga", C.POINTER(C.c_double)),
        ("fokker_planck", C.c_char_p), ("scheme", C.c_char_p), ("scheme_order", C.c_int32),
        ("scheme_y", C.c_char_p), ("scheme_order_y", C.c_int32), ("partial_coupling", C.c_int32),
        ("scheme_z", C.c_char_p), ("scheme_order_z", C.c_int32),
        ("quad_mu", C.POINTER(C.c_double)), ("quad_eta", C.POINTER(C.c_double)),
        ("quad_xi", C.POINTER(C.c_double)), ("quad_w", C.POINTER(C.c_double)), ("quad_n", C.c_int32),
        ("scheme_e", C.c_char_p), ("scheme_order_e", C.c_int32),
        ("em", C.c_int32), ("em_B", C.c_double * 3), ("em_charge", C.c_double), ("Eb", C.POINTER(C.c_double)),
    ]

# SN.jl get_solver_type
SOLVER_TYPES = {"BTE": 1, "BFP": 2, "BCSD": 3, "FP": 4, "CSD": 5, "BFP-EF": 6}


class OracleStats(C.Structure):
    _fields_ = [("niter", C.c_int64), ("resid", C.c_double), ("rho", C.c_double),
                ("converged", C.c_int32), ("passes", C.c_int32)]


def build():
    """Compile the oracle library (g++); called by ``__graft_entry__.build()``."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int64)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_quadrature.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int]
        L.oracle_basis.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_char_p, C.c_int, dp, dp, ip, ip, C.c_int]
        L.oracle_surface_basis.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_int,
                                           dp, dp, ip, C.POINTER(C.c_int), C.c_int]
        L.oracle_real_sh.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, dp]
        L.oracle_legendre.argtypes = [C.c_int, C.c_double, C.c_int, dp]
        L.oracle_problem_sizes.argtypes = [C.POINTER(OracleProblem), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.oracle_solve.argtypes = [C.POINTER(OracleProblem), dp, C.POINTER(OracleStats)]
        L.oracle_solve_cutoff.argtypes = [C.POINTER(OracleProblem), dp, dp, C.POINTER(OracleStats)]
        L.oracle_pass.argtypes = [C.POINTER(OracleProblem), C.c_int, C.c_int, dp, dp]
        L.oracle_pass_timed.argtypes = [C.POINTER(OracleProblem), C.c_int, C.c_int, dp, dp, dp]
        L.oracle_fokker_planck.argtypes = [C.POINTER(OracleProblem), dp, dp, C.c_int]
        L.oracle_cell_1d.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.oracle_cell_3d.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                                     C.c_double, dp, dp, dp, dp, dp]
        L.oracle_cell_2d.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.oracle_cell_bfp.argtypes = [C.c_char_p] * 4 + [C.c_int] * 6 + [dp, dp, C.c_double, C.c_double, C.c_double,
                                                                       dp, dp, dp, dp, dp, dp]
        L.oracle_cell_bfp1d_awd1.argtypes = [C.c_double] * 6 + [dp] * 4
        L.oracle_constant_linear.argtypes = [C.c_double, C.c_double, dp]
        L.oracle_em_matrix.argtypes = [C.POINTER(OracleProblem), dp, C.c_int]
        L.oracle_ferrer.argtypes = [C.c_int, C.c_int, C.c_double]
        L.oracle_ferrer.restype = C.c_double
        L.oracle_cell_bfp_awd.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp,
                                          dp, dp, dp]
        L.oracle_voronoi_neighbours.argtypes = [C.c_int, dp, dp, dp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                                C.c_int]
        L.oracle_legendre_triple.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oracle_legendre_triple.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc < 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())
    return rc


def quadrature(N, qtype, ndims=1, qdims=None):
    if qdims is None:
        qdims = ndims
    cap = 4 * (N + 2) * (N + 2) + 8
    mu, eta, xi, w = (np.zeros(cap) for _ in range(4))
    nd = _check(lib().oracle_quadrature(N, qtype.encode(), ndims, qdims, _p(mu), _p(eta), _p(xi), _p(w), cap))
    return mu[:nd].copy(), eta[:nd].copy(), xi[:nd].copy(), w[:nd].copy()


def basis(mu, eta, xi, w, L, sn_type, qdims):
    nd = len(w)
    cap = max(nd, (L + 1) ** 2) + 1
    M = np.zeros(nd * cap)
    D = np.zeros(nd * cap)
    pl = np.zeros(cap, np.int64)
    pm = np.zeros(cap, np.int64)
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (mu, eta, xi, w)]
    np_ = _check(lib().oracle_basis(*[_p(a) for a in arr], nd, L, sn_type.encode(), qdims, _p(M), _p(D),
                                    pl.ctypes.data_as(C.POINTER(C.c_int64)),
                                    pm.ctypes.data_as(C.POINTER(C.c_int64)), cap))
    Mm = M[: nd * np_].reshape((np_, nd)).T.copy()   # Nd x Np
    Dm = D[: nd * np_].reshape((nd, np_)).T.copy()   # Np x Nd
    return np_, Mm, Dm, pl[:np_].copy(), pm[:np_].copy()


def surface_basis(mu, eta, xi, w, L, sn_type, qdims, face):
    nd = len(w)
    cap = nd * nd + 1
    M = np.zeros(cap)
    D = np.zeros(cap)
    n2p = np.zeros(nd, np.int64)
    ndp = C.c_int(0)
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (mu, eta, xi, w)]
    nps = _check(lib().oracle_surface_basis(*[_p(a) for a in arr], nd, L, sn_type.encode(), qdims, face,
                                            _p(M), _p(D), n2p.ctypes.data_as(C.POINTER(C.c_int64)),
                                            C.byref(ndp), cap))
    k = ndp.value
    Mm = M[: k * nps].reshape((nps, k)).T.copy()     # Nd+ x Nps
    Dm = D[: k * nps].reshape((k, nps)).T.copy()     # Nps x Nd+
    return nps, Mm, Dm, n2p


def real_sh(L, mu, phi, half_range=False):
    R = np.zeros((L + 1) ** 2)
    _check(lib().oracle_real_sh(L, mu, phi, int(half_range), _p(R)))
    return R


def legendre(L, x, half_range=False):
    P = np.zeros(L + 1)
    _check(lib().oracle_legendre(L, x, int(half_range), _p(P)))
    return P


class _Keep:
    """Holds the numpy arrays referenced by an OracleProblem alive."""


def make_problem(prob, nthreads=1, verbose=0, **over):
    """Build an OracleProblem from a radiant_hip.problems.Problem-like object."""
    k = _Keep()
    k.dx = np.ascontiguousarray(prob.dx, np.float64)
    k.dy = np.ascontiguousarray(prob.dy if prob.ndims >= 2 else [1.0], np.float64)
    k.dz = np.ascontiguousarray(prob.dz if prob.ndims >= 3 else [1.0], np.float64)
    k.mat = np.ascontiguousarray(np.asarray(prob.mat).reshape(-1, order="F"), np.int64)
    k.sigt = np.ascontiguousarray(np.asarray(prob.sigt).reshape(-1, order="F"), np.float64)
    k.sigs = np.ascontiguousarray(np.asarray(prob.sigs).reshape(-1, order="F"), np.float64)
    k.qvol = np.ascontiguousarray(np.asarray(prob.qvol).reshape(-1, order="F"), np.float64)
    k.bc = np.ascontiguousarray(prob.bc, np.int32)
    k.qt = prob.quad_type.encode()
    k.st = prob.sn_type.encode()
    acc = over.get("accel", prob.accel)
    k.ac = acc.encode()
    op = OracleProblem(
        prob.ndims, prob.Nx, prob.Ny, prob.Nz, _p(k.dx), _p(k.dy), _p(k.dz),
        k.mat.ctypes.data_as(C.POINTER(C.c_int64)), prob.Nmat, prob.G, _p(k.sigt), _p(k.sigs), prob.Lxs,
        _p(k.qvol), k.bc.ctypes.data_as(C.POINTER(C.c_int32)), k.qt, prob.quad_order, prob.qdims,
        prob.legendre_order, k.st, k.ac, over.get("restart", prob.restart), over.get("depth", prob.depth),
        over.get("maxit", prob.maxit), over.get("tol", prob.tol), nthreads, verbose)
    ss = getattr(prob, "ssrc", None)
    if ss:
        k.ss = {}
        for f, arr in ss.items():
            k.ss[f] = np.ascontiguousarray(arr, np.float64).ravel()
            op.ssrc[f] = _p(k.ss[f])
        op.Np_source = int(prob.Np_source)
    if getattr(prob, "csd", False):
        k.Sb = np.ascontiguousarray(np.asarray(prob.Sb, np.float64).reshape(-1, order="F"))
        k.dE = np.ascontiguousarray(prob.dE, np.float64)
        op.csd, op.Sb, op.dE = 1, _p(k.Sb), _p(k.dE)
    st = getattr(prob, "solver_type", None)
    if st:
        op.solver_type = SOLVER_TYPES[st.upper()]
    for nm in ("T", "siga"):
        v = getattr(prob, nm, None)
        if v is not None:
            setattr(k, nm, np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1, order="F")))
            setattr(op, nm, _p(getattr(k, nm)))
    fp = getattr(prob, "fokker_planck", None)
    if fp:
        k.fp = fp.encode()
        op.fokker_planck = k.fp
    if prob.quad_type == "lebedev" and prob.quad_order > 11:
        # rules above order 11 come point by point from the host's orbit generation (pinned by
        # exactness in tests/test_quadratures.py)
        from radiant_hip.angular import lebedev
        k.q = [np.ascontiguousarray(a, np.float64) for a in lebedev(prob.quad_order)]
        op.quad_mu, op.quad_eta, op.quad_xi, op.quad_w = (_p(a) for a in k.q)
        op.quad_n = len(k.q[3])
    k.scheme = (getattr(prob, "scheme", None) or "DD").encode()
    op.scheme = k.scheme
    op.scheme_order = int(getattr(prob, "scheme_order", 1) or 1)
    k.scheme_y = (getattr(prob, "scheme_y", None) or "DD").encode()
    op.scheme_y = k.scheme_y
    op.scheme_order_y = int(getattr(prob, "scheme_order_y", 1) or 1)
    op.partial_coupling = 0 if getattr(prob, "full_coupling", True) else 1
    k.scheme_z = (getattr(prob, "scheme_z", None) or "DD").encode()
    op.scheme_z = k.scheme_z
    op.scheme_order_z = int(getattr(prob, "scheme_order_z", 1) or 1)
    k.scheme_e = (getattr(prob, "scheme_e", None) or "DD").encode()
    op.scheme_e = k.scheme_e
    op.scheme_order_e = int(getattr(prob, "scheme_order_e", 1) or 1)
    B = getattr(prob, "em_B", None)
    if B is not None:  # an external magnetic field (Electromagnetic_Field)
        op.em = 1
        for i in range(3):
            op.em_B[i] = float(B[i])
        op.em_charge = float(prob.em_charge)
        k.Eb = np.ascontiguousarray(prob.Eb, np.float64)
        op.Eb = _p(k.Eb)
    op._keep = k
    return op


def voronoi_neighbours(P):
    """The oracle's voronoi_sphere: per node of P [Nd, 3], its neighbours in polygon order."""
    P = np.ascontiguousarray(P, np.float64)
    Nd = len(P)
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    cnt = np.zeros(Nd, np.int32)
    cap = 16 * Nd
    nbr = np.zeros(cap, np.int32)
    ip = C.POINTER(C.c_int32)
    r = lib().oracle_voronoi_neighbours(Nd, _p(x), _p(y), _p(z), cnt.ctypes.data_as(ip), nbr.ctypes.data_as(ip), cap)
    _check(r)
    out, k = [], 0
    for n in range(Nd):
        out.append([int(v) for v in nbr[k:k + cnt[n]]])
        k += cnt[n]
    return out


def em_matrix(prob):
    """ℳ_EM [G, Np, Np] of a problem with a magnetic field, as the oracle builds it."""
    op = make_problem(prob)
    nd, np_, nps, S = sizes(prob)
    out = np.zeros(prob.G * np_ * np_)
    _check(lib().oracle_em_matrix(C.byref(op), _p(out), len(out)))
    return np.stack([out[g * np_ * np_:(g + 1) * np_ * np_].reshape((np_, np_), order="F") for g in range(prob.G)])


def cell_bfp1d_awd1(mu, dx, St, Sm, Sp, dE, Qn, fe, fx):
    """One 1D slowing-down cell of orders (2, 2) with the coupled linear AWD1: (psi[4], fe[2], fx[2])."""
    arr = lambda v: np.ascontiguousarray(v, np.float64).copy()
    q, e, x = arr(Qn), arr(fe), arr(fx)
    psi = np.zeros(4)
    _check(lib().oracle_cell_bfp1d_awd1(mu, dx, St, Sm, Sp, dE, _p(q), _p(e), _p(x), _p(psi)))
    return psi, e, x


def constant_linear(x, y):
    """(subdomain 0..11, [Qx, Qy, Tx, Ty]) of the coupled linear closure at centroid (x, y)."""
    out = np.zeros(4)
    r = lib().oracle_constant_linear(float(x), float(y), _p(out))
    _check(r)
    return r, out


def ferrer(l, m, x):
    return lib().oracle_ferrer(l, m, x)


def fokker_planck(prob):
    """(ℳ [Np, Np], λ₀) of a BFP / FP problem as the oracle builds it."""
    op = make_problem(prob)
    nd, np_, nps, S = sizes(prob)
    M = np.zeros(np_ * np_)
    lam = C.c_double()
    r = lib().oracle_fokker_planck(C.byref(op), _p(M), C.byref(lam), np_)
    _check(r)
    return M.reshape((np_, np_), order="F"), lam.value


def spatial_moments(prob):
    """Legendre moments per cell (SN.get_schemes Nm[5]): over x[, y[, z]] and, for the slowing-down
    types, the energy axis; the product of the orders with full coupling, 1 + the sum of
    (order - 1) without."""
    orders = [max(1, int(getattr(prob, "scheme_order", 1) or 1))]
    if prob.ndims >= 2:
        orders.append(max(1, int(getattr(prob, "scheme_order_y", 1) or 1)))
    if prob.ndims >= 3:
        orders.append(max(1, int(getattr(prob, "scheme_order_z", 1) or 1)))
    if getattr(prob, "csd", False):
        orders.append(max(1, int(getattr(prob, "scheme_order_e", 1) or 1)))
    if getattr(prob, "full_coupling", True):
        return int(np.prod(orders))
    return 1 + sum(o - 1 for o in orders)


def cell_1d(scheme, K, mu, St, dx, Qn, fx_in):
    """One 1D cell of the given closure and order: (spatial moments [K], outgoing face flux)."""
    q = np.ascontiguousarray(Qn, np.float64)
    psi = np.zeros(K)
    fx = C.c_double(fx_in)
    _check(lib().oracle_cell_1d(scheme.encode(), K, mu, St, dx, _p(q), C.byref(fx), _p(psi)))
    return psi, fx.value


def cell_2d(schemes, orders, full_coupling, mu, eta, St, dx, dy, Qn, fx_in, fy_in):
    """One 2D cell: (spatial moments, outgoing x-face moments [Oy], outgoing y-face moments [Ox])."""
    ox, oy = orders
    nm = ox * oy if full_coupling else ox + oy - 1
    q = np.ascontiguousarray(Qn, np.float64)
    fx = np.ascontiguousarray(fx_in, np.float64).copy()
    fy = np.ascontiguousarray(fy_in, np.float64).copy()
    psi = np.zeros(nm)
    _check(lib().oracle_cell_2d(schemes[0].encode(), schemes[1].encode(), ox, oy, int(not full_coupling), mu, eta,
                                St, dx, dy, _p(q), _p(fx), _p(fy), _p(psi)))
    return psi, fx, fy


def cell_3d(schemes, orders, full_coupling, om, h, St, Qn, fx_in, fy_in, fz_in):
    """One 3D cell: (spatial moments, outgoing x-, y- and z-face moments)."""
    ox, oy, oz = orders
    nm = ox * oy * oz if full_coupling else ox + oy + oz - 2
    arr = lambda v: np.ascontiguousarray(v, np.float64).copy()
    q, o, hh, fx, fy, fz = arr(Qn), arr(om), arr(h), arr(fx_in), arr(fy_in), arr(fz_in)
    psi = np.zeros(nm)
    _check(lib().oracle_cell_3d(schemes[0].encode(), schemes[1].encode(), schemes[2].encode(), ox, oy, oz,
                                int(not full_coupling), _p(o), _p(hh), St, _p(q), _p(fx), _p(fy), _p(fz), _p(psi)))
    return psi, fx, fy, fz


def cell_bfp(schemes, orders, full_coupling, ndims, om, h, St, Sm, Sp, Qn, fe_in, fx_in, fy_in, fz_in):
    """One slowing-down cell: schemes/orders over (E, x, y, z), Sm/Sp the stopping powers at the
    group's upper/lower energy per unit group width; returns (moments, outgoing energy, x, y and z
    face moments)."""
    oe, ox, oy, oz = orders
    nm = oe * ox * oy * oz if full_coupling else oe + ox + oy + oz - 3
    arr = lambda v: np.ascontiguousarray(v, np.float64).copy()
    q, o, hh = arr(Qn), arr(om), arr(h)
    fe, fx, fy, fz = arr(fe_in), arr(fx_in), arr(fy_in), arr(fz_in)
    psi = np.zeros(nm)
    _check(lib().oracle_cell_bfp(*(s.encode() for s in schemes), oe, ox, oy, oz, int(not full_coupling), ndims,
                                 _p(o), _p(hh), St, Sm, Sp, _p(q), _p(fe), _p(fx), _p(fy), _p(fz), _p(psi)))
    return psi, fe, fx, fy, fz


def cell_bfp_awd(ndims, om, h, St, Sm, Sp, dE, Q, fe, fx, fy, fz):
    """One slowing-down cell of order 1 with the adaptive closure: (psi, fe, fx, fy, fz) outgoing."""
    arr = lambda v: np.ascontiguousarray(v, np.float64).copy()
    o, hh, q = arr(om), arr(h), arr([Q])
    f = [arr([v]) for v in (fe, fx, fy, fz)]
    psi = np.zeros(1)
    _check(lib().oracle_cell_bfp_awd(ndims, _p(o), _p(hh), St, Sm, Sp, dE, _p(q), *[_p(v) for v in f], _p(psi)))
    return psi[0], f[0][0], f[1][0], f[2][0], f[3][0]


def legendre_triple(a, b, c):
    """½ ∫ P_a P_b P_c dx as the oracle computes it."""
    return lib().oracle_legendre_triple(a, b, c)


def sizes(prob):
    op = make_problem(prob)
    nd, np_, nps = C.c_int(), C.c_int(), C.c_int()
    S = C.c_int64()
    _check(lib().oracle_problem_sizes(C.byref(op), C.byref(nd), C.byref(np_), C.byref(nps), C.byref(S)))
    return nd.value, np_.value, nps.value, S.value


def solve(prob, nthreads=1, verbose=0, **over):
    """Full solve; returns (phi[G,Np,Nc] in Radiant order, stats list of dicts)."""
    op = make_problem(prob, nthreads=nthreads, verbose=verbose, **over)
    nd, np_, nps, S = sizes(prob)
    nc = prob.Nx * prob.Ny * prob.Nz * spatial_moments(prob)
    phi = np.zeros(prob.G * np_ * nc)
    st = (OracleStats * prob.G)()
    _check(lib().oracle_solve(C.byref(op), _p(phi), st))
    stats = [dict(niter=s.niter, resid=s.resid, rho=s.rho, converged=bool(s.converged), passes=s.passes) for s in st]
    return phi.reshape((prob.G, np_, nc), order="F"), stats


def solve_cutoff(prob, nthreads=1, verbose=0, **over):
    """Full solve that also returns 𝚽cutoff (sn_flux.jl:213-218): (phi[G,Np,Nc], cutoff[Np,Nc],
    stats).  cutoff is zero unless continuous slowing down is on."""
    op = make_problem(prob, nthreads=nthreads, verbose=verbose, **over)
    nd, np_, nps, S = sizes(prob)
    nc = prob.Nx * prob.Ny * prob.Nz * spatial_moments(prob)
    phi = np.zeros(prob.G * np_ * nc)
    cut = np.zeros(np_ * nc)
    st = (OracleStats * prob.G)()
    _check(lib().oracle_solve_cutoff(C.byref(op), _p(phi), _p(cut), st))
    stats = [dict(niter=s.niter, resid=s.resid, rho=s.rho, converged=bool(s.converged), passes=s.passes) for s in st]
    return phi.reshape((prob.G, np_, nc), order="F"), cut.reshape((np_, nc), order="F"), stats


def transport_pass(prob, g, homogeneous, z, qlout, nthreads=1):
    """One pass T(z) / A z of group g (0-based) on state z (length S); returns the new state."""
    op = make_problem(prob, nthreads=nthreads)
    zz = np.ascontiguousarray(z, np.float64).copy()
    q = np.ascontiguousarray(qlout, np.float64)
    _check(lib().oracle_pass(C.byref(op), g, int(homogeneous), _p(zz), _p(q)))
    return zz


def timed_pass(prob, g, homogeneous, z, qlout, nthreads=1):
    """transport_pass that also returns the wall time (s) of the pass alone, set-up excluded."""
    op = make_problem(prob, nthreads=nthreads)
    zz = np.ascontiguousarray(z, np.float64).copy()
    q = np.ascontiguousarray(qlout, np.float64)
    sec = C.c_double()
    _check(lib().oracle_pass_timed(C.byref(op), g, int(homogeneous), _p(zz), _p(q), C.byref(sec)))
    return zz, sec.value
