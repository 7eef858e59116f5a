"""Host-side angular setup: quadratures and Galerkin moment<->direction operators.

This is the part of Radiant's SN setup that stays on the host (sn_flux.jl:38-48): it produces
the direction cosines, weights, the volume operators M (Nd x Np) and D (Np x Nd) with the
Legendre order of each moment, and the half-range surface operators of every face.  A Julia
caller keeps Radiant's own functions for this and passes the arrays through the C ABI; this
numpy restatement lets Python callers build the same arrays.  Results are cached per
(quadrature, order, dims, L, type) because the Gram-Schmidt selection is O(Nd^3).
"""
import functools
import math

import numpy as np

PI = math.pi
EPS = np.finfo(np.float64).eps


# ----------------------------------------------------------------------------- quadratures
def gauss_legendre(N):
    """tools/gauss_legendre.jl:18-74: Tricomi guess + Newton on Bonnet's recursion (descending x)."""
    x = np.zeros(N)
    for i in range(1, N // 2 + 1):
        phi = ((PI / 2) * (4 * i - 1)) / (2 * N + 1)
        s = math.sin(phi)
        x[i - 1] = ((1 - (N - 1) / (8 * N ** 3)) - 1 / (384 * N ** 4) * (39 - 28 / (s * s))) * math.cos(phi)
        x[N - i] = -x[i - 1]
    w = np.zeros(N)
    dPdx = 0.0
    for i in range(N):
        xi = float(x[i])
        xold = 2.0
        while abs(xi - xold) > EPS:
            xold = xi
            P1, P2 = 1.0, xi
            for n in range(2, N + 1):
                P1, P2 = P2, ((2 * n - 1) * xi * P2 - (n - 1) * P1) / n
            dPdx = N * (P1 - xi * P2) / (1 - xi * xi)
            xi -= P2 / dPdx
        x[i] = xi
        w[i] = 2 / ((1 - xi * xi) * dPdx * dPdx)
    return x, w


def gauss_lobatto(N, nmax=200):
    """tools/gauss_lobatto.jl:18-76: nodes +-1 and the roots of P'_{N-1}, found by the same
    iteration from x = cos(pi i/(N-1)); weights 2/(N(N-1) P_{N-1}(x)^2); the centre node of an
    odd rule is set to 0."""
    if N < 2:
        raise ValueError("Gauss-Lobatto needs at least 2 points")
    x = np.cos(PI * np.arange(N) / (N - 1))
    w = np.zeros(N)
    w[0] = w[-1] = 2.0 / (N * (N - 1))
    for i in range(1, N - 1):
        xi = float(x[i])
        xold = 2.0
        P2 = 0.0
        it = 0
        while abs(xi - xold) > EPS and it < nmax:
            xold = xi
            P1, P2 = 1.0, xi
            for n in range(2, N):
                P1, P2 = P2, ((2 * n - 1) * xi * P2 - (n - 1) * P1) / n
            d1 = N * (P1 - xi * P2) / (1 - xi * xi)
            d2 = (2 * xi * d1 - N * (N + 1) * P2) / (1 - xi * xi)
            xi -= d1 / d2
            it += 1
        x[i] = xi
        w[i] = 2.0 / (N * (N - 1) * P2 * P2)
    if N % 2 == 1:
        x[(N + 1) // 2 - 1] = 0.0
    if len(set(x.tolist())) != N:
        raise ValueError(f"Unable to build Lobatto quadrature at order N = {N}.")
    return x, w


# Lebedev rules available in closed form (order: [(symmetry class, parameter, weight)]); the
# weights are fractions of the sphere and the classes are those of the octahedral group O_h:
# 1 = (1,0,0) 6 points, 2 = (0,a,a) a=1/sqrt2 12, 3 = (a,a,a) a=1/sqrt3 8, 4 = (a,a,b)
# b=sqrt(1-2a^2) 24, 5 = (a,b,0) b=sqrt(1-a^2) 24.  Each rule integrates every spherical
# harmonic of degree <= order exactly.
LEBEDEV_RULES = {
    3: [(1, 0.0, 1 / 6)],
    5: [(1, 0.0, 1 / 15), (3, 0.0, 3 / 40)],
    7: [(1, 0.0, 1 / 21), (2, 0.0, 4 / 105), (3, 0.0, 9 / 280)],
    9: [(1, 0.0, 1 / 105), (3, 0.0, 9 / 280), (5, math.sqrt((1 - 1 / math.sqrt(3)) / 2), 1 / 35)],
    11: [(1, 0.0, 4 / 315), (2, 0.0, 64 / 2835), (3, 0.0, 27 / 1280), (4, 1 / math.sqrt(11), 14641 / 725760)],
}


def _oh_orbit(code, a):
    s2 = ((1, 1), (1, -1), (-1, 1), (-1, -1))
    pts = []
    if code == 1:
        for ax in range(3):
            for sg in (1.0, -1.0):
                v = [0.0, 0.0, 0.0]
                v[ax] = sg
                pts.append(v)
    elif code == 2:
        c = 1 / math.sqrt(2)
        for z in range(3):
            o = [k for k in range(3) if k != z]
            for s1, s_2 in s2:
                v = [0.0, 0.0, 0.0]
                v[o[0]], v[o[1]] = s1 * c, s_2 * c
                pts.append(v)
    elif code == 3:
        c = 1 / math.sqrt(3)
        for s1, s_2 in s2:
            for s3 in (1, -1):
                pts.append([s1 * c, s_2 * c, s3 * c])
    elif code == 4:
        b = math.sqrt(1 - 2 * a * a)
        for bi in range(3):
            for s1, s_2 in s2:
                for s3 in (1, -1):
                    v = [s1 * a, s_2 * a, s3 * a]
                    v[bi] = (s1, s_2, s3)[bi] * b
                    pts.append(v)
    elif code == 5:
        b = math.sqrt(1 - a * a)
        for z in range(3):
            o = [k for k in range(3) if k != z]
            for p_, q_ in ((a, b), (b, a)):
                for s1, s_2 in s2:
                    v = [0.0, 0.0, 0.0]
                    v[o[0]], v[o[1]] = s1 * p_, s_2 * q_
                    pts.append(v)
    return pts


def _signed(p, zero_free):
    """The sign patterns of a point, the first coordinate's sign outermost, zeros left unsigned."""
    out = [[]]
    for k, x in enumerate(p):
        out = [q + [s * x] for q in out for s in ((1, -1) if x != 0.0 else (1,))]
    return out


def _oh_orbit_ab(code, a, b):
    """Orbit of the octahedral group with its sign changes, for the table-based rules: type 4
    (a, a, b) with b = sqrt(1 - 2 a^2), 5 (a, b, 0) with b = sqrt(1 - a^2), 6 (a, b, c) with
    c = sqrt(1 - a^2 - b^2); types 1-3 as the closed-form rules.  The coordinate permutations are
    taken in the order of tools/lebedev.jl gen_oh!."""
    if code in (1, 2, 3):
        return _oh_orbit(code, a)
    if code == 4:
        c = math.sqrt(1 - 2 * a * a)
        perms = [(a, a, c), (a, c, a), (c, a, a)]
    elif code == 5:
        c = math.sqrt(1 - a * a)
        perms = [(a, c, 0.0), (c, a, 0.0), (a, 0.0, c), (c, 0.0, a), (0.0, a, c), (0.0, c, a)]
    elif code == 6:
        c = math.sqrt(1 - a * a - b * b)
        perms = [(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)]
    else:
        raise ValueError("orbit type must be 1..6")
    pts = []
    for p in perms:
        pts.extend(_signed(p, True))
    return pts


def lebedev(N):
    """tools/lebedev.jl:21-4751: the Lebedev-Laikov rule of algebraic order N (31 rules, 6 to 5294
    points), weights scaled to 4 pi as Radiant does.  Orders 3-11 are written from their closed
    forms; higher orders generate their orbits from the published parameters (lebedev_rules)."""
    from .lebedev_rules import ORDER_POINTS, RULES
    if N not in ORDER_POINTS:
        raise ValueError(f"Lebedev quadrature is not available at order {N} "
                         f"(available: {', '.join(str(k) for k in sorted(ORDER_POINTS))}).")
    mu, eta, xi, w = [], [], [], []
    if N in LEBEDEV_RULES:
        orbits = [(code, a, v, _oh_orbit(code, a)) for code, a, v in LEBEDEV_RULES[N]]
    else:
        orbits = [(code, a, v, _oh_orbit_ab(code, a, b)) for code, a, b, v in RULES[ORDER_POINTS[N]]]
    for code, a, v, pts in orbits:
        for p in pts:
            mu.append(p[0]); eta.append(p[1]); xi.append(p[2]); w.append(4 * PI * v)
    if len(w) != ORDER_POINTS[N]:
        raise RuntimeError("Lebedev rule with a wrong number of points")
    return np.array(mu), np.array(eta), np.array(xi), np.array(w)


def gauss_legendre_chebychev(N):
    """tools/gauss_legendre_chebychev.jl:19-39: GL in mu (ascending) x 2N Chebyshev azimuths."""
    mun, wn = gauss_legendre(N)
    idx = np.argsort(mun, kind="stable")
    mun, wn = mun[idx], wn[idx]
    m = np.arange(1, 2 * N + 1)
    phi = PI / N * (m - 0.5)
    mu = np.repeat(mun, 2 * N)
    st = np.repeat(np.sqrt(1 - mun * mun), 2 * N)
    eta = st * np.tile(np.cos(phi), N)
    xi = st * np.tile(np.sin(phi), N)
    w = np.repeat(wn * PI / N, 2 * N)
    return mu, eta, xi, w


def _newton_bisection(f, dfdx, x1, x2, eps0=EPS, nmax=100):
    """tools/newton_bisection.jl:20-75."""
    f1, f2 = f(x1), f(x2)
    if f1 * f2 > 0:
        raise ValueError("Root must be bracketed")
    if f1 == 0:
        return x1
    if f2 == 0:
        return x2
    xl, xh = (x1, x2) if f1 < 0 else (x2, x1)
    rts = 0.5 * (x1 + x2)
    dxold = abs(x2 - x1)
    dx = dxold
    fr, dfr = f(rts), dfdx(rts)
    for _ in range(nmax):
        if ((rts - xh) * dfr - fr) * ((rts - xl) * dfr - fr) > 0 or abs(2 * fr) > abs(dxold * dfr):
            dxold = dx
            dx = 0.5 * (xh - xl)
            rts = xl + dx
        else:
            dxold = dx
            dx = fr / dfr
            rts -= dx
        if abs(dx) < eps0:
            return rts
        fr, dfr = f(rts), dfdx(rts)
        if fr < 0:
            xl = rts
        else:
            xh = rts
    raise RuntimeError("newton_bisection did not converge")


def carlson(N):
    """tools/carlson.jl:20-94: equal-weight set with N(N+2) directions, octant-major order."""
    if N < 2 or N % 2:
        raise ValueError("Carlson quadrature available for even N >= 2.")
    H = N // 2
    n = np.arange(1, H + 1)
    wn = 4 * (N - 2 * n + 2) / (N * (N + 2))
    mum = 1 - (N - 2 * n + 2) * (N - 2 * n + 4) / (N * (N + 2))
    mu0 = 1 - (N - 2 * n + 2) ** 2 / (N * (N + 2))
    if N == 2:
        mun = np.array([1 / math.sqrt(3)])
    else:
        def f(x):
            s = 0.0
            for k in range(H):
                t = mu0[k] + x * mum[k]
                s += wn[k] * (t * t)
            return 1 / 3 - s

        def df(x):
            s = 0.0
            for k in range(H):
                s += 2 * wn[k] * mum[k] * (mu0[k] + x * mum[k])
            return -s

        x = _newton_bisection(f, df, 1.0, -1.0)
        mun = mu0 + x * mum
    w0 = 1 / (N * (N + 2))
    if N == 2:
        A = 1.0
    else:
        def g(a):
            s = 0.0
            for k in range(1, H + 1):
                m = mun[k - 1]
                s += (PI / 2) * wn[k - 1] * m - (4 * PI) * w0 / math.sqrt(2) * math.sin(PI * a / 4) * \
                    math.sqrt(1 - m * m) / math.sin(PI * a / (2 * N - 4 * k + 4))
            return s

        def dg(a):
            s = 0.0
            for k in range(1, H + 1):
                m = mun[k - 1]
                sq = math.sqrt(1 - m * m)
                d = 2 * N - 4 * k + 4
                sd = math.sin(PI * a / d)
                s += -4 * PI * w0 * math.sqrt(2) * PI / 8 * (
                    math.cos(PI * a / 4) * sq / sd
                    - 2 * math.sin(PI * a / 4) * sq * math.cos(PI * a / d) / (sd * sd) / (N - 2 * k + 2))
            return s

        A = _newton_bisection(g, dg, 0.5, 2.0)
    sg = [(1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1), (-1, 1, 1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1)]
    mu, eta, xi = [], [], []
    for s in sg:
        for k in range(1, H + 1):
            for m in range(1, H - k + 2):
                phi = PI / 2 * ((2 * m - 1) / (N - 2 * k + 2) * A + (1 - A) / 2)
                u = s[0] * mun[k - 1]
                st = math.sqrt(1 - u * u)
                mu.append(u)
                eta.append(s[1] * st * math.cos(phi))
                xi.append(s[2] * st * math.sin(phi))
    nd = len(mu)
    return np.array(mu), np.array(eta), np.array(xi), np.full(nd, (4 * PI) * w0)


def quadrature_dimension(quad_type, ndims, qdims=0):
    """structures/SN.jl:638-659 (0 = default)."""
    if quad_type in ("gauss-legendre", "gauss-lobatto"):
        return 1
    if ndims in (1, 2):
        return 2 if qdims == 2 else 3
    return 3


def quadrature(N, quad_type, ndims=1, qdims=None):
    """tools/quadrature.jl:19-59; returns mu, eta, xi, w (1D sets expanded to [mu,0,0])."""
    if qdims is None:
        qdims = ndims
    if quad_type == "gauss-legendre" and ndims == 1:
        mu, w = gauss_legendre(N)
        eta = np.zeros_like(mu)
        xi = np.zeros_like(mu)
    elif quad_type == "gauss-lobatto" and ndims == 1:
        mu, w = gauss_lobatto(N)
        eta = np.zeros_like(mu)
        xi = np.zeros_like(mu)
    elif quad_type == "gauss-legendre-chebychev":
        mu, eta, xi, w = gauss_legendre_chebychev(N)
    elif quad_type == "lebedev":
        mu, eta, xi, w = lebedev(N)
    elif quad_type == "carlson":
        mu, eta, xi, w = carlson(N)
    else:
        raise ValueError(f"Unknown {quad_type} quadrature in {ndims}D geometry.")
    if qdims == 2:
        nd = len(w)
        w0 = w.copy()
        for i in range(nd):
            if xi[i] >= 0:
                continue
            e = np.abs(mu[i] - mu) ** 2 + np.abs(eta[i] - eta) ** 2 + np.abs(xi[i] + xi) ** 2
            w0[int(np.argmin(e))] += w0[i]
        keep = xi >= 0
        mu, eta, xi, w = mu[keep], eta[keep], xi[keep], w0[keep]
    return mu, eta, xi, w


# ----------------------------------------------------------------------------- harmonics
@functools.lru_cache(maxsize=None)
def _ff_int(num, den):
    """The integer part of factorial_factor: the same sequential sum of logs, computed once per
    (num, den)."""
    f = 0.0
    for n in num:
        for i in range(1, n + 1):
            f += math.log(i)
    for n in den:
        for i in range(1, n + 1):
            f -= math.log(i)
    return f


def _ff(num, den, ex=()):
    """tools/factorial_factor.jl:17-33."""
    f = _ff_int(tuple(num), tuple(den))
    for x, m in ex:
        if m == 0:
            continue
        if x == 0:
            return 0.0 if m > 0 else math.inf
        f += m * math.log(x)
    return math.exp(f)


def legendre_up_to_L(L, x):
    """tools/legendre_polynomials.jl:52-75."""
    P = np.zeros(L + 1)
    if x == 1:
        P[:] = 1
    elif x == -1:
        P[:] = [(-1) ** l for l in range(L + 1)]
    else:
        P[0] = 1
        if L >= 1:
            P[1] = x
        for l in range(2, L + 1):
            P[l] = ((2 * l - 1) * x * P[l - 1] - (l - 1) * P[l - 2]) / l
    return P


def half_range_legendre_up_to_L(L, x):
    """tools/legendre_polynomials.jl:91-93."""
    return np.sqrt(2 * np.arange(L + 1) + 1) * legendre_up_to_L(L, 2 * x - 1)


def jacobi_up_to_L(L, a, b, x):
    """tools/jacobi_polynomials.jl:16-49."""
    P = np.zeros(L + 1)
    if x == 1:
        for l in range(L + 1):
            P[l] = _ff([l + a], [l, a])
    elif x == -1:
        for l in range(L + 1):
            P[l] = (-1) ** l * _ff([l + b], [l, b])
    else:
        P[0] = 1
        if L >= 1:
            P[1] = (a + 1) + 0.5 * (a + b + 2) * (x - 1)
        for l in range(2, L + 1):
            if l < 125:
                s = 2 * l + a + b
                P[l] = ((s - 1) * ((s * (s - 2)) * x + a * a - b * b) * P[l - 1]
                        - (2 * (l + a - 1) * (l + b - 1) * s) * P[l - 2]) / (2 * l * (l + a + b) * (s - 2))
            else:
                th = math.acos(x)
                k = 1 / (math.sqrt(PI) * math.sin(0.5 * th) ** (a + 0.5) * math.cos(0.5 * th) ** (b + 0.5))
                Nn = l + 0.5 * (a + b + 1)
                P[l] = 1 / math.sqrt(l) * k * math.cos(Nn * th - 0.5 * PI * (a + 0.5))
    return P


def _plm(L, x, use_ends):
    P = [np.zeros(l + 1) for l in range(L + 1)]
    if use_ends:
        J = jacobi_up_to_L(L, 0, 0, x)
        for l in range(L + 1):
            P[l][0] = J[l]
    else:
        for m in range(L + 1):
            J = jacobi_up_to_L(L - m, m, m, x)
            for l in range(m, L + 1):
                P[l][m] = _ff([l + m], [l], [(2, -m), (1 - x * x, m / 2)]) * J[l - m]
    return P


def real_sh_up_to_L(L, mu, phi):
    """tools/real_spherical_harmonics.jl:52-88 -> list R[l][l+m]."""
    P = _plm(L, mu, mu in (-1.0, 1.0))
    R = []
    for l in range(L + 1):
        r = np.zeros(2 * l + 1)
        for m in range(-l, l + 1):
            am = abs(m)
            T = math.cos(m * phi) if m >= 0 else math.sin(am * phi)
            C = math.sqrt((2 - (m == 0)) * _ff([l - am], [l + am]))
            r[l + m] = C * P[l][am] * T
        R.append(r)
    return R


def real_half_range_sh_up_to_L(L, mu, phi):
    """tools/real_spherical_harmonics.jl:104-139 -> list R[l][l+m] on mu in [0,1]."""
    x = 2 * mu - 1
    P = _plm(L, x, mu in (-1.0, 1.0))
    R = []
    for l in range(L + 1):
        r = np.zeros(2 * l + 1)
        for m in range(-l, l + 1):
            am = abs(m)
            T = math.cos(m * phi) if m >= 0 else math.sin(am * phi)
            C = math.sqrt((2 - (m == 0)) / (2 * PI) * (2 * l + 1) * _ff([l - am], [l + am]))
            r[l + m] = C * P[l][am] * T
        R.append(r)
    return R


# ----------------------------------------------------------------------------- operators
def galerkin_minimum_order(Nd, qdims):
    """angular_polynomial_basis.jl:9-17."""
    if qdims == 1:
        return Nd - 1
    n = 0
    while True:
        cnt = (n + 1) * (n + 2) // 2 if qdims == 2 else (n + 1) ** 2
        if cnt >= Nd:
            return n
        n += 1


def _harm_matrix(R, L, qdims):
    """Stack harmonics into rows in Radiant's traversal order (l, m); returns rows, (l, m) list."""
    rows, lm = [], []
    for l in range(L + 1):
        mmin = 0 if (qdims == 2 and l > 0) else -l
        for m in range(mmin, l + 1):
            rows.append(np.array([R[n][l][m + l] for n in range(len(R))]))
            lm.append((l, m))
    return rows, lm


def _gram_schmidt(Nd, L, R, w, qdims, gtype):
    """angular_polynomial_basis.jl:465-548 -- classical GS selection with the 1e-5 threshold."""
    rows, lm = _harm_matrix(R, L, qdims)
    U = np.zeros((Nd, Nd))
    v0 = rows[0]
    U[:, 0] = v0 / math.sqrt(float(np.sum(v0 * v0)))
    keep = [0]
    for k in range(1, len(rows)):
        if len(keep) == Nd:
            break
        v = rows[k]
        i = len(keep)
        ip = U[:, :i].T @ v
        vk = v - U[:, :i] @ ip
        nk = math.sqrt(float(np.sum(vk * vk)))
        if nk > 1e-5:
            U[:, i] = vk / nk
            keep.append(k)
    if len(keep) < Nd:
        return None
    pl = np.array([lm[k][0] for k in keep], np.int64)
    pm = np.array([lm[k][1] for k in keep], np.int64)
    H = np.array([rows[k] for k in keep])  # Np x Nd
    if gtype == 1:
        D = H * w[None, :]
        M = np.linalg.inv(D)
    elif gtype == 2:
        M = (H * ((2 * pl + 1) / (4 * PI))[:, None]).T
        D = np.linalg.inv(M)
    else:
        M = H.T.copy()
        D = np.linalg.inv(M)
    return M, D, pl, pm


def angular_polynomial_basis(mu, eta, xi, w, L, sn_type, qdims):
    """angular_polynomial_basis.jl:51-179 -> (Np, M[Nd,Np], D[Np,Nd], pl, pm)."""
    Nd = len(w)
    gal = sn_type in ("galerkin-m", "galerkin-d")
    Lb = max(L, galerkin_minimum_order(Nd, qdims)) if gal else L
    if qdims == 1:
        Pl = np.array([legendre_up_to_L(Lb, m) for m in mu])  # Nd x (Lb+1)
        if sn_type == "standard":
            l = np.arange(L + 1)
            M = (2 * l + 1) / 2 * Pl[:, : L + 1]
            D = (w[:, None] * Pl[:, : L + 1]).T
            return L + 1, M, D, l.astype(np.int64), np.zeros(L + 1, np.int64)
        l = np.arange(Nd)
        if sn_type == "galerkin-m":
            M = (2 * l + 1) / 2 * Pl[:, :Nd]
            D = np.linalg.inv(M)
        elif sn_type == "galerkin-d":
            D = (w[:, None] * Pl[:, :Nd]).T.copy()
            M = np.linalg.inv(D)
        else:
            raise ValueError("Unknown method.")
        return Nd, M, D, l.astype(np.int64), np.zeros(Nd, np.int64)
    phi = np.arctan2(xi, eta)
    if sn_type == "standard":
        R = [real_sh_up_to_L(L, mu[n], phi[n]) for n in range(Nd)]
        rows, lm = [], []
        for l in range(L + 1):
            for m in range(-l, l + 1):
                if qdims == 2 and (l + m) % 2 == 1:
                    continue
                rows.append([R[n][l][m + l] for n in range(Nd)])
                lm.append((l, m))
        H = np.array(rows)
        pl = np.array([a for a, _ in lm], np.int64)
        pm = np.array([b for _, b in lm], np.int64)
        M = (H * ((2 * pl + 1) / (4 * PI))[:, None]).T
        D = H * w[None, :]
        return len(lm), M, D, pl, pm
    if not gal:
        raise ValueError("Unknown method.")
    gt = 1 if sn_type == "galerkin-d" else 2
    Lg = Lb
    while True:
        R = [real_sh_up_to_L(Lg, mu[n], phi[n]) for n in range(Nd)]
        out = _gram_schmidt(Nd, Lg, R, w, qdims, gt)
        if out is not None:
            M, D, pl, pm = out
            return Nd, M, D, pl, pm
        Lg += 1
        if Lg > Lb + Nd + 10:
            raise RuntimeError("Galerkin: unable to find enough independent spherical harmonics")


def surface_basis(mu, eta, xi, w, L, sn_type, qdims, face, with_indices=False):
    """angular_polynomial_basis.jl:262-436 for face 0..5 (X-,X+,Y-,Y+,Z-,Z+).
    Returns (Nps, M[Nd+,Nps], D[Nps,Nd+], n_to_np[Nd] (1-based, 0 = not entering)), and with
    with_indices also the degree and order (pl, pm) of each half-range moment."""
    out = _surface_basis(mu, eta, xi, w, L, sn_type, qdims, face)
    return out if with_indices else out[:4]


def _surface_basis(mu, eta, xi, w, L, sn_type, qdims, face):
    Nd = len(w)
    if qdims == 1:
        if face > 1:
            raise ValueError("Unknown surface for 1D angular domain.")
        sel = mu > 0 if face == 0 else mu < 0
        mup = mu[sel] if face == 0 else -mu[sel]
    else:
        a, ph = [(mu, np.arctan2(xi, eta)), (eta, np.arctan2(mu, xi)), (xi, np.arctan2(eta, mu))][face // 2]
        sel = a > 0 if face % 2 == 0 else a < 0
        mup = a[sel] if face % 2 == 0 else -a[sel]
        php = ph[sel]
    idx = np.nonzero(sel)[0]
    ndp = len(idx)
    wp = w[idx]
    n2p = np.zeros(Nd, np.int64)
    n2p[idx] = np.arange(1, ndp + 1)
    gal = sn_type in ("galerkin-m", "galerkin-d")
    Lb = max(L, galerkin_minimum_order(ndp, qdims)) if gal else L
    if qdims == 1:
        Pl = np.array([half_range_legendre_up_to_L(Lb, m) for m in mup])
        if sn_type == "standard":
            M = Pl[:, : L + 1].copy()
            D = (wp[:, None] * Pl[:, : L + 1]).T.copy()
            return L + 1, M, D, n2p, np.arange(L + 1), np.zeros(L + 1, np.int64)
        if sn_type == "galerkin-m":
            M = Pl[:, :ndp].copy()
            D = np.linalg.inv(M)
        elif sn_type == "galerkin-d":
            D = (wp[:, None] * Pl[:, :ndp]).T.copy()
            M = np.linalg.inv(D)
        else:
            raise ValueError("Unknown method.")
        return ndp, M, D, n2p, np.arange(ndp), np.zeros(ndp, np.int64)
    if sn_type == "standard":
        R = [real_half_range_sh_up_to_L(L, abs(mup[i]), php[i]) for i in range(ndp)]
        rows, lm = [], []
        for l in range(L + 1):
            for m in range(-l, l + 1):
                if qdims == 2 and (l + m) % 2 == 1:
                    continue
                rows.append([R[i][l][m + l] for i in range(ndp)])
                lm.append((l, m))
        H = np.array(rows)
        return (len(rows), H.T.copy(), H * wp[None, :], n2p, np.array([a for a, _ in lm]),
                np.array([b for _, b in lm]))
    if not gal:
        raise ValueError("Unknown method.")
    gt = 1 if sn_type == "galerkin-d" else 3
    Lg = Lb
    while True:
        R = [real_half_range_sh_up_to_L(Lg, abs(mup[i]), php[i]) for i in range(ndp)]
        out = _gram_schmidt(ndp, Lg, R, wp, qdims, gt)
        if out is not None:
            M, D, pl, pm = out
            return ndp, M, D, n2p, np.asarray(pl), np.asarray(pm)
        Lg += 1
        if Lg > Lb + ndp + 10:
            raise RuntimeError("Surface Galerkin: unable to find enough independent half-range harmonics")


class AngularSet:
    """All angular arrays of one (quadrature, L, type, dims) combination."""

    def __init__(self, quad_type, quad_order, ndims, qdims, L, sn_type):
        Q = quadrature_dimension(quad_type, ndims, qdims)
        self.qdims = Q
        self.mu, self.eta, self.xi, self.w = quadrature(quad_order, quad_type, ndims, Q)
        self.Nd = len(self.w)
        self.Np, self.M, self.D, self.pl, self.pm = angular_polynomial_basis(
            self.mu, self.eta, self.xi, self.w, L, sn_type, Q)
        self.faces = []
        self.face_lm = []  # (pl, pm) of each face's half-range moments
        for f in range(2 * ndims):
            sb = surface_basis(self.mu, self.eta, self.xi, self.w, L, sn_type, Q, f, with_indices=True)
            self.faces.append(sb[:4])
            self.face_lm.append(sb[4:])
        self.Nps = self.faces[-1][0]
        if any(f[0] != self.Nps for f in self.faces):
            raise ValueError("faces with different half-range basis sizes are not supported")


@functools.lru_cache(maxsize=64)
def _angular_set(quad_type, quad_order, ndims, qdims, L, sn_type):
    return AngularSet(quad_type, quad_order, ndims, qdims, L, sn_type)


def angular_set(quad_type, quad_order, ndims, qdims, L, sn_type):
    """Cached AngularSet.  For the Galerkin types the bases are built to max(L, minimum order)
    (angular_polynomial_basis.jl:64 and the surface twin), so every L up to the smallest of those
    minima gives the same set; such L share one cache entry."""
    if sn_type in ("galerkin-m", "galerkin-d"):
        Q = quadrature_dimension(quad_type, ndims, qdims)
        mu, eta, xi, w = quadrature(quad_order, quad_type, ndims, Q)
        mo = galerkin_minimum_order(len(w), Q)
        for f in range(2 * ndims):  # directions entering each face, as surface_basis selects them
            a = (mu, eta, xi)[f // 2] if Q > 1 else mu
            ndp = int(np.count_nonzero(a > 0 if f % 2 == 0 else a < 0))
            mo = min(mo, galerkin_minimum_order(ndp, Q))
        L = max(L, mo)
    return _angular_set(quad_type, quad_order, ndims, qdims, L, sn_type)
