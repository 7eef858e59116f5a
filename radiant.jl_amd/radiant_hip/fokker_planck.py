"""Angular Fokker-Planck operator of the BFP and FP solver types, in moment space, and the
removal cross section each SN solver type uses: the part of Radiant's ``compute_flux(::SN)`` set-up
that runs before the flux loop (src/particle_transport/sn_flux.jl:69-112), so the caller side of
the C ABI, which receives Σtot, Σs, T and ℳ ready to use.

ℳ (``[Np, Np]``) acts on the flux moments of a cell: the pass adds ``T[g, mat] * ℳ @ Φ`` to the
source (fokker_planck_source.jl:26-36), and Σtot grows by ``T * λ₀`` so that the direction-space
operator ``ℳ_n`` shifted by ``λ₀`` has a non-negative diagonal.  Discretisations
(fokker_planck_scattering_matrix.jl:29-40):

* ``galerkin``  -- ``diag(-l(l+1))`` mapped to directions with M and D (fokker_planck_galerkin.jl);
* ``finite-difference`` -- Morel's moment-preserving three-point scheme on the mu levels in 1D;
  for the gauss-legendre-chebychev product set in 2D/3D the same polar operator plus a three-point
  azimuthal operator per level, periodic over the full circle (Qdims 3) or closed at the ends of
  the upper half (Qdims 2), scaled so that the first moments keep their eigenvalue
  (fokker_planck_finite_difference.jl); for the other 2D/3D sets (lebedev, carlson) a coupling
  γ_nm / w_n between nodes whose Voronoi cells on the unit sphere share an edge, with one γ per
  edge fixed by the first moments being eigenfunctions of eigenvalue -2 (tools/voronoi.jl);
* ``differential-quadrature`` -- ``(1 - mu²) W² - 2 mu W`` with W the Lagrange differentiation
  matrix on the 1D nodes (fokker_planck_differential_quadrature.jl).
"""
import numpy as np

SOLVER_TYPES = {"BTE": 1, "BFP": 2, "BCSD": 3, "FP": 4, "CSD": 5, "BFP-EF": 6}   # SN.jl get_solver_type


def solver_type_of(prob):
    """SN solver type number of a problem (1..6); None means BTE, or BCSD with slowing down."""
    st = getattr(prob, "solver_type", None)
    if st is None:
        return 3 if getattr(prob, "csd", False) else 1
    if st.upper() not in SOLVER_TYPES:
        raise ValueError("Unknown type of solver.")
    return SOLVER_TYPES[st.upper()]


def _shift_min_diag(Mn):
    """λ₀ = max(0, max_n -ℳ[n, n]) added on the diagonal."""
    lam = max(0.0, float(np.max(-np.diag(Mn))))
    return Mn + lam * np.eye(len(Mn)), lam


def _galerkin(ang):
    lfac = -(ang.pl * (ang.pl + 1)).astype(np.float64)
    Mn, lam = _shift_min_diag((ang.M * lfac[None, :]) @ ang.D)
    return Mn, lam


def _morel_polar(mu, w):
    """Morel's 1D operator on ascending nodes: coupling to the lower and upper neighbour."""
    N = len(mu)
    beta = np.zeros(N + 1)
    beta[1:N] = -2.0 * np.cumsum(w[:N - 1] * mu[:N - 1])
    lower = np.zeros(N)
    upper = np.zeros(N)
    lower[1:] = beta[1:N] / (w[1:] * np.diff(mu))
    upper[:-1] = beta[1:N] / (w[:-1] * np.diff(mu))
    return beta, lower, upper


def _fd_1d(mu, w):
    order = np.argsort(mu, kind="stable")
    m, ww = mu[order], w[order]
    _, lower, upper = _morel_polar(m, ww)
    lam = float(np.max(lower + upper))
    N = len(m)
    T = np.diag(lam - (lower + upper)) + np.diag(lower[1:], -1) + np.diag(upper[:-1], 1)
    Mn = np.zeros((N, N))
    Mn[np.ix_(order, order)] = T
    return Mn, lam


def _fd_glc(N, qdims):
    x, wx = np.polynomial.legendre.leggauss(N)
    order = np.argsort(x, kind="stable")
    mu, wi = x[order], wx[order]
    Na = N if qdims == 2 else 2 * N
    wj = np.pi / N
    om = np.pi * (np.arange(1, Na + 1) - 0.5) / N
    beta, lower, upper = _morel_polar(mu, wi)
    s = np.sqrt(1.0 - mu ** 2)
    d12 = np.zeros(N + 1)
    d12[1:N] = np.diff(s) / np.diff(mu)
    cn = (beta[1:] * d12[1:] - beta[:-1] * d12[:-1]) / wi
    gam = np.pi ** 2 * (2.0 * (1.0 - mu ** 2) + cn * s) / (2.0 * N ** 2 * (1.0 - np.cos(np.pi / N)))
    # azimuthal three-point operator (unit factor), periodic for the full circle
    A = np.zeros((Na, Na))
    if qdims == 2:
        gaps = np.diff(om)
        for i in range(Na):
            if i > 0:
                A[i, i - 1] += 1.0 / gaps[i - 1]
            if i < Na - 1:
                A[i, i + 1] += 1.0 / gaps[i]
    else:
        gaps = np.append(np.diff(om), om[0] + 2 * np.pi - om[-1])   # gap i: between i and i+1 (mod Na)
        for i in range(Na):
            A[i, (i + 1) % Na] += 1.0 / gaps[i]
            A[i, (i - 1) % Na] += 1.0 / gaps[i - 1]
    A -= np.diag(A.sum(axis=1))
    Nd = N * Na
    Mn = np.zeros((Nd, Nd))
    for n in range(N):
        blk = slice(n * Na, (n + 1) * Na)
        Mn[blk, blk] += gam[n] / ((1.0 - mu[n] ** 2) * wj) * A
        Mn[blk, blk] -= (lower[n] + upper[n]) * np.eye(Na)
        if n > 0:
            Mn[blk, slice((n - 1) * Na, n * Na)] += lower[n] * np.eye(Na)
        if n < N - 1:
            Mn[blk, slice((n + 1) * Na, (n + 2) * Na)] += upper[n] * np.eye(Na)
    lam = float(np.max(np.abs(np.diag(Mn))))
    return Mn + lam * np.eye(Nd), lam


def _dq_1d(mu):
    N = len(mu)
    diff = mu[:, None] - mu[None, :]
    np.fill_diagonal(diff, 1.0)
    Pi = np.prod(diff, axis=1)
    W = Pi[:, None] / (diff * Pi[None, :])
    np.fill_diagonal(W, 0.0)
    np.fill_diagonal(W, -W.sum(axis=1))
    H = np.diag(mu)
    return _shift_min_diag((np.eye(N) - H @ H) @ W @ W - 2.0 * H @ W)


def voronoi_edges(P, tol=1e-7):
    """Pairs (n, m), n < m, of nodes on the unit sphere (P [Nd, 3]) whose Voronoi cells share an
    edge (the neighbours that voronoi.jl:13-205 finds from vertex triplets).  For nodes on a sphere
    the Delaunay triangulation is the convex hull, and each hull facet's outward normal is the
    Voronoi vertex of its three nodes; facets of nodes on one circle (coplanar facets) are one
    vertex, so their diagonals are not edges.  Two cells share an edge exactly when they share two
    vertices."""
    from scipy.spatial import ConvexHull
    hull = ConvexHull(P)
    nrm = hull.equations[:, :3]
    vid = np.empty(len(nrm), dtype=np.int64)
    reps = []
    for f, v in enumerate(nrm):
        for r, rv in enumerate(reps):
            if np.max(np.abs(rv - v)) < tol:
                vid[f] = r
                break
        else:
            reps.append(v)
            vid[f] = len(reps) - 1
    verts = [set() for _ in range(len(P))]
    for f, tri in enumerate(hull.simplices):
        for n in tri:
            verts[n].add(int(vid[f]))
    edges = set()
    for tri in hull.simplices:
        for a in range(3):
            for b in range(a + 1, 3):
                n, m = sorted((int(tri[a]), int(tri[b])))
                if len(verts[n] & verts[m]) >= 2:
                    edges.add((n, m))
    return sorted(edges)


def _fd_voronoi(ang, quad_type, quad_order, qdims):
    """fokker_planck_finite_difference.jl:139-179 (qdims 2) and :255-275 (qdims 3) with
    fokker_planck_weights_{2D,3D} of voronoi.jl: unknowns γ per Voronoi edge with
    sum_m γ_nm (Ω_k[m] - Ω_k[n]) = -2 w_n Ω_k[n] for every node and axis (least squares, full
    column rank required), ℳ[n, m] = γ_nm / w_n, rows summing to zero, shifted by max |ℳ[n, n]|.
    In 2D the edges of the full-sphere rule between upper-hemisphere nodes, mapped onto the folded set."""
    from .angular import quadrature
    mu, eta, xi, w = (np.asarray(a, float) for a in (ang.mu, ang.eta, ang.xi, ang.w))
    Nd = len(w)
    X3, Y3, Z3, _ = quadrature(quad_order, quad_type, 3, 3)
    if qdims == 3:
        if len(X3) != Nd:
            raise ValueError("Fokker-Planck operator does not match the quadrature")
        to2 = np.arange(Nd)
        up = np.ones(Nd, bool)
    else:
        d = (X3[:, None] - mu[None]) ** 2 + (Y3[:, None] - eta[None]) ** 2 + (np.abs(Z3)[:, None] - xi[None]) ** 2
        to2 = np.argmin(d, axis=1)
        up = Z3 >= 0
    pairs = set()
    for n3, m3 in voronoi_edges(np.stack([X3, Y3, Z3], 1)):
        if up[n3] and up[m3] and to2[n3] != to2[m3]:
            pairs.add((min(to2[n3], to2[m3]), max(to2[n3], to2[m3])))
    pairs = sorted(pairs)
    O = np.stack([mu, eta, xi], 1)
    G = np.zeros((3 * Nd, len(pairs)))
    Q = (-2.0 * w[:, None] * O).ravel()
    for e, (n, m) in enumerate(pairs):
        G[3 * n:3 * n + 3, e] += O[m] - O[n]
        G[3 * m:3 * m + 3, e] += O[n] - O[m]
    pG = np.linalg.pinv(G)
    if np.linalg.norm(pG @ G - np.eye(len(pairs))) > 1e-7:
        raise ValueError("The equation system to find γ-parameters does not have a solution.")
    gam = pG @ Q
    Mn = np.zeros((Nd, Nd))
    for e, (n, m) in enumerate(pairs):
        Mn[n, m] += gam[e] / w[n]
        Mn[m, n] += gam[e] / w[m]
        Mn[n, n] -= gam[e] / w[n]
        Mn[m, m] -= gam[e] / w[m]
    lam = float(np.max(np.abs(np.diag(Mn))))
    return Mn + lam * np.eye(Nd), lam


def direction_operator(ang, method, quad_type, quad_order, qdims):
    """(ℳ_n [Nd, Nd] in the direction order of the angular set, λ₀)."""
    if method == "galerkin":
        return _galerkin(ang)
    if method == "finite-difference":
        if qdims == 1:
            return _fd_1d(np.asarray(ang.mu, float), np.asarray(ang.w, float))
        if quad_type == "gauss-legendre-chebychev":
            Mn, lam = _fd_glc(quad_order, qdims)
            if Mn.shape[0] != ang.Nd:
                raise ValueError("Fokker-Planck operator does not match the quadrature")
            return Mn, lam
        return _fd_voronoi(ang, quad_type, quad_order, qdims)
    if method == "differential-quadrature":
        if qdims != 1:
            raise ValueError("Differential quadrature unavailable in multidimensional geometry.")
        return _dq_1d(np.asarray(ang.mu, float))
    raise ValueError("Unknown method to treat the Fokker-Planck term.")


def moment_operator(ang, method, quad_type, quad_order, qdims):
    """(ℳ = D ℳ_n M [Np, Np], λ₀) (fokker_planck_scattering_matrix.jl:29-40)."""
    Mn, lam = direction_operator(ang, method, quad_type, quad_order, qdims)
    return ang.D @ Mn @ ang.M, lam


def prepare(prob, ang, qdims):
    """What sn_flux.jl:69-112 hands to the flux loop: (solver type, Σtot [G, Nmat], Σs, T or
    None, ℳ or None).  FP and CSD remove with the absorption and drop Boltzmann scattering; BFP
    and FP add T λ₀ to the removal; BFP-EF removes the self-scattering moment l = 0."""
    st = solver_type_of(prob)
    csd = bool(getattr(prob, "csd", False))
    if (st >= 2) != csd:
        raise ValueError("solver types BFP, BCSD, FP, CSD and BFP-EF carry the slowing-down term (csd), BTE does not")
    sigs = np.asarray(prob.sigs, np.float64)
    if st in (4, 5):
        if getattr(prob, "siga", None) is None:
            raise ValueError("solver types FP and CSD need the absorption cross sections")
        sigt = np.array(prob.siga, np.float64)
        sigs = np.zeros_like(sigs)
    else:
        sigt = np.array(prob.sigt, np.float64)
    T = Mfp = None
    if st in (2, 4):
        if getattr(prob, "T", None) is None:
            raise ValueError("solver types BFP and FP need the momentum transfer")
        T = np.asarray(prob.T, np.float64)
        Mfp, lam = moment_operator(ang, getattr(prob, "fokker_planck", None) or "finite-difference",
                                   prob.quad_type, prob.quad_order, qdims)
        sigt = sigt + T * lam
    if st == 6:
        for m in range(prob.Nmat):
            sigt[:, m] -= np.diag(sigs[m, :, :, 0])
    return st, sigt, sigs, T, Mfp
