"""Seeded synthetic SN problems in Radiant's array layouts.

The workloads are the fixed-source photon (BTE) problems of the benchmark suite: multigroup
cross-sections drawn from a seeded splitmix64 stream, downscatter-only Legendre scattering,
Cartesian voxel grids, box volume sources and void boundaries.

Layouts follow the arrays Radiant hands to its solver (column-major, 1-based materials):

* ``sigt[g, m]``            -- Σtot[Ng, Nmat]               (src/particle_transport/sn_flux.jl:69-74)
* ``sigs[m, gi, gf, l]``    -- Σs[Nmat, Ng, Ng, L+1]         (src/structures/Cross_Sections.jl:928-942)
* ``mat[ix, iy, iz]``       -- material per voxel, 1-based   (src/particle_transport/geometry.jl:240-246)
* ``qvol[g, c]``            -- volume source of moment p=1   (src/structures/Source.jl:132-137)

Scattering-moment convention: Radiant stores Σs_l without the (2l+1) factor (the factor lives in
M).  The Henyey-Greenstein-like kernel σ_l = (2l+1) g^l σ_s is therefore entered as
Σs_l = g^l σ_s, the same physical kernel in Radiant's normalisation.
"""
from dataclasses import dataclass, field

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """splitmix64 stream; ``uniform(n)`` returns the next n draws in [0, 1)."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLD
            self.state = self.state + np.uint64(n) * _GOLD
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n=None):
        m = 1 if n is None else n
        u = (self.next_u64(m) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return float(u[0]) if n is None else u

    def randint(self, lo, hi, n=None):
        """Integers uniform in {lo..hi}."""
        u = self.uniform(1 if n is None else n)
        v = lo + np.floor(u * (hi - lo + 1)).astype(np.int64)
        v = np.minimum(v, hi)
        return int(v[0]) if n is None else v


@dataclass
class Problem:
    ndims: int
    Nx: int
    Ny: int
    Nz: int
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    mat: np.ndarray          # int64 [Nx, Ny, Nz], 1-based
    Nmat: int
    G: int
    sigt: np.ndarray         # [G, Nmat]
    sigs: np.ndarray         # [Nmat, G, G, Lxs+1]
    Lxs: int
    qvol: np.ndarray         # [G, Nc] (c = ix + Nx*(iy + Ny*iz))
    bc: np.ndarray           # int32 [2*ndims]
    quad_type: str
    quad_order: int
    qdims: int = 0           # 0 = Radiant's default for the quadrature/geometry pair
    legendre_order: int = 0
    sn_type: str = "galerkin-d"
    accel: str = "gmres"
    restart: int = 30
    depth: int = 3
    maxit: int = 1000
    tol: float = 1e-10
    name: str = ""
    meta: dict = field(default_factory=dict)
    # surface sources: face (0..5 = X-, X+, Y-, Y+, Z-, Z+) -> moments [G, Np_source, npos]
    # (radiant_hip.sources.add_surface_source builds them)
    ssrc: dict = None
    Np_source: int = 0
    # continuous slowing down (SN solver "BCSD", DD of order 1 in energy): stopping powers at the
    # group boundaries Sb [G+1, Nmat] and group widths dE [G]
    csd: bool = False
    Sb: np.ndarray = None
    dE: np.ndarray = None
    # SN solver type (SN.set_solver_type): None = "BTE", or "BCSD" when csd is set; "BFP", "FP",
    # "CSD" and "BFP-EF" carry the slowing-down term too.  T: momentum transfer [G, Nmat] (BFP,
    # FP), siga: absorption [G, Nmat] (the removal of FP and CSD), fokker_planck: the angular
    # Fokker-Planck discretisation (SN.set_angular_fokker_planck)
    solver_type: str = None
    T: np.ndarray = None
    siga: np.ndarray = None
    fokker_planck: str = "finite-difference"
    # spatial closure along x (SN.set_scheme("x", scheme, order)): "DD", "DG", "DG-" or "DG+" (the
    # last two 1D BTE only) and the number of spatial Legendre moments per cell along x
    scheme: str = "DD"
    scheme_order: int = 1
    # the y axis of 2D problems ("DD" or "DG") and the coupling of the high-order moments
    # (SN.set_is_full_coupling; without it a cell holds Ox + Oy - 1 moments instead of Ox * Oy)
    scheme_y: str = "DD"
    scheme_order_y: int = 1
    full_coupling: bool = True
    scheme_z: str = "DD"     # the z axis of 3D problems
    scheme_order_z: int = 1
    # the energy axis of the slowing-down types (SN.set_scheme("E", scheme, order)): "DD" or "DG";
    # a cell then holds the moments over the spatial axes and energy, energy fastest
    scheme_e: str = "DD"
    scheme_order_e: int = 1

    @property
    def Nc(self):
        return self.Nx * self.Ny * self.Nz


def multigroup_xs(rng, G, L, ghg=0.8, st_range=(0.1, 1.0), c_range=(0.5, 0.95)):
    """One material: Σt_g ~ U[st], c_g ~ U[c], downscatter-only row-normalised U[0,1] weights,
    Σs_l[g->g'] = ghg^l c_g Σt_g W[g,g']  (Radiant convention, see module docstring).
    Returns (sigt[G], sigs[G(from), G(to), L+1])."""
    st = st_range[0] + (st_range[1] - st_range[0]) * rng.uniform(G)
    c = c_range[0] + (c_range[1] - c_range[0]) * rng.uniform(G)
    W = np.zeros((G, G))
    for g in range(G):
        row = rng.uniform(G - g)
        W[g, g:] = row / row.sum()
    s0 = (c * st)[:, None] * W
    sigs = np.zeros((G, G, L + 1))
    for l in range(L + 1):
        sigs[:, :, l] = (ghg ** l) * s0
    return st, sigs


def _pack_xs(mats, G, L):
    Nmat = len(mats)
    sigt = np.zeros((G, Nmat))
    sigs = np.zeros((Nmat, G, G, L + 1))
    for m, (st, ss) in enumerate(mats):
        sigt[:, m] = st
        sigs[m] = ss
    return sigt, sigs


def _centres(N, width):
    """Voxel centres as Radiant accumulates them (geometry.jl:41-49)."""
    d = width / N
    xp = 0.0
    x = np.zeros(N)
    for i in range(N):
        xp += d
        x[i] = xp - d / 2
    return x, np.full(N, d)


def _box_source(G, group, axes, bounds, intensity=1.0):
    """Box volume source on cell centres, inclusive bounds (volume_source.jl:50)."""
    masks = [(x >= lo) & (x <= hi) for x, (lo, hi) in zip(axes, bounds)]
    m = masks[0]
    for mk in masks[1:]:
        m = np.multiply.outer(m, mk)
    nc = m.size
    q = np.zeros((G, nc))
    q[group, :] = intensity * m.reshape(-1, order="F")
    return q


def config0(seed=20260101):
    """1D slab 10 cm, 200 cells, 10 groups, P7, Gauss-Legendre S16, source in cells 80-120."""
    rng = SplitMix64(seed)
    G, L = 10, 7
    x, dx = _centres(200, 10.0)
    sigt, sigs = _pack_xs([multigroup_xs(rng, G, L)], G, L)
    return Problem(ndims=1, Nx=200, Ny=1, Nz=1, dx=dx, dy=np.ones(1), dz=np.ones(1),
                   mat=np.ones((200, 1, 1), np.int64), Nmat=1, G=G, sigt=sigt, sigs=sigs, Lxs=L,
                   qvol=_box_source(G, 0, [x], [(3.95, 6.0)]), bc=np.zeros(2, np.int32),
                   quad_type="gauss-legendre", quad_order=16, qdims=1, legendre_order=L, name="C0")


def config_2d(N=128, width=10.0, G=20, L=5, quad_order=8, block=16, seed=20260102, name="C1"):
    """2D checkerboard of two materials in block x block cells, central-quarter source."""
    rng = SplitMix64(seed)
    x, dx = _centres(N, width)
    y, dy = _centres(N, width)
    mats = [multigroup_xs(rng, G, L) for _ in range(2)]
    sigt, sigs = _pack_xs(mats, G, L)
    bi = (np.arange(N) // block)
    mat = (((bi[:, None] + bi[None, :]) % 2) + 1).astype(np.int64).reshape(N, N, 1)
    lo, hi = 0.375 * width, 0.625 * width
    return Problem(ndims=2, Nx=N, Ny=N, Nz=1, dx=dx, dy=dy, dz=np.ones(1), mat=mat, Nmat=2, G=G, sigt=sigt,
                   sigs=sigs, Lxs=L, qvol=_box_source(G, 0, [x, y], [(lo, hi), (lo, hi)]),
                   bc=np.zeros(4, np.int32), quad_type="gauss-legendre-chebychev", quad_order=quad_order,
                   qdims=0, legendre_order=L, name=name)


def config1(seed=20260102):
    """2D 128x128 on 10x10 cm, 16x16-cell checkerboard, 20 groups, P5, GLC 8 (128 directions)."""
    return config_2d(seed=seed)


def config_3d(N=64, width=6.4, G=40, L=3, quad_order=12, Nmat=3, seed=20260103, name="C2"):
    """3D cube with a seeded random voxel map of Nmat materials and a central (N/4)^3 source."""
    rng = SplitMix64(seed)
    x, dx = _centres(N, width)
    mats = [multigroup_xs(rng, G, L) for _ in range(Nmat)]
    sigt, sigs = _pack_xs(mats, G, L)
    mat = (1 + np.minimum(np.floor(rng.uniform(N ** 3) * Nmat), Nmat - 1)).astype(np.int64)
    mat = mat.reshape((N, N, N), order="F")
    h = width / N
    q = N // 4
    lo = (N // 2 - q // 2) * h          # cell edges: centres of the central block lie strictly inside
    hi = (N // 2 + q // 2) * h
    return Problem(ndims=3, Nx=N, Ny=N, Nz=N, dx=dx, dy=dx.copy(), dz=dx.copy(), mat=mat, Nmat=Nmat, G=G,
                   sigt=sigt, sigs=sigs, Lxs=L, qvol=_box_source(G, 0, [x, x, x], [(lo, hi)] * 3),
                   bc=np.zeros(6, np.int32), quad_type="carlson", quad_order=quad_order, qdims=3,
                   legendre_order=L, name=name)


def config2(seed=20260103):
    """3D 64^3, 40 groups, Carlson S12 (168 directions), P3, 3 random materials."""
    return config_3d(seed=seed)


def config3(n_problems=256, seed=20260104, quad_orders=(8, 10, 12, 14, 16)):
    """Batch of uneven 2D problems: N~U{32..160}, G~U{8..32}, L~U{1..5}, GLC order drawn from
    quad_orders.  With Galerkin-D on the full sphere, GLC 12, 14 and 16 give cond(D) ~ 1e18 (D is
    numerically singular, so M = inv(D) is not determined in FP64); pass quad_orders=(8, 10) for a
    batch whose members are well posed."""
    out = []
    for i in range(n_problems):
        r = SplitMix64(seed * 1000003 + i)
        N = r.randint(32, 160)
        G = r.randint(8, 32)
        L = r.randint(1, 5)
        qo = list(quad_orders)[r.randint(0, len(quad_orders) - 1)]
        p = config_2d(N=N, G=G, L=L, quad_order=qo, block=max(1, N // 8), seed=seed * 7919 + i, name=f"C3[{i}]")
        out.append(p)
    return out


def config4_request(i, seed=20260105):
    """One request of the served stream: N~U{32..96} per axis, G~U{10..40}, Carlson 8/10/12, P3."""
    r = SplitMix64(seed * 1000003 + i)
    N = r.randint(32, 96)
    G = r.randint(10, 40)
    qo = [8, 10, 12][r.randint(0, 2)]
    return config_3d(N=N, width=0.1 * N, G=G, L=3, quad_order=qo, Nmat=3, seed=seed * 7919 + i, name=f"C4[{i}]")


def poisson_arrivals(n, rate=50.0, seed=20260106):
    """Exponential inter-arrival times (s) of a Poisson stream."""
    u = SplitMix64(seed).uniform(n)
    return np.cumsum(-np.log1p(-u) / rate)


def small_problem(ndims, N=6, G=3, L=2, quad_type=None, quad_order=None, seed=7, Nmat=2, accel="gmres",
                  tol=1e-12, bc=None, source_group=0, dims=None):
    """Tiny seeded problem for unit tests (oracle runs in milliseconds): N cells of width 1 per
    axis, or dims = (Nx, Ny, Nz) cells for a non-cubic grid."""
    rng = SplitMix64(seed)
    if quad_type is None:
        quad_type = {1: "gauss-legendre", 2: "gauss-legendre-chebychev", 3: "carlson"}[ndims]
    if quad_order is None:
        quad_order = {1: 8, 2: 4, 3: 4}[ndims]
    n_ax = [N] * ndims if dims is None else [int(v) for v in dims[:ndims]]
    axes = []
    ws = []
    for a in range(ndims):
        x, d = _centres(n_ax[a], 1.0 * n_ax[a])
        axes.append(x)
        ws.append(d)
    mats = [multigroup_xs(rng, G, L) for _ in range(Nmat)]
    sigt, sigs = _pack_xs(mats, G, L)
    shape = n_ax + [1] * (3 - ndims)
    mat = (1 + np.minimum(np.floor(rng.uniform(int(np.prod(shape))) * Nmat), Nmat - 1)).astype(np.int64)
    mat = mat.reshape(shape, order="F")
    q = _box_source(G, source_group, axes, [(0.3 * n, 0.7 * n) for n in n_ax])
    return Problem(ndims=ndims, Nx=shape[0], Ny=shape[1], Nz=shape[2], dx=ws[0],
                   dy=ws[1] if ndims >= 2 else np.ones(1), dz=ws[2] if ndims >= 3 else np.ones(1),
                   mat=mat, Nmat=Nmat, G=G, sigt=sigt, sigs=sigs, Lxs=L, qvol=q,
                   bc=np.zeros(2 * ndims, np.int32) if bc is None else np.asarray(bc, np.int32),
                   quad_type=quad_type, quad_order=quad_order, qdims=1 if quad_type == "gauss-legendre" else 0,
                   legendre_order=L, accel=accel, tol=tol, name=f"small{ndims}d")


CONFIGS = {"C0": config0, "C1": config1, "C2": config2}
