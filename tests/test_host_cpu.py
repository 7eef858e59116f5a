"""CPU tests of the host side: seeded generator, workload definitions, descriptor packing."""
import numpy as np
import pytest


def test_splitmix64_known_vector():
    """splitmix64 seeded with 0 (published reference outputs of Vigna's splitmix64.c)."""
    from radiant_hip.problems import SplitMix64
    r = SplitMix64(0)
    v = r.next_u64(3)
    assert [int(x) for x in v] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix64_stream_is_sequential():
    from radiant_hip.problems import SplitMix64
    a = SplitMix64(123).uniform(10)
    r = SplitMix64(123)
    b = np.concatenate([r.uniform(4), r.uniform(6)])
    assert np.array_equal(a, b) and np.all((a >= 0) & (a < 1))


def test_config0_definition():
    from radiant_hip import problems as P
    p = P.config0()
    assert (p.ndims, p.Nx, p.G, p.legendre_order, p.quad_order) == (1, 200, 10, 7, 16)
    src = np.nonzero(p.qvol[0])[0]
    assert src[0] == 79 and src[-1] == 119 and len(src) == 41       # cells 80..120 (1-based)
    assert np.all(p.qvol[1:] == 0)
    st = p.sigt[:, 0]
    assert np.all((st >= 0.1) & (st <= 1.0))
    s0 = p.sigs[0, :, :, 0]
    assert np.allclose(np.triu(s0), s0)                               # downscatter only
    c = s0.sum(axis=1) / st
    assert np.all((c >= 0.5) & (c <= 0.95))
    for l in range(8):
        assert np.allclose(p.sigs[0, :, :, l], 0.8 ** l * s0)


def test_config1_and_2_shapes():
    from radiant_hip import problems as P
    p1 = P.config1()
    assert (p1.Nx, p1.Ny, p1.G, p1.Nmat) == (128, 128, 20, 2)
    blocks = p1.mat[::16, ::16, 0]
    assert np.all(blocks[:-1] != blocks[1:]) and np.all(blocks[:, :-1] != blocks[:, 1:])
    assert np.count_nonzero(p1.qvol[0]) == 32 * 32
    p2 = P.config2()
    assert (p2.Nx, p2.Ny, p2.Nz, p2.G, p2.Nmat) == (64, 64, 64, 40, 3)
    assert np.count_nonzero(p2.qvol[0]) == 16 ** 3
    assert set(np.unique(p2.mat)) == {1, 2, 3}


def test_batch_and_stream_definitions():
    from radiant_hip import problems as P
    b = P.config3(n_problems=8)
    for p in b:
        assert 32 <= p.Nx <= 160 and p.Nx == p.Ny and 8 <= p.G <= 32 and 1 <= p.Lxs <= 5
        assert p.quad_order in (8, 10, 12, 14, 16)
    r = P.config4_request(3)
    assert 32 <= r.Nx <= 96 and 10 <= r.G <= 40 and r.quad_order in (8, 10, 12)
    t = P.poisson_arrivals(1000)
    assert abs(1000 / t[-1] - 50.0) < 5.0


def test_descriptor_packing():
    """rt_problem_desc built from a Problem points at arrays in Radiant's layouts."""
    from radiant_hip import problems as P
    from radiant_hip.solver import Desc
    p = P.small_problem(2, N=5, G=3, L=2)
    d = Desc(p)
    D = d.desc
    assert (D.ndims, D.Nx, D.Ny, D.Nz, D.G, D.Nmat, D.Nd, D.Np) == (2, 5, 5, 1, 3, 2, d.ang.Nd, d.ang.Np)
    assert D.mat[0] == p.mat[0, 0, 0] and D.mat[1] == p.mat[1, 0, 0] and D.mat[5] == p.mat[0, 1, 0]
    assert D.sigt[1] == p.sigt[1, 0] and D.sigt[3] == p.sigt[0, 1]
    assert D.sigs[1] == p.sigs[1, 0, 0, 0] and D.sigs[2] == p.sigs[0, 1, 0, 0]
    assert D.M[1] == d.ang.M[1, 0] and D.M[d.ang.Nd] == d.ang.M[0, 1]


def test_angular_cache_key_equivalence():
    """Every Legendre order below the Galerkin minimum orders shares one cached set; the shared set
    equals the one built directly for that order (angular_polynomial_basis.jl:64)."""
    import numpy as np
    from radiant_hip import angular as A
    direct = A.AngularSet("gauss-legendre-chebychev", 4, 2, 3, 2, "galerkin-d")
    cached = A.angular_set("gauss-legendre-chebychev", 4, 2, 3, 2, "galerkin-d")
    assert np.array_equal(direct.M, cached.M) and np.array_equal(direct.D, cached.D)
    assert np.array_equal(direct.pl, cached.pl)
    for a, b in zip(direct.faces, cached.faces):
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
