"""FMAC-M libraries (cross_sections/read_fmac_m.jl, write_fmac_m.jl).

CPU:
- the reader against a hand-composed file whose values are known (tests/golden);
- the writer's number format against the rules of write_fmac_m.jl:283-313;
- write -> read round trips of synthetic libraries (six significant digits);
- a library read from a file feeding the oracle's solve.
GPU: a solve with a library read from an FMAC-M file against the oracle."""
import os
import sys

import numpy as np
import pytest

from util import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
HAND = os.path.join(HERE, "golden", "fmac_hand_2g2m.txt")


def _F():
    from radiant_hip import fmac
    return fmac


def test_reader_hand_written_file():
    F = _F()
    lib = F.read_fmac_m(HAND)
    assert lib.particles == ["photon"] and lib.groups == [2] and lib.legendre_order == 1 and lib.Nmat == 2
    assert lib.title == "Hand-written test library"
    assert np.allclose(lib.energy_boundaries[0], [1.0, 0.1, 0.01], rtol=1e-15)
    sigt, sigs = lib.transport("photon")
    assert np.allclose(sigt, [[1.5, 0.75], [2.0, 1.25]], rtol=1e-15)
    ref = np.zeros((2, 2, 2, 2))  # [material, from-group, to-group, l]
    ref[0, 0, 0] = [0.5, 0.1]
    ref[1, 0, 0] = [0.25, 0.05]
    ref[0, 0, 1] = [0.3, 0.0]       # truncated to P0
    ref[0, 1, 1] = [0.8, 0.2]
    ref[1, 1, 1] = [0.6, -0.1]      # negative P1 moment
    assert np.allclose(sigs, ref, rtol=1e-15, atol=0)
    # absent records read as zeros, as the reference leaves them
    assert not lib.absorption[0].any() and not lib.stopping_powers[0].any()
    # a higher requested order is zero-padded (sn_flux.jl:79-83 via get_scattering)
    _, s3 = lib.transport("photon", L=3)
    assert s3.shape == (2, 2, 2, 4) and np.array_equal(s3[..., :2], sigs) and not s3[..., 2:].any()


@pytest.mark.parametrize("x,s", [(1.0, " 1.00000E+00"), (-0.000123456, "-1.23456E-04"), (0.0, " 0.00000E+00"),
                                 (123456.7, " 1.23457E+05"), (2.5e-7, " 2.50000E-07"), (-7.0, "-7.00000E+00"),
                                 (9.999996, " 10.0000E+00"), (1e-120, " 0.00000E-00"), (3.25e11, " 3.25000E+11")])
def test_writer_number_format(x, s):
    assert _F()._real_field(x) == s


def _synthetic(seed, npart=2, G=(4, 3), Nmat=3, L=3):
    F = _F()
    rng = np.random.default_rng(seed)
    names = ["photon", "electron", "positron"][:npart]
    G = list(G[:npart])
    lib = F.Library(particles=names, groups=G, legendre_order=L, energy_boundaries=[], total=[], absorption=[],
                    stopping_powers=[], momentum_transfer=[], energy_deposition=[], charge_deposition=[])
    for n in range(npart):
        lib.energy_boundaries.append(np.sort(rng.uniform(0.01, 10.0, G[n] + 1))[::-1])
        lib.total.append(rng.uniform(0.1, 2.0, (G[n], Nmat)))
        lib.absorption.append(np.zeros((G[n], Nmat)))  # not written by the reference writer
        lib.stopping_powers.append(rng.uniform(0.5, 3.0, (G[n] + 1, Nmat)))
        lib.momentum_transfer.append(rng.uniform(0.0, 1.0, (G[n], Nmat)))
        lib.energy_deposition.append(rng.uniform(0.0, 1.0, (G[n], Nmat)))
        lib.charge_deposition.append(rng.uniform(-1.0, 1.0, (G[n], Nmat)))
    for i in range(npart):
        for j in range(npart):
            a = rng.uniform(-0.2, 1.0, (Nmat, G[i], G[j], L + 1))
            a[rng.uniform(size=a.shape) < 0.3] = 0.0       # zero moments anywhere
            a[:, :, :, 0] = np.abs(a[:, :, :, 0])
            a[0, 0, -1] = 0.0                              # a pair with no scattering
            if L >= 2:
                a[1, -1, 0, 2:] = 0.0                      # a truncated moment list
            lib.scattering[(i, j)] = a
    return lib


@pytest.mark.parametrize("npart", [1, 2])
def test_write_read_round_trip(tmp_path, npart):
    F = _F()
    lib = _synthetic(7 + npart, npart=npart)
    path = tmp_path / "lib.fmac"
    F.write_fmac_m(path, lib)
    got = F.read_fmac_m(path)
    assert got.particles == lib.particles and got.groups == lib.groups and got.legendre_order == lib.legendre_order
    close = lambda a, b: np.allclose(a, b, rtol=5e-6, atol=0)  # six significant digits in the file
    for n in range(npart):
        for name in ("energy_boundaries", "total", "stopping_powers", "momentum_transfer", "energy_deposition",
                     "charge_deposition"):
            assert close(getattr(got, name)[n], getattr(lib, name)[n]), name
    for key, a in lib.scattering.items():
        assert close(got.scattering[key], a), key
        assert np.array_equal(got.scattering[key] == 0.0, a == 0.0)  # zeros stay zeros
    # a second round trip is exact: the values already have six digits
    path2 = tmp_path / "lib2.fmac"
    F.write_fmac_m(path2, got)
    assert open(path).read() == open(path2).read()


def test_library_feeds_the_solve(oracle, tmp_path):
    """A problem whose cross sections come from an FMAC-M file solves like the same arrays given
    directly (the oracle on both)."""
    F = _F()
    from radiant_hip import problems as P
    lib = _synthetic(21, npart=1, G=(3,), Nmat=2, L=2)
    for g in range(3):  # keep the medium subcritical: scattering below the total
        lib.scattering[(0, 0)][:, g, :, 0] *= 0.5 / max(1e-30, lib.scattering[(0, 0)][:, g, :, 0].sum() /
                                                         lib.total[0][g].min())
    path = tmp_path / "lib.fmac"
    F.write_fmac_m(path, lib)
    got = F.read_fmac_m(path)
    p = P.small_problem(2, N=7, G=3, L=2, Nmat=2, tol=1e-12, seed=3)
    p.sigt, p.sigs = got.transport("photon", L=p.Lxs)
    phi, st = oracle.solve(p)
    assert all(s["converged"] for s in st) and np.isfinite(phi).all()
    q = P.small_problem(2, N=7, G=3, L=2, Nmat=2, tol=1e-12, seed=3)
    q.sigt, q.sigs = lib.transport("photon", L=q.Lxs)
    phq, _ = oracle.solve(q)
    assert rel_err(phi, phq) < 1e-4  # the file keeps six digits


@pytest.mark.gpu
def test_fmac_solve_parity(oracle):
    F = _F()
    from radiant_hip import problems as P, solver as S
    lib = F.read_fmac_m(HAND)
    for ndims in (1, 2, 3):
        p = P.small_problem(ndims, N=5 if ndims == 3 else 8, G=2, L=1, Nmat=2, tol=1e-12, seed=61)
        p.sigt, p.sigs = lib.transport("photon", L=p.Lxs)
        ref, _ = oracle.solve(p)
        phi, st = S.compute_flux(p)
        e = rel_err(phi, ref)
        print("fmac solve", ndims, e, flush=True)
        sys.stdout.flush()
        assert e < 1e-10 and all(s["converged"] for s in st)
