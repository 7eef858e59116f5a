"""MATXS libraries (cross_sections/read_matxs.jl, write_matxs.jl).

CPU:
- the reader against a hand-composed text file whose values are known (tests/golden), including
  the density scaling, the banded transfer matrices (a full band, a one-group band), a blank
  hollerith continuation word, the charge deposition from *char with 0 at the cut-off and the
  optional vectors defaulting to zero;
- the refusals the reference makes (material count, densities, missing required vectors,
  particle codes without a provided particle);
- write -> read round trips of synthetic one- and two-particle libraries in both encodings
  (six significant digits in text, the per-density doubles exact in binary), a second text round trip byte-identical;
- an FMAC-M library converted to MATXS and read back;
- the band of a transfer matrix with a gap and an empty outgoing group;
- a library read from a file feeding the oracle's solve.
GPU: solves with the hand-composed library against the oracle in 1D/2D/3D."""
import os
import sys

import numpy as np
import pytest

from util import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
HAND = os.path.join(HERE, "golden", "matxs_hand_1p2g2m.txt")
RHO = [2.0, 1.0]  # water, lead


def _X():
    from radiant_hip import matxs
    return matxs


def test_reader_hand_written_file():
    X = _X()
    lib = X.read_matxs(HAND, RHO)
    assert lib.particles == ["photon"] and lib.groups == [2] and lib.legendre_order == 1 and lib.Nmat == 2
    eq = lambda a, b: np.allclose(a, b, rtol=1e-15, atol=0)
    assert eq(lib.energy_boundaries[0], [1.0, 0.1, 0.01])
    assert eq(lib.total[0], [[1.2, 1.5], [1.8, 2.0]])             # [group, material], times density
    assert eq(lib.absorption[0], [[0.1, 0.2], [0.2, 0.3]])
    assert eq(lib.energy_deposition[0], [[0.2, 0.3], [0.4, 0.4]])
    assert eq(lib.energy_deposition_cutoff[0], [0.6, 0.5])
    assert eq(lib.charge_deposition[0], [[0.0, 0.01], [0.0, 0.02]])
    assert eq(lib.charge_deposition_cutoff[0], [0.0, 0.0])        # *char: zero at the cut-off
    assert not lib.momentum_transfer[0].any() and lib.stopping_powers[0].shape == (3, 2)
    assert not lib.stopping_powers[0].any()                       # optional vectors absent
    ref = np.zeros((2, 2, 2, 2))  # [material, from-group, to-group, l]
    ref[0, 0, 0] = [0.5, 0.1]
    ref[0, 1, 1] = [0.8, 0.2]
    ref[0, 0, 1] = [0.3, 0.0]
    ref[1, 0, 0] = [0.5, 0.1]
    ref[1, 1, 1] = [0.6, -0.1]
    assert eq(lib.scattering[(0, 0)], ref)
    sigt, sigs = lib.transport("photon", L=3)                     # zero above the file's order
    assert eq(sigt, lib.total[0]) and sigs.shape == (2, 2, 2, 4) and not sigs[..., 2:].any()


def test_reader_refusals(tmp_path):
    X = _X()
    with pytest.raises(ValueError, match="material count"):
        X.read_matxs(HAND, [1.0])
    with pytest.raises(ValueError, match="positive material densities"):
        X.read_matxs(HAND, [1.0, 0.0])
    with pytest.raises(ValueError, match="no matching provided particle"):
        X.read_matxs(HAND, RHO, particles=["electron"])
    text = open(HAND).read()
    bad = tmp_path / "no_tot0.txt"
    bad.write_text(text.replace("gtot0   ", "gtot1   "))
    with pytest.raises(ValueError, match=r"required vector '\*tot0'"):
        X.read_matxs(bad, RHO)
    bad2 = tmp_path / "no_charge.txt"
    bad2.write_text(text.replace("gchar   ", "gxxxx   "))
    with pytest.raises(ValueError, match="charge-deposition"):
        X.read_matxs(bad2, RHO)


def _synthetic(seed, npart=2, G=(3, 4), Nmat=2, L=2):
    from radiant_hip.fmac import Library
    rng = np.random.default_rng(seed)
    names = ["photon", "electron"][:npart]
    G = list(G[:npart])
    lib = Library(particles=names, groups=G, legendre_order=L, energy_boundaries=[], total=[], absorption=[],
                  stopping_powers=[], momentum_transfer=[], energy_deposition=[], charge_deposition=[])
    for n in range(npart):
        lib.energy_boundaries.append(np.sort(rng.uniform(0.01, 10.0, G[n] + 1))[::-1])
        lib.total.append(rng.uniform(0.1, 2.0, (G[n], Nmat)))
        lib.absorption.append(rng.uniform(0.0, 0.5, (G[n], Nmat)))
        charged = names[n] != "photon"
        lib.stopping_powers.append(rng.uniform(0.5, 3.0, (G[n] + 1, Nmat)) if charged else np.zeros((G[n] + 1, Nmat)))
        lib.momentum_transfer.append(rng.uniform(0.0, 1.0, (G[n], Nmat)) if charged else np.zeros((G[n], Nmat)))
        lib.energy_deposition.append(rng.uniform(0.0, 1.0, (G[n], Nmat)))
        lib.energy_deposition_cutoff.append(rng.uniform(0.0, 1.0, Nmat))
        lib.charge_deposition.append(rng.uniform(-1.0, 1.0, (G[n], Nmat)))
        lib.charge_deposition_cutoff.append(np.zeros(Nmat))  # *char carries no cut-off value
    for i in range(npart):
        for j in range(npart):
            a = rng.uniform(-0.2, 1.0, (Nmat, G[i], G[j], L + 1))
            a[rng.uniform(size=a.shape) < 0.3] = 0.0
            a[:, :, :, 0] = np.abs(a[:, :, :, 0])
            a[0, :, -1] = 0.0            # an outgoing group nothing scatters into
            a[1, 1, 0] = 0.0             # a gap inside a band
            lib.scattering[(i, j)] = a
    return lib


@pytest.mark.parametrize("npart", [1, 2])
@pytest.mark.parametrize("binary", [False, True])
def test_write_read_round_trip(tmp_path, npart, binary):
    X = _X()
    lib = _synthetic(40 + npart, npart=npart)
    rho = [2.5, 0.75]
    path = tmp_path / "lib.matxs"
    X.write_matxs(path, lib, rho, material_names=["water", "tissue"], binary=binary)
    got = X.read_matxs(path, rho)
    assert got.particles == lib.particles and got.groups == lib.groups and got.legendre_order == lib.legendre_order
    # binary carries the per-density doubles exactly: the only rounding is x / rho * rho (one ulp)
    tol = 4e-16 if binary else 5e-6  # six significant digits in the text encoding
    close = lambda a, b: np.allclose(a, b, rtol=tol, atol=1e-300)
    for n in range(npart):
        for name in ("energy_boundaries", "total", "absorption", "stopping_powers", "momentum_transfer",
                     "energy_deposition", "energy_deposition_cutoff", "charge_deposition", "charge_deposition_cutoff"):
            assert close(getattr(got, name)[n], getattr(lib, name)[n]), (name, n)
    for key, a in lib.scattering.items():
        assert close(got.scattering[key], a), key
        assert np.array_equal(got.scattering[key] == 0.0, a == 0.0)  # zeros stay zeros
    if not binary:  # a second round trip writes the same bytes
        path2 = tmp_path / "lib2.matxs"
        X.write_matxs(path2, got, rho, material_names=["water", "tissue"])
        assert open(path, "rb").read() == open(path2, "rb").read()
    else:
        assert open(path, "rb").read(4) != b" 0v "


@pytest.mark.parametrize("binary", [False, True])
def test_fmac_library_through_matxs(tmp_path, binary):
    """An FMAC-M library written as MATXS reads back as the same library (the conversion of
    Cross_Sections.jl:170-181): the hand-typed FMAC-M fixture's arrays, with its absent upscatter
    and truncated P0 block as zeros."""
    X = _X()
    from radiant_hip import fmac
    lib = fmac.read_fmac_m(os.path.join(HERE, "golden", "fmac_hand_2g2m.txt"))
    rho = [1.5, 11.35]
    path = tmp_path / "from_fmac.matxs"
    X.write_matxs(path, lib, rho, material_names=["m1", "m2"], binary=binary)
    got = X.read_matxs(path, rho, particles=lib.particles)
    tol = 4e-16 if binary else 5e-6
    close = lambda a, b: np.allclose(a, b, rtol=tol, atol=1e-300)
    assert got.groups == lib.groups and got.legendre_order == lib.legendre_order
    for name in ("energy_boundaries", "total", "absorption", "energy_deposition", "charge_deposition"):
        assert close(getattr(got, name)[0], getattr(lib, name)[0]), name
    assert close(got.scattering[(0, 0)], lib.scattering[(0, 0)])
    s1, s2 = got.transport(lib.particles[0], L=1), lib.transport(lib.particles[0], L=1)
    assert close(s1[0], s2[0]) and close(s1[1], s2[1])


def test_band():
    X = _X()
    a = np.zeros((5, 3, 2))
    a[1, 0, 0] = 1.0
    a[3, 0, 1] = 2.0          # groups 2..4 into group 1, a gap at group 3
    a[4, 2, 0] = 3.0          # group 5 only into group 3; nothing into group 2
    jband, ijj = X._band(a)
    assert jband == [3, 0, 1] and ijj == [4, 1, 5]


def test_library_feeds_the_solve(oracle):
    """A problem with the hand-composed library's cross sections converges in the oracle and
    matches the same arrays given directly."""
    X = _X()
    from radiant_hip import problems as P
    lib = X.read_matxs(HAND, RHO)
    p = P.small_problem(2, N=6, G=2, L=1, Nmat=2, tol=1e-12, seed=5)
    p.sigt, p.sigs = lib.transport("photon", L=p.Lxs)
    phi, st = oracle.solve(p)
    assert all(s["converged"] for s in st) and np.isfinite(phi).all()
    q = P.small_problem(2, N=6, G=2, L=1, Nmat=2, tol=1e-12, seed=5)
    q.sigt = np.array([[1.2, 1.5], [1.8, 2.0]])
    q.sigs = np.zeros_like(p.sigs)
    q.sigs[0, 0, 0, :2] = [0.5, 0.1]
    q.sigs[0, 1, 1, :2] = [0.8, 0.2]
    q.sigs[0, 0, 1, :2] = [0.3, 0.0]
    q.sigs[1, 0, 0, :2] = [0.5, 0.1]
    q.sigs[1, 1, 1, :2] = [0.6, -0.1]
    phq, _ = oracle.solve(q)
    assert rel_err(phi, phq) < 1e-14


@pytest.mark.gpu
def test_matxs_solve_parity(oracle):
    X = _X()
    from radiant_hip import problems as P, solver as S
    lib = X.read_matxs(HAND, RHO)
    for ndims in (1, 2, 3):
        p = P.small_problem(ndims, N=5 if ndims == 3 else 8, G=2, L=1, Nmat=2, tol=1e-12, seed=62)
        p.sigt, p.sigs = lib.transport("photon", L=p.Lxs)
        ref, _ = oracle.solve(p)
        phi, st = S.compute_flux(p)
        e = rel_err(phi, ref)
        print("matxs solve", ndims, e, flush=True)
        sys.stdout.flush()
        assert e < 1e-10 and all(s["converged"] for s in st)
