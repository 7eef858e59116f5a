"""3D scan sweep with each direction's pencils split by z over 2 or 4 workgroups
(sn_sweep_3D.jl:57-240 order-1 path; the split changes the schedule only).

A slab hands its top layer's outgoing z-faces to the slab above through device memory as tagged
words, so the split sweep must give the same bits as one workgroup per direction: solves with the
split forced (RADIANT_HIP_SCAN_SPLIT=2, 4) equal the unsplit ones exactly, in FP64 (faithful state)
and mixed FP32, on grids whose slabs are uneven, thinner than the courier's lag, or one layer
thick, and the forced split matches the oracle like the unsplit sweep.  Transport passes repeated
on one problem check that the tags of earlier launches never pass for current words."""
import os
import sys

import numpy as np
import pytest

from util import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _split:
    def __init__(self, v):
        self.v = v

    def __enter__(self):
        self.old = os.environ.get("RADIANT_HIP_SCAN_SPLIT")
        os.environ["RADIANT_HIP_SCAN_SPLIT"] = str(self.v)

    def __exit__(self, *a):
        if self.old is None:
            os.environ.pop("RADIANT_HIP_SCAN_SPLIT", None)
        else:
            os.environ["RADIANT_HIP_SCAN_SPLIT"] = self.old


# (Nx, Ny, Nz): two pencils per wave needs Nx <= 64
GRIDS = [
    (16, 12, 10),   # S = 4: slabs of 3, 3, 3, 1 layers
    (33, 5, 9),     # Ny below the courier's lag; S = 4 leaves no top slab (not offered), S = 2: 5 + 4
    (64, 8, 12),    # two cells per lane
    (8, 20, 8),     # S = 4: slabs of 2 layers
]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("precision", ["fp64", "mixed-fp32"])
def test_split_equals_unsplit(dims, precision):
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(3, dims=dims, G=2, L=2, seed=811 + dims[0], tol=1e-12)
    state = "faithful" if precision == "fp64" else "reduced"
    try:
        S.autotune(0)
        with _split(1):
            ref, rst = S.compute_flux(prob, precision=precision, state=state)
        for nsl in (2, 4):
            with _split(nsl):
                got, st = S.compute_flux(prob, precision=precision, state=state)
            d = float(np.abs(got - ref).max())
            _say("split", nsl, dims, precision, "max diff", d, [s["niter"] for s in st])
            assert np.array_equal(got, ref), (nsl, dims, precision)
            assert [s["niter"] for s in st] == [s["niter"] for s in rst]
    finally:
        S.autotune(1)


def test_split_matches_oracle(oracle):
    """Carlson S8 on a 14 x 10 x 13 grid with the sweep split in 4: FP64 solve against the oracle."""
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(3, dims=(14, 10, 13), G=2, L=2, seed=829, tol=1e-12, quad_order=8)
    ref, rst = oracle.solve(prob, nthreads=8)
    try:
        S.autotune(0)
        with _split(4):
            phi, st = S.compute_flux(prob)
    finally:
        S.autotune(1)
    e = rel_err(phi, ref)
    _say("split 4 vs oracle", e, [s["niter"] for s in st], [s["niter"] for s in rst])
    assert e < 1e-10
    assert [s["niter"] for s in st] == [s["niter"] for s in rst]


def test_split_repeated_passes():
    """Many transport passes of one problem, alternating 2 and 4 slabs and the unsplit sweep:
    every pass equals the unsplit one bit for bit (the handed words of each launch carry a new
    tag; words of the previous launches are never taken)."""
    from radiant_hip import problems as P, solver as S
    prob = P.small_problem(3, dims=(24, 9, 16), G=1, L=2, seed=833)
    h = S.ResidentProblem(prob)
    try:
        S.autotune(0)
        Sl = h.state_length()
        rng = np.random.default_rng(5)
        for it in range(12):
            z = rng.standard_normal(Sl)
            q = np.zeros(len(h.d.ang.pl) * prob.Nc)
            with _split(1):
                ref = h.transport_pass(0, 1, q, z)
            with _split(2 if it % 2 == 0 else 4):
                got = h.transport_pass(0, 1, q, z)
            assert np.array_equal(got, ref), it
        _say("12 split passes equal the unsplit ones")
    finally:
        S.autotune(1)
        h.close()


def test_split_autotuned_solve():
    """With the tuner choosing among the unsplit and split shapes (mode 1, then its choice),
    the C2-shaped sweep (Carlson S12, 32^3 cells) gives the unsplit solve's bits."""
    from radiant_hip import problems as P, solver as S
    prob = P.config_3d(N=32, width=3.2, G=2, seed=20260107)
    try:
        S.autotune(0)
        with _split(1):
            ref, _ = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
        S.autotune(1)
        for _ in range(2):
            got, _ = S.compute_flux(prob, precision="mixed-fp32", state="reduced")
            assert np.array_equal(got, ref)
        rep = S.autotune_report()
        _say(rep)
        assert "splits=2" in rep
    finally:
        S.autotune(1)
