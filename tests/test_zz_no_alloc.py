"""rt_problem_pass_device, the allocation-free variant of rt_problem_pass: on device buffers the
caller allocated (torch tensors here), its result equals rt_problem_pass bit for bit, on two calls
in a row on one GPU (the second on a smaller problem, so the shared scratch array holds the first
call's larger contents), in every pass mode; a problem whose pass needs more scratch than the fixed
array is refused with a clear error."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _qlout(rp, lxs, seed):
    """A random out-of-group source on the moments that carry a scattering source (Legendre order
    <= the cross-section order), zero on the others, as the group loop produces it."""
    d = rp.d
    rng = np.random.default_rng(seed)
    q = np.zeros((d.Np, d.Nc))
    act = np.nonzero(d.ang.pl <= lxs)[0]
    q[act] = rng.uniform(0.0, 1.0, (len(act), d.Nc))
    return q.ravel(order="F")


class _Dev:
    """Device buffers of the caller, allocated through the HIP runtime the library itself uses."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host -> device
        return p

    def get(self, p, n):
        out = np.empty(n, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self, *ps):
        for p in ps:
            assert self.hip.hipFree(p) == 0


def _device_pass(dev, rp, g, mode, q, z):
    from radiant_hip import _lib
    dq, dz = dev.put(q), dev.put(z)
    do = dev.alloc(z.size * 8)
    try:
        _lib.check(_lib.lib().rt_problem_pass_device(rp.h, g, mode, dq, dz, do))
        return dev.get(do, z.size)
    finally:
        dev.free(dq, dz, do)


def test_device_pass_equals_host_pass():
    from radiant_hip import problems, solver
    cases = [problems.small_problem(2, N=20, G=3, L=2, quad_order=6, seed=11),  # larger first
             problems.small_problem(1, N=12, G=2, L=3, seed=12)]               # then smaller
    dev = _Dev()
    for k, p in enumerate(cases):
        rp = solver.ResidentProblem(p, precision="fp64", state="faithful")
        S = rp.state_length()
        rng = np.random.default_rng(100 + k)
        for g in range(p.G):
            q = _qlout(rp, p.Lxs, 10 * k + g)
            z = rng.uniform(-1.0, 1.0, S)
            for mode in (0, 1, 2):
                _say(f"case {k} ({p.ndims}D, S={S}) group {g} mode {mode}")
                want = rp.transport_pass(g, mode, q, z)
                got = _device_pass(dev, rp, g, mode, q, z)
                assert np.array_equal(got, want), (k, g, mode, float(np.abs(got - want).max()))
        rp.close()


def test_device_pass_refuses_oversized_problem():
    from radiant_hip import problems, solver, _lib
    p = problems.config_3d(N=40, width=4.0, G=1, L=1, quad_order=12, Nmat=2, seed=5)
    rp = solver.ResidentProblem(p, precision="fp64", state="faithful")
    try:
        S = rp.state_length()
        assert 2 * S + rp.d.Np * rp.d.Nc > (8 << 20)
        dev = _Dev()
        buf, q = dev.alloc(S * 8), dev.alloc(rp.d.Np * rp.d.Nc * 8)
        rc = _lib.lib().rt_problem_pass_device(rp.h, 0, 0, q, buf, buf)
        dev.free(buf, q)
        assert rc == _lib.RT_EINVAL
        msg = _lib.lib().rt_last_error().decode()
        _say(msg)
        assert "scratch" in msg and "maximum" in msg
    finally:
        rp.close()
