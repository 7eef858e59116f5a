"""ctypes binding of libradiant_hip.so (C ABI in include/radiant_hip.h).

The library is loaded from this directory (built in-tree by radiant.jl_amd/build.py); there is
no fallback: a missing library raises immediately.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "libradiant_hip.so")

RT_OK, RT_EINVAL, RT_EDEVICE, RT_ENOMEM, RT_ECANCELLED, RT_EPENDING = 0, -1, -2, -3, -4, -5
ACCEL = {"none": 0, "livolant": 1, "anderson": 2, "gmres": 3, "bicgstab": 4}
PREC = {"fp64": 0, "mixed-fp32": 1, "mixed-bf16": 2}
STATE = {"faithful": 0, "reduced": 1}

dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_int64)
ip = C.POINTER(C.c_int32)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("ndims", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("dx", dp), ("dy", dp), ("dz", dp),
        ("mat", lp), ("bc", ip),
        ("Nmat", C.c_int32), ("G", C.c_int32),
        ("sigt", dp), ("sigs", dp), ("Ls", C.c_int32),
        ("qvol", dp),
        ("Nd", C.c_int32), ("Np", C.c_int32), ("Nps", C.c_int32),
        ("mu", dp), ("eta", dp), ("xi", dp), ("w", dp),
        ("M", dp), ("D", dp), ("pl", lp),
        ("Msurf", dp * 6), ("Dsurf", dp * 6), ("n_to_np", lp * 6),
        ("accel", C.c_int32), ("restart", C.c_int32), ("depth", C.c_int32), ("maxit", C.c_int32),
        ("tol", C.c_double),
        ("precision", C.c_int32), ("state_mode", C.c_int32), ("verbose", C.c_int32),
        ("ssrc", dp * 6), ("Np_source", C.c_int32),
        ("csd", C.c_int32), ("Sb", dp), ("dE", dp),
        ("solver_type", C.c_int32), ("T", dp), ("Mfp", dp),
        ("order_x", C.c_int32), ("omega_x", dp),
        ("order_y", C.c_int32), ("omega_y", dp), ("partial_coupling", C.c_int32),
        ("order_z", C.c_int32), ("omega_z", dp),
        ("order_e", C.c_int32), ("omega_e", dp), ("adaptive", C.c_int32), ("Mem", dp),
    ]


class GroupStats(C.Structure):
    _fields_ = [("niter", C.c_int64), ("resid", C.c_double), ("rho", C.c_double), ("converged", C.c_int32),
                ("passes", C.c_int32), ("passes_lo", C.c_int32), ("refinements", C.c_int32)]

    def as_dict(self):
        return dict(niter=self.niter, resid=self.resid, rho=self.rho, converged=bool(self.converged),
                    passes=self.passes, passes_lo=self.passes_lo, refinements=self.refinements)


class SolveInfo(C.Structure):
    _fields_ = [("ms_device", C.c_double), ("passes", C.c_int64), ("passes_lo", C.c_int64),
                ("updates_raw", C.c_double), ("updates_problem", C.c_double), ("ms_sweep", C.c_double),
                ("ms_mgemm", C.c_double), ("ms_dgemm", C.c_double), ("ms_krylov", C.c_double),
                ("graph_launches", C.c_int64), ("graph_instantiations", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int32), ("precision", C.c_int32)]


EXPORTS = ["rt_solve", "rt_problem_create", "rt_problem_solve", "rt_problem_get_flux", "rt_problem_destroy",
           "rt_problem_device_bytes", "rt_problem_pass", "rt_problem_state_length", "rt_solve_batch",
           "rt_server_create", "rt_server_submit", "rt_server_poll", "rt_server_cancel", "rt_server_destroy",
           "rt_last_error", "rt_version", "rt_test_contraction", "rt_problem_kernel_times",
           "rt_batch_create", "rt_batch_solve", "rt_batch_get_flux", "rt_batch_destroy", "rt_batch_kernel_times",
           "rt_autotune", "rt_autotune_report", "rt_solve_cutoff", "rt_problem_get_flux_cutoff",
           "rt_batch_get_flux_cutoff", "rt_server_poll_cutoff", "rt_problem_pass_device"]

_lib = None


def lib():
    """Load the library (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO_PATH):
        raise OSError(f"{SO_PATH} is missing: run radiant.jl_amd/build.py (or __graft_entry__.build())")
    L = C.CDLL(SO_PATH)
    PD = C.POINTER(ProblemDesc)
    GS = C.POINTER(GroupStats)
    H = C.c_void_p
    L.rt_last_error.restype = C.c_char_p
    L.rt_version.restype = C.c_char_p
    L.rt_solve.argtypes = [PD, dp, GS]
    L.rt_problem_create.argtypes = [PD, C.POINTER(H)]
    L.rt_problem_solve.argtypes = [H, GS, C.POINTER(SolveInfo)]
    L.rt_problem_get_flux.argtypes = [H, dp]
    L.rt_problem_destroy.argtypes = [H]
    L.rt_problem_device_bytes.argtypes = [PD, lp]
    L.rt_problem_pass.argtypes = [H, C.c_int32, C.c_int32, dp, dp, dp]
    L.rt_problem_state_length.argtypes = [H, lp]
    L.rt_problem_pass_device.argtypes = [H, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.rt_solve_batch.argtypes = [C.c_int32, C.POINTER(PD), C.POINTER(dp), C.POINTER(GS), C.POINTER(SolveInfo)]
    L.rt_problem_kernel_times.argtypes = [H, dp, lp]
    L.rt_test_contraction.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp, dp]
    L.rt_batch_create.argtypes = [C.c_int32, C.POINTER(PD), C.POINTER(H)]
    L.rt_batch_solve.argtypes = [H, C.POINTER(GS), C.POINTER(SolveInfo)]
    L.rt_batch_get_flux.argtypes = [H, C.c_int32, dp]
    L.rt_batch_destroy.argtypes = [H]
    L.rt_batch_kernel_times.argtypes = [H, dp, lp, dp, dp, dp]
    L.rt_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(H)]
    L.rt_server_submit.argtypes = [H, PD, lp]
    L.rt_server_poll.argtypes = [H, C.c_int64, dp, GS, dp]
    L.rt_server_cancel.argtypes = [H, C.c_int64]
    L.rt_server_destroy.argtypes = [H]
    L.rt_solve_cutoff.argtypes = [PD, dp, dp, GS]
    L.rt_problem_get_flux_cutoff.argtypes = [H, dp]
    L.rt_batch_get_flux_cutoff.argtypes = [H, C.c_int32, dp]
    L.rt_server_poll_cutoff.argtypes = [H, C.c_int64, dp, dp, GS, dp]
    L.rt_autotune.argtypes = [C.c_int32]
    L.rt_autotune_report.argtypes = [C.c_char_p, C.c_int64]
    L.rt_autotune_report.restype = C.c_int64
    _lib = L
    return L


class RTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"radiant_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != RT_OK:
        raise RTError(rc, lib().rt_last_error().decode(errors="replace"))
    return rc


def ptr(a, kind=dp):
    return a.ctypes.data_as(kind)
