"""Python host of the GPU solve: the role Radiant's `compute_flux(::SN)` setup plays
(src/particle_transport/sn_flux.jl:27-148) in front of the C ABI.

``compute_flux(problem)`` builds the angular operators (``angular.py``), packs the arrays in
Radiant's layouts into an ``rt_problem_desc`` and calls ``rt_solve``; it returns 𝚽l as
``phi[G, Np, Nc]`` (Radiant's 𝚽l[Ng, Np, 1, Nx, Ny, Nz] with the voxels flattened) plus the
per-group statistics that sn_one_speed.jl prints.
"""
import ctypes as C
import time

import numpy as np

from . import _lib, fokker_planck, schemes
from .angular import angular_set, quadrature_dimension
from ._lib import ACCEL, PREC, STATE, check, lib, ptr


class Desc:
    """An rt_problem_desc together with the numpy arrays it points into."""

    def __init__(self, prob, precision="fp64", state="faithful", accel=None, tol=None, maxit=None, restart=None,
                 depth=None, verbose=0):
        qd = quadrature_dimension(prob.quad_type, prob.ndims, prob.qdims)
        self.ang = ang = angular_set(prob.quad_type, prob.quad_order, prob.ndims, qd, prob.legendre_order,
                                     prob.sn_type)
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))
        i64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64).reshape(-1, order="F"))
        k = self.keep = {}
        k["dx"] = f64(prob.dx)
        k["dy"] = f64(prob.dy if prob.ndims >= 2 else [1.0])
        k["dz"] = f64(prob.dz if prob.ndims >= 3 else [1.0])
        k["mat"] = i64(prob.mat)
        k["bc"] = np.ascontiguousarray(prob.bc, dtype=np.int32)
        # removal, scattering and Fokker-Planck operator of the solver type (sn_flux.jl:69-112)
        stype, sigt, sigs, T, Mfp = fokker_planck.prepare(prob, ang, qd)
        k["sigt"] = f64(sigt)
        k["sigs"] = f64(sigs)
        k["qvol"] = f64(prob.qvol)
        for nm in ("mu", "eta", "xi", "w"):
            k[nm] = f64(getattr(ang, nm))
        k["M"] = f64(ang.M)          # Nd x Np, column-major
        k["D"] = f64(ang.D)          # Np x Nd, column-major
        k["pl"] = i64(ang.pl)
        d = self.desc = _lib.ProblemDesc()
        d.ndims, d.Nx, d.Ny, d.Nz = prob.ndims, prob.Nx, prob.Ny, prob.Nz
        d.dx, d.dy, d.dz = ptr(k["dx"]), ptr(k["dy"]), ptr(k["dz"])
        d.mat = ptr(k["mat"], _lib.lp)
        d.bc = ptr(k["bc"], _lib.ip)
        d.Nmat, d.G = prob.Nmat, prob.G
        d.sigt, d.sigs = ptr(k["sigt"]), ptr(k["sigs"])
        d.Ls = prob.sigs.shape[3] - 1
        d.qvol = ptr(k["qvol"])
        d.Nd, d.Np, d.Nps = ang.Nd, ang.Np, ang.Nps
        d.mu, d.eta, d.xi, d.w = ptr(k["mu"]), ptr(k["eta"]), ptr(k["xi"]), ptr(k["w"])
        d.M, d.D, d.pl = ptr(k["M"]), ptr(k["D"]), ptr(k["pl"], _lib.lp)
        for f, (nps, Ms, Ds, n2p) in enumerate(ang.faces):
            k[f"Ms{f}"] = f64(Ms)
            k[f"Ds{f}"] = f64(Ds)
            k[f"n2p{f}"] = i64(n2p)
            d.Msurf[f] = ptr(k[f"Ms{f}"])
            d.Dsurf[f] = ptr(k[f"Ds{f}"])
            d.n_to_np[f] = ptr(k[f"n2p{f}"], _lib.lp)
        if getattr(prob, "ssrc", None):  # surface sources [G, Np_source, npos] per face, position fastest
            for f, arr in prob.ssrc.items():
                k[f"ssrc{f}"] = np.ascontiguousarray(arr, np.float64).ravel()
                d.ssrc[f] = ptr(k[f"ssrc{f}"])
            d.Np_source = int(prob.Np_source)
        if getattr(prob, "csd", False):  # continuous slowing down: Sb [G+1, Nmat], dE [G]
            k["Sb"] = f64(prob.Sb)
            k["dE"] = f64(prob.dE)
            d.csd, d.Sb, d.dE = 1, ptr(k["Sb"]), ptr(k["dE"])
        d.solver_type = stype
        K = schemes.spatial_moments(prob)
        scheme = (getattr(prob, "scheme", None) or "DD").upper()
        ox = max(1, int(getattr(prob, "scheme_order", 1) or 1))
        axes = [scheme] + [(getattr(prob, nm, None) or "DD").upper() for nm in ("scheme_y", "scheme_z")][:prob.ndims - 1]
        if getattr(prob, "csd", False):
            axes.append((getattr(prob, "scheme_e", None) or "DD").upper())
        if "AWD" in axes:  # scheme_weights.jl:32-37: one adaptive axis makes every axis adaptive
            d.adaptive = 1
            scheme = "AWD"
        if prob.ndims >= 2:
            scheme_y = "AWD" if d.adaptive else (getattr(prob, "scheme_y", None) or "DD").upper()
            oy = max(1, int(getattr(prob, "scheme_order_y", 1) or 1))
            scheme_z = ("AWD" if d.adaptive else (getattr(prob, "scheme_z", None) or "DD").upper()) if prob.ndims == 3 else "DD"
            oz = max(1, int(getattr(prob, "scheme_order_z", 1) or 1)) if prob.ndims == 3 else 1
            if ox > 1 or oy > 1 or oz > 1 or scheme != "DD" or scheme_y != "DD" or scheme_z != "DD":
                for sc in (scheme, scheme_y, scheme_z):
                    if sc not in ("DD", "DG", "AWD"):
                        raise ValueError("Closure relation not implemented yet.")
                k["omega_x"] = np.ascontiguousarray(schemes.weights_1d(scheme, ox))
                k["omega_y"] = np.ascontiguousarray(schemes.weights_1d(scheme_y, oy))
                d.order_x, d.omega_x = ox, ptr(k["omega_x"])
                d.order_y, d.omega_y = oy, ptr(k["omega_y"])
                if prob.ndims == 3:
                    k["omega_z"] = np.ascontiguousarray(schemes.weights_1d(scheme_z, oz))
                    d.order_z, d.omega_z = oz, ptr(k["omega_z"])
                d.partial_coupling = 0 if getattr(prob, "full_coupling", True) else 1
        elif ox > 1 or scheme != "DD":  # closure weights of the x axis (scheme_weights.jl)
            k["omega_x"] = np.ascontiguousarray(schemes.weights_1d(scheme, ox))
            d.order_x, d.omega_x = ox, ptr(k["omega_x"])
        scheme_e = (getattr(prob, "scheme_e", None) or "DD").upper()
        oe = max(1, int(getattr(prob, "scheme_order_e", 1) or 1))
        if oe > 1 or scheme_e != "DD":  # the energy closure of the slowing-down types
            if not getattr(prob, "csd", False):
                raise ValueError("an energy closure needs the slowing-down term (csd)")
            if scheme_e not in ("DD", "DG", "AWD"):
                raise ValueError("Closure relation not implemented yet.")
            k["omega_e"] = np.ascontiguousarray(schemes.weights_1d(scheme_e, oe))
            d.order_e, d.omega_e = oe, ptr(k["omega_e"])
        if getattr(prob, "csd", False):
            d.partial_coupling = 0 if getattr(prob, "full_coupling", True) else 1
        if getattr(prob, "em_B", None) is not None:  # magnetic field: ℳ_EM [G, Np, Np] column-major
            from . import electromagnetic
            Mem, _ = electromagnetic.moment_operator(ang, prob.em_B, prob.em_charge, prob.Eb, qd)
            k["Mem"] = np.ascontiguousarray(Mem.reshape(-1, order="F"))
            d.Mem = ptr(k["Mem"])
        if Mfp is not None:  # T [G, Nmat], ℳ [Np, Np], both column-major
            k["T"] = f64(T)
            k["Mfp"] = f64(Mfp)
            d.T, d.Mfp = ptr(k["T"]), ptr(k["Mfp"])
        d.accel = ACCEL[accel or prob.accel]
        d.restart = restart if restart is not None else prob.restart
        d.depth = depth if depth is not None else prob.depth
        d.maxit = maxit if maxit is not None else prob.maxit
        d.tol = tol if tol is not None else prob.tol
        d.precision = PREC[precision]
        d.state_mode = STATE[state]
        d.verbose = verbose
        # moment columns: spatial moment + K * cell
        self.G, self.Np, self.Nc, self.Nd = prob.G, ang.Np, prob.Nc * K, ang.Nd

    def byref(self):
        return C.byref(self.desc)

    def device_bytes(self):
        b = C.c_int64()
        check(lib().rt_problem_device_bytes(self.byref(), C.byref(b)))
        return b.value


def _phi(desc, flat):
    return flat.reshape((desc.G, desc.Np, desc.Nc), order="F")


def compute_flux(prob, precision="fp64", state="faithful", **kw):
    """One-shot solve through rt_solve; returns (phi[G, Np, Nc], stats list)."""
    d = Desc(prob, precision, state, **kw)
    phi = np.zeros(d.G * d.Np * d.Nc)
    st = (_lib.GroupStats * d.G)()
    check(lib().rt_solve(d.byref(), ptr(phi), st))
    return _phi(d, phi), [s.as_dict() for s in st]


def compute_flux_cutoff(prob, precision="fp64", state="faithful", **kw):
    """One-shot solve through rt_solve_cutoff; returns (phi[G, Np, Nc], cutoff[Np, Nc], stats):
    𝚽cutoff of the slowing-down solver types (sn_flux.jl:213-218), zeros without slowing down."""
    d = Desc(prob, precision, state, **kw)
    phi = np.zeros(d.G * d.Np * d.Nc)
    cut = np.zeros(d.Np * d.Nc)
    st = (_lib.GroupStats * d.G)()
    check(lib().rt_solve_cutoff(d.byref(), ptr(phi), ptr(cut), st))
    return _phi(d, phi), cut.reshape((d.Np, d.Nc), order="F"), [s.as_dict() for s in st]


class ResidentProblem:
    """A problem uploaded once (rt_problem_create); solve() keeps everything in HBM."""

    def __init__(self, prob, precision="fp64", state="faithful", **kw):
        self.d = Desc(prob, precision, state, **kw)
        self.h = C.c_void_p()
        check(lib().rt_problem_create(self.d.byref(), C.byref(self.h)))

    def solve(self):
        st = (_lib.GroupStats * self.d.G)()
        info = _lib.SolveInfo()
        check(lib().rt_problem_solve(self.h, st, C.byref(info)))
        return [s.as_dict() for s in st], info

    def flux(self):
        phi = np.zeros(self.d.G * self.d.Np * self.d.Nc)
        check(lib().rt_problem_get_flux(self.h, ptr(phi)))
        return _phi(self.d, phi)

    def flux_cutoff(self):
        """𝚽cutoff[Np, Nc] of the last solve (rt_problem_get_flux_cutoff)."""
        cut = np.zeros(self.d.Np * self.d.Nc)
        check(lib().rt_problem_get_flux_cutoff(self.h, ptr(cut)))
        return cut.reshape((self.d.Np, self.d.Nc), order="F")

    def kernel_times(self):
        """Per-kernel (ms, launches) of the last solve, [precision][kernel] (RADIANT_HIP_PROFILE=1)."""
        ms = np.zeros(12)
        n = np.zeros(12, np.int64)
        check(lib().rt_problem_kernel_times(self.h, ptr(ms), ptr(n, _lib.lp)))
        return ms.reshape(3, 4), n.reshape(3, 4)

    def state_length(self):
        s = C.c_int64()
        check(lib().rt_problem_state_length(self.h, C.byref(s)))
        return s.value

    def transport_pass(self, g, mode, qlout, z):
        """rt_problem_pass: mode 0 T(z), 1 A z, 2 z - A z on the faithful state layout."""
        zin = np.ascontiguousarray(z, np.float64)
        zout = np.zeros_like(zin)
        q = np.ascontiguousarray(qlout, np.float64)
        check(lib().rt_problem_pass(self.h, g, mode, ptr(q), ptr(zin), ptr(zout)))
        return zout

    def close(self):
        if self.h:
            lib().rt_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(probs, precision="mixed-fp32", **kw):
    """rt_solve_batch over a list of problems; returns ([phi], [stats], info)."""
    descs = [Desc(p, precision, "reduced" if precision != "fp64" else "faithful", **kw) for p in probs]
    n = len(descs)
    PD = C.POINTER(_lib.ProblemDesc)
    parr = (PD * n)(*[C.pointer(d.desc) for d in descs])
    phis = [np.zeros(d.G * d.Np * d.Nc) for d in descs]
    oarr = (_lib.dp * n)(*[ptr(p) for p in phis])
    stats = [(_lib.GroupStats * d.G)() for d in descs]
    sarr = (C.POINTER(_lib.GroupStats) * n)(*[C.cast(s, C.POINTER(_lib.GroupStats)) for s in stats])
    info = _lib.SolveInfo()
    check(lib().rt_solve_batch(n, parr, oarr, sarr, C.byref(info)))
    return ([_phi(d, p) for d, p in zip(descs, phis)], [[s.as_dict() for s in st] for st in stats], info)


class ResidentBatch:
    """rt_batch_*: problems uploaded once, solved together (lockstep in the mixed modes)."""

    def __init__(self, probs, precision="mixed-fp32", **kw):
        self.descs = [Desc(p, precision, "reduced" if precision != "fp64" else "faithful", **kw) for p in probs]
        n = len(self.descs)
        PD = C.POINTER(_lib.ProblemDesc)
        parr = (PD * n)(*[C.pointer(d.desc) for d in self.descs])
        self.h = C.c_void_p()
        check(lib().rt_batch_create(n, parr, C.byref(self.h)))

    def solve(self):
        n = len(self.descs)
        stats = [(_lib.GroupStats * d.G)() for d in self.descs]
        sarr = (C.POINTER(_lib.GroupStats) * n)(*[C.cast(s, C.POINTER(_lib.GroupStats)) for s in stats])
        info = _lib.SolveInfo()
        check(lib().rt_batch_solve(self.h, sarr, C.byref(info)))
        return [[s.as_dict() for s in st] for st in stats], info

    def kernel_times(self):
        """Per-kernel (ms, launches, updates, flops, algorithmic bytes) of the last solve, each
        [precision][kernel] (RADIANT_HIP_PROFILE=1; lockstep batches)."""
        ms, upd, fl, by = np.zeros(12), np.zeros(12), np.zeros(12), np.zeros(12)
        n = np.zeros(12, np.int64)
        check(lib().rt_batch_kernel_times(self.h, ptr(ms), ptr(n, _lib.lp), ptr(upd), ptr(fl), ptr(by)))
        return ms.reshape(3, 4), n.reshape(3, 4), upd.reshape(3, 4), fl.reshape(3, 4), by.reshape(3, 4)

    def flux(self, i):
        d = self.descs[i]
        phi = np.zeros(d.G * d.Np * d.Nc)
        check(lib().rt_batch_get_flux(self.h, i, ptr(phi)))
        return _phi(d, phi)

    def flux_cutoff(self, i):
        """𝚽cutoff[Np, Nc] of problem i after the last solve (rt_batch_get_flux_cutoff)."""
        d = self.descs[i]
        cut = np.zeros(d.Np * d.Nc)
        check(lib().rt_batch_get_flux_cutoff(self.h, i, ptr(cut)))
        return cut.reshape((d.Np, d.Nc), order="F")

    def close(self):
        if self.h:
            lib().rt_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """rt_server_*: continuous batching of fixed-source solves under an HBM budget."""

    def __init__(self, hbm_budget_bytes, max_active=8, precision="mixed-fp32"):
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(max_active), PREC[precision])
        self.precision = precision
        self.h = C.c_void_p()
        check(lib().rt_server_create(C.byref(cfg), C.byref(self.h)))
        self.descs = {}
        self.last_collect_ms = 0.0

    def describe(self, prob, **kw):
        """The problem description a request of this server carries (built on the host)."""
        return Desc(prob, self.precision, "reduced" if self.precision != "fp64" else "faithful", **kw)

    def submit(self, prob, **kw):
        """Submits a problem, or a description built with describe(); returns the request id."""
        d = prob if isinstance(prob, Desc) else self.describe(prob, **kw)
        rid = C.c_int64()
        check(lib().rt_server_submit(self.h, d.byref(), C.byref(rid)))
        self.descs[rid.value] = (d.G, d.Np, d.Nc)
        return rid.value

    def poll(self, rid, cutoff=False):
        """Returns None while pending, else (phi, stats, latency_ms), or with cutoff=True
        (phi, stats, latency_ms, cutoff[Np, Nc]); raises on failure/cancel.

        latency_ms is submit -> flux final in device memory; the device -> host copy of the
        collection that follows takes ``self.last_collect_ms``."""
        G, Np, Nc = self.descs[rid]
        st = (_lib.GroupStats * G)()
        lat = C.c_double()
        rc = lib().rt_server_poll(self.h, rid, None, st, C.byref(lat))  # status query
        if rc == _lib.RT_EPENDING:
            return None
        if rc != _lib.RT_OK:
            del self.descs[rid]
            check(rc)
        t0 = time.perf_counter()
        phi = np.empty(G * Np * Nc)
        cut = np.empty(Np * Nc) if cutoff else None
        # collects and releases
        rc = lib().rt_server_poll_cutoff(self.h, rid, ptr(phi), ptr(cut) if cutoff else None, st, C.byref(lat))
        self.last_collect_ms = (time.perf_counter() - t0) * 1e3
        del self.descs[rid]
        check(rc)
        out = (phi.reshape((G, Np, Nc), order="F"), [s.as_dict() for s in st], lat.value)
        return out + (cut.reshape((Np, Nc), order="F"),) if cutoff else out

    def finished(self, rid):
        """Status query (rt_server_poll with no flux buffer): True once the request's flux is final
        in device memory; the result stays there for a later poll.  Raises on failure/cancel."""
        G = self.descs[rid][0]
        st = (_lib.GroupStats * G)()
        lat = C.c_double()
        rc = lib().rt_server_poll(self.h, rid, None, st, C.byref(lat))
        if rc == _lib.RT_EPENDING:
            return False
        if rc != _lib.RT_OK:
            del self.descs[rid]
            check(rc)
        return True

    def discard(self, rid):
        """Drops a finished request's result without copying it out (rt_server_cancel releases its
        device memory), then releases the id; returns its latency in ms (submit -> flux final)."""
        G = self.descs[rid][0]
        st = (_lib.GroupStats * G)()
        lat = C.c_double()
        rc = lib().rt_server_poll(self.h, rid, None, st, C.byref(lat))
        if rc != _lib.RT_OK:
            del self.descs[rid]
            check(rc)
        check(lib().rt_server_cancel(self.h, rid))
        rc = lib().rt_server_poll(self.h, rid, None, st, None)
        del self.descs[rid]
        if rc != _lib.RT_ECANCELLED:
            check(rc)
        return lat.value, [s.as_dict() for s in st]

    def cancel(self, rid):
        check(lib().rt_server_cancel(self.h, rid))

    def close(self):
        if self.h:
            lib().rt_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_alone(prob, precision="mixed-fp32", timeout_s=120.0):
    """The flux a server computes for this request when it is the only one (the reference for
    "the same alone and in the mix")."""
    srv = Server(hbm_budget_bytes=max(1, Desc(prob, precision, "reduced" if precision != "fp64" else "faithful")
                                      .device_bytes()), max_active=1, precision=precision)
    try:
        rid = srv.submit(prob)
        t0 = time.perf_counter()
        while True:
            r = srv.poll(rid)
            if r is not None:
                return r[0]
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("request not served in time")
            time.sleep(0.002)
    finally:
        srv.close()


def autotune(mode):
    """rt_autotune: 0 default launch shapes, 1 online tuning (default), 2 a new variant every launch."""
    check(lib().rt_autotune(int(mode)))


def autotune_report():
    """rt_autotune_report: one line per tuned sweep launch shape."""
    n = lib().rt_autotune_report(None, 0)
    buf = C.create_string_buffer(int(n) + 1)
    lib().rt_autotune_report(buf, n + 1)
    return buf.value.decode()


def version():
    return lib().rt_version().decode()
