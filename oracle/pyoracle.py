"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of liboaxis_oracle.so, the plain-C restatement of the
reference's evolution hot path (see or_internal.h).  Loaded only by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline legs (the throughput sample
and the horizon probe's search), as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboaxis_oracle.so")

_lib = None


def build():
    """Compile the oracle with the committed Makefile (gcc, IEEE, no FMA)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = ctypes.CDLL(LIBPATH)
        H = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            "or_create": (H, [ctypes.c_void_p]),
            "or_destroy": (None, [H]),
            "or_nlevels": (ctypes.c_int, [H]),
            "or_level_dims": (ctypes.c_int, [H, ctypes.c_int, ip, ip, ip, dp, dp, dp]),
            "or_field": (dp, [H, ctypes.c_int, ctypes.c_char_p]),
            "or_source": (dp, [H, ctypes.c_int, ctypes.c_char_p]),
            "or_initial_data": (ctypes.c_int, [H]),
            "or_initial_aux": (ctypes.c_int, [H]),
            "or_onestep": (ctypes.c_int, [H, ctypes.c_int]),
            "or_coarse_step": (ctypes.c_int, [H]),
            "or_evolve": (ctypes.c_int, [H, ctypes.c_int]),
            "or_get_time": (None, [H, ctypes.c_int, dp, dp, dp, ip]),
            "or_set_time": (None, [H, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
            "or_set_ext_derivs": (None, [H, ctypes.c_int, ctypes.c_int]),
            "or_apply_stencil": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp, dp]),
            "or_apply_dissipation": (ctypes.c_int, [H, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
            "or_interp_at": (ctypes.c_double, [H, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double]),
            "or_run_auxiliary": (ctypes.c_int, [H, ctypes.c_int]),
            "or_run_constraints": (ctypes.c_int, [H, ctypes.c_int]),
            "or_run_sources": (ctypes.c_int, [H, ctypes.c_int]),
            "or_run_saveold": (ctypes.c_int, [H, ctypes.c_int]),
            "or_run_finebound": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_double]),
            "or_run_restrict": (ctypes.c_int, [H, ctypes.c_int]),
            "or_run_update": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            "or_run_symmetries": (ctypes.c_int, [H, ctypes.c_int]),
            "or_ahfind_flat": (ctypes.c_int, [H, ctypes.c_int, ip, dp, ip, dp, dp, dp]),
            "or_ah_integrate": (ctypes.c_int, [H, ctypes.c_double, ctypes.c_double, dp]),
            "or_ah_set_tolerance": (None, [ctypes.c_double, ctypes.c_double]),
            "or_ell_stats": (None, [H, ip, ip, dp]),
            "or_set_ell_stationary": (None, [ctypes.c_int]),
            "or_ell_field": (dp, [H, ctypes.c_int, ctypes.c_char_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ell_stationary(on):
    """Stationary exit of the oracle's WaveElliptic solve (off by default:
    the reference's stopping rule alone; process-wide)."""
    lib().or_set_ell_stationary(1 if on else 0)


def ah_tolerance(rel=1.0e-4, abs_=1.0e-8):
    """Step-control tolerances of the oracle's horizon integration (the
    reference's by default; process-wide)."""
    lib().or_ah_set_tolerance(float(rel), float(abs_))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """The reference algorithm on the CPU for one parameter set (box 0)."""

    def __init__(self, params):
        self._params = params          # keep the ctypes struct alive
        self._h = lib().or_create(ctypes.addressof(params))
        self.nlevels = lib().or_nlevels(self._h)

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self, level=0):
        Nr, Nz, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dr, dz, dt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().or_level_dims(self._h, level, ctypes.byref(Nr), ctypes.byref(Nz), ctypes.byref(g),
                            ctypes.byref(dr), ctypes.byref(dz), ctypes.byref(dt))
        return dict(Nr=Nr.value, Nz=Nz.value, ghost=g.value, dr=dr.value, dz=dz.value, dt=dt.value)

    def shape(self, level=0):
        d = self.dims(level)
        return (d["Nr"] + d["ghost"], d["Nz"] + d["ghost"])

    def ell_stats(self):
        """Iterations of the WaveElliptic solve (base level alone, then all
        levels; 0 without refinement) and its last residual."""
        a, b, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        lib().or_ell_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(r))
        return dict(steps0=a.value, steps1=b.value, residual=r.value)

    def view(self, name, level=0, source=False):
        """Live numpy view (Fortran order) of an evolved or auxiliary array
        (or, after the initial data's elliptic solve, ell_u, ell_v, ell_S0,
        ell_S1, ell_S5)."""
        if name.startswith("ell_"):
            fn = lib().or_ell_field
        else:
            fn = lib().or_source if source else lib().or_field
        ptr = fn(self._h, level, name.encode())
        if not ptr:
            raise KeyError("oracle has no array %r on level %d" % (name, level))
        shp = self.shape(level)
        flat = np.ctypeslib.as_array(ptr, shape=(shp[0] * shp[1],))
        return flat.reshape(shp, order="F")

    def get(self, name, level=0):
        return np.array(self.view(name, level), order="F")

    def set(self, name, value, level=0):
        self.view(name, level)[...] = value

    def constraints(self, level=0):
        """constraints.f90 on a level from its last auxiliary arrays: fills
        ham, mom_r, mom_z, CDelta_r, CDelta_z (read them with get())."""
        lib().or_run_constraints(self._h, level)

    def initial_data(self):
        lib().or_initial_data(self._h)

    def initial_aux(self):
        lib().or_initial_aux(self._h)

    def coarse_step(self):
        lib().or_coarse_step(self._h)

    def evolve(self, n):
        lib().or_evolve(self._h, int(n))

    def time(self, level=0):
        t, t1, t2 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        s = ctypes.c_int()
        lib().or_get_time(self._h, level, ctypes.byref(t), ctypes.byref(t1), ctypes.byref(t2), ctypes.byref(s))
        return t.value, t1.value, t2.value, s.value

    def set_time(self, level, t, t1, t2, s):
        lib().or_set_time(self._h, level, t, t1, t2, s)

    def ext_derivs(self, level, on):
        lib().or_set_ext_derivs(self._h, level, 1 if on else 0)

    def stencil(self, kind, f, symr=1, symz=1, beta=None, level=0):
        """kind: diff1r, diff1z, diff2r, diff2z, diff2rz, diffadvr, diffadvz."""
        kinds = ["diff1r", "diff1z", "diff2r", "diff2z", "diff2rz", "diffadvr", "diffadvz"]
        f = np.asfortranarray(f, dtype=np.float64)
        out = np.zeros_like(f, order="F")
        b = np.asfortranarray(beta if beta is not None else np.zeros_like(f), dtype=np.float64)
        lib().or_apply_stencil(self._h, level, kinds.index(kind), _dptr(f), symr, symz, _dptr(b), _dptr(out))
        return out

    def dissipation(self, ev, src, symr=1, symz=1, diss=0.1, level=0):
        ev = np.asfortranarray(ev, dtype=np.float64)
        src = np.array(src, dtype=np.float64, order="F")
        lib().or_apply_dissipation(self._h, level, _dptr(ev), _dptr(src), symr, symz, diss)
        return src

    def interp(self, f, r0, z0, level=0):
        f = np.asfortranarray(f, dtype=np.float64)
        return lib().or_interp_at(self._h, level, _dptr(f), r0, z0)

    def run_auxiliary(self, level=0):
        lib().or_run_auxiliary(self._h, level)

    def run_sources(self, level=0):
        lib().or_run_sources(self._h, level)

    def run_saveold(self, level):
        lib().or_run_saveold(self._h, level)

    def run_update(self, level, it, niter, dtw, w):
        lib().or_run_update(self._h, level, it, niter, dtw, w)

    def run_symmetries(self, level):
        lib().or_run_symmetries(self._h, level)

    def run_finebound(self, level, dtw):
        lib().or_run_finebound(self._h, level, dtw)

    def run_restrict(self, level):
        lib().or_run_restrict(self._h, level)

    AH_KEYS = ("found", "ntheta", "shots", "errorflag", "zcenter", "zinit", "dradlast", "area", "radius", "mass",
               "method", "iterations", "residual")

    def ahfind(self, k=0, shape=False):
        """horizon_finder for horizon k on the current state and auxiliary
        arrays (every level).  ahfound(k), the cosine coefficients ack(k,:)
        and the spectral finder's first-call flag carry over between calls
        as in the reference.  Returns a dict; with shape=True also theta and
        rad of the last integration (shooting) or of the last fast-flow
        iteration (spectral, both)."""
        if not hasattr(self, "_ahfound"):
            self._ahfound = [ctypes.c_int(0) for _ in range(3)]
            self._ahfirst = [ctypes.c_int(1) for _ in range(3)]
            self._ahack = np.zeros((3, 51))
        res = np.zeros(13)
        th = np.zeros(50001)
        rd = np.zeros(50001)
        ack = self._ahack[k]
        rc = lib().or_ahfind_flat(self._h, int(k), ctypes.byref(self._ahfound[k]), _dptr(ack),
                                  ctypes.byref(self._ahfirst[k]), _dptr(res), _dptr(th), _dptr(rd))
        if rc:
            raise RuntimeError("oracle ahfind: error %d" % rc)
        out = dict(zip(self.AH_KEYS, res.tolist()))
        for key in ("found", "ntheta", "shots", "errorflag", "method", "iterations"):
            out[key] = int(out[key])
        if shape:
            n = out["ntheta"] + 1
            # ah_radius_SH / _SP hold h = 0 when no horizon was found (:962-966, :1783-1787)
            out["theta"] = th[:n].copy()
            out["rad"] = rd[:n].copy() if out["found"] else np.zeros(n)
        return out

    def ah_coefficients(self, k=0):
        """ack(k,:), the cosine coefficients kept for horizon k (51 values)."""
        if not hasattr(self, "_ahack"):
            return np.zeros(51)
        return self._ahack[k].copy()

    def ah_integrate(self, zinit, zcenter=0.0):
        """One integration of the horizon ODE from the axis point zinit:
        (dradlast, area, errorflag, ntheta)."""
        res = np.zeros(4)
        lib().or_ah_integrate(self._h, float(zinit), float(zcenter), _dptr(res))
        return res[0], res[1], int(res[2]), int(res[3])

    def coords(self, level=0):
        return self.get("r", level), self.get("z", level)
