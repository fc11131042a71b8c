"""Python host side of the MI355X evolution hot path of OllinAxis-BiB.

It mirrors what the reference's Fortran main loop does around the hot path
(src/base/evolve.f90, onestep.f90): build the parameters from a par file,
create the device-resident grid hierarchy, set initial data, run the
Berger-Oliger steps, read arrays back.  All numerical work happens in
liboaxis.so (HIP, gfx950) behind the C ABI of include/oaxis.h.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import ENUMS, Params, declared_symbols

_lib = None


def lib():
    """The loaded C ABI (raises ImportError when liboaxis.so is missing)."""
    global _lib
    if _lib is None:
        _lib = _abi.load_library()
    return _lib


class OaxisError(RuntimeError):
    pass


def _check(rc, what):
    if rc != _abi.OA_OK:
        raise OaxisError("%s failed (code %d): %s" % (what, rc, lib().oa_last_error().decode()))


def default_params():
    p = Params()
    _check(lib().oa_params_default(ctypes.byref(p)), "oa_params_default")
    return p


def params_from_text(text, base=None):
    """Parameters from par-file text (reference syntax), on top of the
    reference defaults or of ``base``."""
    p = default_params() if base is None else _copy(base)
    _check(lib().oa_params_parse_text(ctypes.byref(p), text.encode()), "par-file parse")
    return p


def params_from_file(path):
    p = default_params()
    _check(lib().oa_params_parse_file(ctypes.byref(p), path.encode()), "par-file parse")
    return p


def params_from_dict(d, base=None):
    p = default_params() if base is None else _copy(base)
    for k, v in d.items():
        _check(lib().oa_params_set(ctypes.byref(p), str(k).encode(), str(v).encode()), "set %s" % k)
    return p


def param_text(p, key):
    buf = ctypes.create_string_buffer(256)
    _check(lib().oa_params_get(ctypes.byref(p), key.encode(), buf, 256), "get %s" % key)
    return buf.value.decode()


def footprint(params):
    """Device bytes a Spacetime with these parameters allocates."""
    v = ctypes.c_double()
    _check(lib().oa_params_footprint(ctypes.byref(params), ctypes.byref(v)), "oa_params_footprint")
    return v.value


def _schedule(fn, h):
    """The coarse-step schedule in use and what the autotuner measured."""
    g, x, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ms = (ctypes.c_double * 4)()
    _check(fn(h, ctypes.byref(g), ctypes.byref(x), ctypes.byref(t), ms), "schedule")
    return dict(graphs=bool(g.value), xcd_remap=bool(x.value), tuned=bool(t.value),
                step_ms={"direct": ms[0], "graph": ms[1], "direct+xcd": ms[2], "graph+xcd": ms[3]})


def _copy(p):
    q = Params()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(Params))
    return q


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Spacetime:
    """One device-resident evolution (box 0, all refinement levels).

    Arrays are exchanged in the reference layout: shape (Nr+ghost, Nz+ghost),
    Fortran order, element [i+ghost-1, j+ghost-1] holding point (i, j).
    """

    def __init__(self, params):
        self.params = _copy(params)
        h = ctypes.c_void_p()
        _check(lib().oa_create(ctypes.byref(self.params), ctypes.byref(h)), "oa_create")
        self._h = h
        n = ctypes.c_int()
        _check(lib().oa_num_levels(self._h, ctypes.byref(n)), "oa_num_levels")
        self.nlevels = n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().oa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def level_info(self, level=0):
        Nr, Nz, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dr, dz, dt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().oa_level_info(self._h, 0, level, ctypes.byref(Nr), ctypes.byref(Nz), ctypes.byref(g),
                                   ctypes.byref(dr), ctypes.byref(dz), ctypes.byref(dt)), "oa_level_info")
        return dict(Nr=Nr.value, Nz=Nz.value, ghost=g.value, dr=dr.value, dz=dz.value, dt=dt.value,
                    eqsym=bool(self.params.eqsym))

    def shape(self, level=0):
        i = self.level_info(level)
        return (i["Nr"] + i["ghost"], i["Nz"] + i["ghost"])

    def upload(self, name, array, level=0):
        a = np.asfortranarray(np.asarray(array, dtype=np.float64))
        if a.shape != self.shape(level):
            raise ValueError("array %s has shape %s, level %d needs %s" % (name, a.shape, level, self.shape(level)))
        _check(lib().oa_upload_field(self._h, 0, level, name.encode(), _dptr(a), a.shape[0]), "upload " + name)

    def download(self, name, level=0):
        a = np.zeros(self.shape(level), dtype=np.float64, order="F")
        _check(lib().oa_download_field(self._h, 0, level, name.encode(), _dptr(a), a.shape[0]), "download " + name)
        return a

    def initial_data(self, stationary_exit=True):
        """The reference's initial data on every level.  Brill waves and the
        scalar pulse solve the Hamiltonian constraint (WaveElliptic); with
        stationary_exit=False its all-levels phase keeps the reference's
        stopping rule alone (runs to WE_maxiter when the residual stalls)."""
        _check(lib().oa_set_elliptic_exit(self._h, 1 if stationary_exit else 0), "oa_set_elliptic_exit")
        _check(lib().oa_initial_data(self._h), "oa_initial_data")

    def elliptic_stats(self):
        """Iterations of the initial data's elliptic solve (base level alone,
        then all levels) and its last residual."""
        a, b, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().oa_elliptic_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(r)),
               "oa_elliptic_stats")
        return dict(steps0=a.value, steps1=b.value, residual=r.value)

    def initial_aux(self):
        _check(lib().oa_initial_aux(self._h), "oa_initial_aux")

    def coarse_step(self):
        _check(lib().oa_coarse_step(self._h), "oa_coarse_step")

    def evolve(self, nsteps):
        _check(lib().oa_evolve(self._h, int(nsteps)), "oa_evolve")

    def sync(self):
        _check(lib().oa_sync(self._h), "oa_sync")

    def time(self, level=0):
        t, t1, t2 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        s = ctypes.c_int()
        _check(lib().oa_get_time(self._h, 0, level, ctypes.byref(t), ctypes.byref(t1), ctypes.byref(t2),
                                 ctypes.byref(s)), "oa_get_time")
        return t.value, t1.value, t2.value, s.value

    def set_time(self, level, t, t1, t2, s):
        _check(lib().oa_set_time(self._h, 0, level, t, t1, t2, s), "oa_set_time")

    # pieces of the Berger-Oliger step, in the reference's call order
    def saveold(self, level):
        _check(lib().oa_saveold(self._h, 0, level), "oa_saveold")

    def substep(self, level, it, niter, dtw, w):
        _check(lib().oa_substep(self._h, 0, level, it, niter, dtw, w), "oa_substep")

    def finebound(self, level, dtw):
        _check(lib().oa_finebound(self._h, 0, level, dtw), "oa_finebound")

    def auxiliary(self, level):
        _check(lib().oa_auxiliary(self._h, 0, level), "oa_auxiliary")

    def restrict(self, level):
        _check(lib().oa_restrict(self._h, 0, level), "oa_restrict")

    def advance_time(self, level):
        _check(lib().oa_advance_time(self._h, 0, level), "oa_advance_time")

    def constraints(self, level=0):
        """constraints.f90 on a level's current state: afterwards ham, mom_r,
        mom_z, CDelta_r, CDelta_z can be downloaded or reduced by name."""
        _check(lib().oa_constraints(self._h, 0, level), "oa_constraints")

    def checkpoint(self):
        """The whole evolution state as bytes (oa_checkpoint_save): every
        level's evolved arrays, fine-boundary time levels, clocks and
        auxiliary-pass flags.  Between coarse steps."""
        n = ctypes.c_size_t()
        _check(lib().oa_checkpoint_size(self._h, ctypes.byref(n)), "oa_checkpoint_size")
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().oa_checkpoint_save(self._h, buf, n.value), "oa_checkpoint_save")
        return buf.raw

    def restore(self, data):
        """Load a checkpoint of a spacetime with the same parameters; the
        evolution continues bit for bit as the one that saved it."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(lib().oa_checkpoint_load(self._h, buf, len(data)), "oa_checkpoint_load")

    def restart(self, directory):
        """checkpointrestart: read the state from a checkpoint directory in
        the reference's format (checkparam, checklist, the .2D dumps and the
        _bound_ strip files of every array and level) for a spacetime
        created with idata = checkpoint; then call initial_aux()."""
        _check(lib().oa_checkpoint_restart(self._h, str(directory).encode()), "oa_checkpoint_restart")

    def checkpoint_write(self, directory, parfile="checkpoint.par"):
        """checkpointsave: write the state as a checkpoint directory in the
        reference's format (read back by restart())."""
        _check(lib().oa_checkpoint_write(self._h, str(directory).encode(), str(parfile).encode()),
               "oa_checkpoint_write")

    def ahfind(self, k=0, shape=False):
        """horizon_finder for horizon k on the current state with the
        parameters' ahmethod (shoot, spectral or both): dict(found, ntheta,
        shots, errorflag, zcenter, zinit, dradlast, area, radius, mass,
        method, iterations, residual); with shape=True also theta and rad
        of the last integration (shoot) or of the fast flow's last
        iteration (spectral, both), numpy arrays."""
        res = _abi.Horizon()
        _check(lib().oa_ahfind(self._h, int(k), ctypes.byref(res)), "oa_ahfind")
        out = {name: getattr(res, name) for name, _ in _abi.Horizon._fields_}
        if shape:
            import numpy as np
            n = ctypes.c_int()
            lib().oa_horizon_shape(self._h, int(k), None, None, 0, ctypes.byref(n))
            th = np.zeros(max(n.value, 1))
            rd = np.zeros(max(n.value, 1))
            dp = ctypes.POINTER(ctypes.c_double)
            _check(lib().oa_horizon_shape(self._h, int(k), th.ctypes.data_as(dp), rd.ctypes.data_as(dp),
                                          int(th.size), ctypes.byref(n)), "oa_horizon_shape")
            out["theta"], out["rad"] = th[:n.value], rd[:n.value]
        return out

    def ahfind_begin(self, k=0):
        """Begin the search of horizon k on the current state; the returned
        search's end() gives what ahfind would have given now, after the
        caller has enqueued more work (a coarse step) to run beside its
        shots."""
        h = ctypes.c_void_p()
        _check(lib().oa_ahfind_begin(self._h, int(k), ctypes.byref(h)), "oa_ahfind_begin")
        return HorizonSearch(h, 1, single=True)

    def norms0d(self, name, level=0):
        """save0Ddata's norms of an evolved array: dict(max, min, nm1, nm2)."""
        out = (ctypes.c_double * 4)()
        _check(lib().oa_norms0d(self._h, 0, level, name.encode(), out), "oa_norms0d " + name)
        return dict(max=out[0], min=out[1], nm1=out[2], nm2=out[3])

    # timing
    def last_elapsed_ms(self):
        v = ctypes.c_double()
        _check(lib().oa_last_elapsed_ms(self._h, ctypes.byref(v)), "oa_last_elapsed_ms")
        return v.value

    def set_kernel_timing(self, on):
        _check(lib().oa_set_kernel_timing(self._h, 1 if on else 0), "oa_set_kernel_timing")

    def set_graphs(self, on):
        _check(lib().oa_set_graphs(self._h, 1 if on else 0), "oa_set_graphs")

    def schedule(self):
        return _schedule(lib().oa_schedule, self._h)

    def rhs_kernel_stats(self):
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib().oa_rhs_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(n)), "oa_rhs_kernel_stats")
        return ms.value, n.value

    def point_updates(self):
        v = ctypes.c_double()
        _check(lib().oa_point_updates(self._h, ctypes.byref(v)), "oa_point_updates")
        return v.value


class HorizonSearch:
    """A begun horizon search (oa_ahfind_begin / oa_batch_ahfind_begin):
    end() runs its shots and returns the results (a dict for a spacetime's
    search, a list for a batch's); it must be called exactly once, before
    the spacetimes are closed or leave their batch."""

    def __init__(self, h, n, single=False):
        self._h, self._n, self._single = h, n, single

    def end(self):
        if self._n == 0:
            return []
        if self._h is None:
            raise OaxisError("this horizon search has already ended")
        res = (_abi.Horizon * self._n)()
        h, self._h = self._h, None
        _check(lib().oa_ahfind_end(h, res), "oa_ahfind_end")
        out = [{name: getattr(r, name) for name, _ in _abi.Horizon._fields_} for r in res]
        return out[0] if self._single else out


class Batch:
    """Many independent spacetimes advanced together: each coarse step is one
    launch per kernel kind over the (spacetime, level) work items of every
    member.  Members keep their own state and are read with their own
    Spacetime.download; results equal those of each member evolved alone."""

    def __init__(self, members=()):
        h = ctypes.c_void_p()
        _check(lib().oa_batch_create(ctypes.byref(h)), "oa_batch_create")
        self._b = h
        self.members = []
        for st in members:
            self.add(st)

    def add(self, st):
        _check(lib().oa_batch_add(self._b, st._h), "oa_batch_add")
        self.members.append(st)

    def remove(self, st):
        _check(lib().oa_batch_remove(self._b, st._h), "oa_batch_remove")
        self.members.remove(st)

    def __len__(self):
        return len(self.members)

    def coarse_step(self):
        _check(lib().oa_batch_coarse_step(self._b), "oa_batch_coarse_step")

    def evolve(self, nsteps):
        _check(lib().oa_batch_evolve(self._b, int(nsteps)), "oa_batch_evolve")

    def sync(self):
        _check(lib().oa_batch_sync(self._b), "oa_batch_sync")

    def ahfind(self, searches):
        """horizon_finder for every (spacetime, k) in ``searches`` (members of
        this batch) in lockstep; a list of dicts as Spacetime.ahfind returns."""
        n = len(searches)
        if n == 0:
            return []
        hs = (ctypes.c_void_p * n)(*[st._h.value if isinstance(st._h, ctypes.c_void_p) else st._h
                                     for st, _ in searches])
        ks = (ctypes.c_int * n)(*[int(k) for _, k in searches])
        res = (_abi.Horizon * n)()
        _check(lib().oa_batch_ahfind(self._b, n, hs, ks, res), "oa_batch_ahfind")
        return [{name: getattr(r, name) for name, _ in _abi.Horizon._fields_} for r in res]

    def ahfind_begin(self, searches):
        """ahfind in two parts (HorizonSearch.end() returns the list)."""
        n = len(searches)
        if n == 0:
            return HorizonSearch(None, 0)
        hs = (ctypes.c_void_p * n)(*[st._h.value if isinstance(st._h, ctypes.c_void_p) else st._h
                                     for st, _ in searches])
        ks = (ctypes.c_int * n)(*[int(k) for _, k in searches])
        h = ctypes.c_void_p()
        _check(lib().oa_batch_ahfind_begin(self._b, n, hs, ks, ctypes.byref(h)), "oa_batch_ahfind_begin")
        return HorizonSearch(h, n)

    def last_elapsed_ms(self):
        v = ctypes.c_double()
        _check(lib().oa_batch_last_elapsed_ms(self._b, ctypes.byref(v)), "oa_batch_last_elapsed_ms")
        return v.value

    def point_updates(self):
        v = ctypes.c_double()
        _check(lib().oa_batch_point_updates(self._b, ctypes.byref(v)), "oa_batch_point_updates")
        return v.value

    def launch_stats(self):
        n, m = ctypes.c_long(), ctypes.c_long()
        _check(lib().oa_batch_launch_stats(self._b, ctypes.byref(n), ctypes.byref(m)), "oa_batch_launch_stats")
        return n.value, m.value

    def set_kernel_timing(self, on):
        _check(lib().oa_batch_set_kernel_timing(self._b, 1 if on else 0), "oa_batch_set_kernel_timing")

    def set_graphs(self, on):
        _check(lib().oa_batch_set_graphs(self._b, 1 if on else 0), "oa_batch_set_graphs")

    def schedule(self):
        return _schedule(lib().oa_batch_schedule, self._b)

    def rhs_kernel_stats(self):
        """(total ms, launches, grid points advanced) of the RHS launches of
        the last timed evolve."""
        ms, n, pts = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        _check(lib().oa_batch_rhs_kernel_stats(self._b, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pts)),
               "oa_batch_rhs_kernel_stats")
        return ms.value, n.value, pts.value

    def close(self):
        if getattr(self, "_b", None):
            lib().oa_batch_destroy(self._b)
            self._b = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# names of the evolved state of a configuration (reference array names)
BASE_FIELDS = ["alpha", "chi", "phi", "A", "B", "H", "C", "trK", "KTA", "KTB", "KTC", "Delta_r", "Delta_z"]
SHIFT_FIELDS = ["beta_r", "beta_z", "dtbeta_r", "dtbeta_z"]
SCALAR_FIELDS = ["scalar_phi", "scalar_xi_r", "scalar_xi_z", "scalar_pi"]


_CKPT_MAGIC = b"OAXCKPTF"


def save_checkpoint(st, path):
    """Write the spacetime's parameters and its whole state (Spacetime.
    checkpoint) to one file: the restart file of checkpointsave."""
    data = st.checkpoint()
    raw = bytes(st.params)
    with open(path, "wb") as f:
        f.write(_CKPT_MAGIC)
        f.write(len(raw).to_bytes(8, "little"))
        f.write(raw)
        f.write(data)


def load_checkpoint(path):
    """A new Spacetime from a file save_checkpoint wrote (checkpointload):
    same parameters, same state; it continues bit for bit."""
    with open(path, "rb") as f:
        if f.read(8) != _CKPT_MAGIC:
            raise OaxisError("%s is not a checkpoint of this library" % path)
        n = int.from_bytes(f.read(8), "little")
        raw = f.read(n)
        data = f.read()
    if n != ctypes.sizeof(Params):
        raise OaxisError("%s was written with another parameter layout" % path)
    p = Params.from_buffer_copy(raw)
    st = Spacetime(p)
    try:
        st.restore(data)
    except Exception:
        st.close()
        raise
    return st


def state_fields(params):
    names = list(BASE_FIELDS)
    if params.shift != ENUMS["OA_SHIFT_NONE"]:
        names += SHIFT_FIELDS
    if params.mattertype & ENUMS["OA_MATTER_SCALAR"]:
        names += SCALAR_FIELDS
    return names
