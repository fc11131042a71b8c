"""Shared driver for GPU-vs-oracle parity runs.

Both sides start from the same initial arrays: the oracle's closed-form
initial data is uploaded into the device library before the initial
auxiliary pass, then both advance the same number of coarse steps and every
evolved array of every level is compared in the max norm, relative to the
oracle array's max norm.
"""
import numpy as np

TOL = 1.0e-10      # relative max-norm agreement required for FP64 parity


def rel_err(a, b):
    if not np.all(np.isfinite(a)):
        return float("inf")
    scale = np.abs(b).max()
    diff = np.abs(a - b).max()
    if scale == 0.0:
        return diff
    return diff / scale


def make_params(oaxis, text):
    return oaxis.params_from_text(text)


# lambda = (A-H)/r^2 and Alambda = (KTA-KTH)/r^2 are not evolved under
# nolambda: they are recomputed from A, H, KTA, KT_H, whose last-bit
# differences the 1/r^2 amplifies near the axis.  They are compared through
# r^2*lambda, r^2*Alambda, i.e. against the scale of A-H and KTA-KT_H.
R2_SCALED = {"lambda": ("A", "H"), "Alambda": ("KTA", "KTH")}
# phi = -log(chi)/2 is a logarithm: its absolute error is the relative error
# of chi, so it is compared against max(1, max|phi|) rather than against
# max|phi| alone, which is ~1e-6 for weak scalar pulses.
LOG_SCALED = ("phi",)


def compare_state(oaxis, st, orc, params, extra=("KTH", "lambda", "Alambda"), get=None):
    """Relative max-norm differences of every evolved array of every level,
    st (a Spacetime, or any object read through get(name, level)) against the
    oracle orc."""
    if get is None:
        get = st.download
    out = {}
    for lev in range(orc.nlevels):
        r = orc.get("r", lev)
        for name in list(oaxis.state_fields(params)) + list(extra):
            a = get(name, lev)
            b = orc.get(name, lev)
            if not np.all(np.isfinite(a)):
                out[(lev, name)] = float("inf")
            elif name in LOG_SCALED:
                out[(lev, name)] = np.abs(a - b).max() / max(1.0, np.abs(b).max())
            elif name in R2_SCALED:
                u, v = R2_SCALED[name]
                scale = max(np.abs(orc.get(u, lev)).max(), np.abs(orc.get(v, lev)).max())
                out[(lev, name)] = np.abs((a - b) * r * r).max() / scale if scale > 0 else 0.0
            else:
                out[(lev, name)] = rel_err(a, b)
    return out


def run_pair(oaxis, pyoracle, text, nsteps, check_initial=True):
    p = oaxis.params_from_text(text)
    orc = pyoracle.Oracle(p)
    orc.initial_data()
    st = oaxis.Spacetime(p)
    for lev in range(st.nlevels):
        for name in oaxis.state_fields(p):
            st.upload(name, orc.get(name, lev), lev)
    orc.initial_aux()
    st.initial_aux()
    errs0 = compare_state(oaxis, st, orc, p) if check_initial else {}
    for _ in range(nsteps):
        orc.coarse_step()
        st.coarse_step()
    errs = compare_state(oaxis, st, orc, p)
    times = [(st.time(l), orc.time(l)) for l in range(st.nlevels)]
    st.close()
    orc.close()
    return errs0, errs, times


def worst(errs):
    if not errs:
        return 0.0, None
    k = max(errs, key=lambda q: errs[q])
    return errs[k], k


def run_pair_with_spread(oaxis, pyoracle, text, nsteps, seed=1):
    """run_pair, plus a third run: the oracle again with every evolved array
    perturbed by at most one unit in the last place after each coarse step
    (as two implementations that round differently would).  Returns the
    GPU-vs-oracle errors, the spread between the two oracle runs (both
    relative to the oracle array's max norm, per level and array), and the
    level clocks."""
    p = oaxis.params_from_text(text)
    orc = pyoracle.Oracle(p)
    orc.initial_data()
    pert = pyoracle.Oracle(p)
    st = oaxis.Spacetime(p)
    names = list(oaxis.state_fields(p))
    for lev in range(st.nlevels):
        for name in names:
            a = orc.get(name, lev)
            st.upload(name, a, lev)
            pert.set(name, a, lev)
    rng = np.random.default_rng(seed)
    orc.initial_aux()
    pert.initial_aux()
    st.initial_aux()
    for _ in range(nsteps):
        orc.coarse_step()
        pert.coarse_step()
        st.coarse_step()
        for lev in range(pert.nlevels):
            for name in names:
                v = pert.get(name, lev)
                pert.set(name, v * (1.0 + rng.integers(-1, 2, size=v.shape) * 2.0 ** -52), lev)
    errs = compare_state(oaxis, st, orc, p)
    spread = compare_state(oaxis, None, orc, p, get=pert.get)
    times = [(st.time(l), orc.time(l)) for l in range(st.nlevels)]
    st.close()
    orc.close()
    pert.close()
    return errs, spread, times
