"""Output files in the reference's formats, written from the device state.

The reference writes its output from the module arrays at output steps:
  * 0D norms: ``<var>_b<box>l<level>_{max,min,nm1,nm2}.tl``, one line per
    output time, ``(2ES20.8E3)`` after a comment header
    (save0Ddata.f90:296-364);
  * 1D cuts: ``.rl`` along the equator row (j = eqz) and ``.zl`` along the
    first radial point (i = 1), a ``"Time =`` header in ``(A8,ES14.6)`` and
    rows ``(2ES20.8E3)`` over the whole index range including ghosts
    (save1Ddata.f90:236-268, :435-469);
  * 2D: ``.2D``, a ``#Time =`` header in ``(A8,ES23.15)``, rows of
    ``r, z, var + 1e-30`` over i (fastest) and j including ghosts, in
    ``(3ES23.15)`` when the file is new and ``(3ES16.8)`` when appended, a
    blank line after each j and after each time (save2Ddata.f90:220-280).
Here the norms come from the device reduction (oa_norms0d) and the 1D/2D
arrays from oa_download_field, so an output step moves only what it writes.
Single-process file names (no processor suffix).
"""
import os

import numpy as np

TINY = 1.0e-30


def fortran_es(x, w, d, e=None):
    """Fortran ``ESw.d`` (``ESw.dEe`` with e) of one value, right-justified."""
    s = "%.*E" % (d, x)
    mant, exp = s.split("E")
    ev = int(exp)
    sign = "-" if ev < 0 else "+"
    if e is not None:
        out = "%sE%s%0*d" % (mant, sign, e, abs(ev))
    elif abs(ev) <= 99:
        out = "%sE%s%02d" % (mant, sign, abs(ev))
    else:
        out = "%s%s%03d" % (mant, sign, abs(ev))       # ESw.d drops the E above 99
    if len(out) > w:
        return "*" * w
    return out.rjust(w)


def _filen(box, level):
    return "_b%dl%d" % (box, level)


def _coords(info):
    g, Nr, Nz, dr, dz = info["ghost"], info["Nr"], info["Nz"], info["dr"], info["dz"]
    i = np.arange(1 - g, Nr + 1)
    j = np.arange(1 - g, Nz + 1)
    r = (i - 0.5) * dr
    zoff = 0.5 if info["eqsym"] else 0.5 * (Nz - g + 1)
    z = (j - zoff) * dz
    return r, z


def _eqz(info):
    """j of the equator row: 1 with equatorial symmetry, else the node z = 0."""
    if info["eqsym"]:
        return 1
    g, Nz = info["ghost"], info["Nz"]
    return int(round(0.5 * (Nz - g + 1)))


def save0D(src, outdir, names, level=0, new=False, commenttype="gnuplot"):
    """Append the 0D norms of ``names`` on ``level`` (save0Ddata)."""
    comment = '"' if commenttype == "xgraph" else "#"
    t = src.time(level)[0]
    for name in names:
        n = src.norms0d(name, level)
        for key, val in (("max", n["max"]), ("min", n["min"]), ("nm1", n["nm1"]), ("nm2", n["nm2"])):
            fname = "%s%s_%s.tl" % (name, _filen(0, level), key)
            path = os.path.join(outdir, fname)
            with open(path, "w" if new else "a") as f:
                if new:
                    f.write(" %s%s\n" % (comment, fname))
                f.write(fortran_es(t, 20, 8, 3) + fortran_es(val, 20, 8, 3) + "\n")


def save1D(src, outdir, names, level=0, new=False):
    """Append the .rl and .zl cuts of ``names`` on ``level`` (save1Ddata)."""
    info = src.level_info(level)
    g = info["ghost"]
    t = src.time(level)[0]
    r, z = _coords(info)
    jeq = _eqz(info) + g - 1
    for name in names:
        a = src.download(name, level)
        for ext, xs, vals in ((".rl", r, a[:, jeq]), (".zl", z, a[g, :])):
            path = os.path.join(outdir, "%s%s%s" % (name, _filen(0, level), ext))
            with open(path, "w" if new else "a") as f:
                f.write('"Time = ' + fortran_es(t, 14, 6) + "\n")
                for x, v in zip(xs, vals):
                    f.write(fortran_es(x, 20, 8, 3) + fortran_es(v, 20, 8, 3) + "\n")
                f.write("\n")


def save2D(src, outdir, names, level=0, new=False):
    """Append the .2D data of ``names`` on ``level`` (save2Ddata)."""
    info = src.level_info(level)
    t = src.time(level)[0]
    r, z = _coords(info)
    w, d = (23, 15) if new else (16, 8)
    for name in names:
        a = src.download(name, level)
        path = os.path.join(outdir, "%s%s.2D" % (name, _filen(0, level)))
        lines = ["#Time = " + fortran_es(t, 23, 15)]
        for jj, zj in enumerate(z):
            for ii, ri in enumerate(r):
                lines.append(fortran_es(ri, w, d) + fortran_es(zj, w, d) + fortran_es(a[ii, jj] + TINY, w, d))
            lines.append("")
        lines.append("")
        with open(path, "w" if new else "a") as f:
            f.write("\n".join(lines) + "\n")
