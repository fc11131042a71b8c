"""The output writers (oaxis/output.py) against the reference's Fortran
edit descriptors and file layouts (save0Ddata, save1Ddata, save2Ddata),
on a stand-in state (no device needed: the writers only call download,
norms0d, level_info and time)."""
import os

import numpy as np
import pytest

from oaxis import output


@pytest.mark.parametrize("x,w,d,e,want", [
    (1.0, 20, 8, 3, "     1.00000000E+000"),
    (-2.5, 23, 15, None, " -2.500000000000000E+00"),
    (0.0, 14, 6, None, "  0.000000E+00"),
    (123456.0, 20, 8, 3, "     1.23456000E+005"),
    (-1.0e-5, 16, 8, None, " -1.00000000E-05"),
    (1.5e-120, 16, 8, None, "  1.50000000-120"),       # ESw.d drops the E above exponent 99
    (9.9999999999, 14, 6, None, "  1.000000E+01"),      # rounding carries into the exponent
    (3.0e-200, 20, 8, 3, "     3.00000000E-200"),
])
def test_es_edit_descriptor(x, w, d, e, want):
    assert output.fortran_es(x, w, d, e) == want


class Fake:
    """A level 0 of 5 x 6 (equatorial symmetry) or 5 x 8 points, ghost 3, with a known array."""

    def __init__(self, eqsym=False):
        self.g, self.Nr, self.Nz = 3, 5, 6 if eqsym else 8     # Nz + ghost odd: z = 0 is a node
        self.eqsym = eqsym
        rng = np.random.default_rng(4)
        self.a = rng.normal(size=(self.Nr + self.g, self.Nz + self.g))

    def level_info(self, level=0):
        return dict(Nr=self.Nr, Nz=self.Nz, ghost=self.g, dr=0.1, dz=0.1, dt=0.025, eqsym=self.eqsym)

    def time(self, level=0):
        return (0.125, 0.1, 0.075, 5)

    def download(self, name, level=0):
        return self.a

    def norms0d(self, name, level=0):
        return dict(max=2.0, min=-1.0, nm1=2.0, nm2=0.5)


def _rows(path):
    return open(path).read().split("\n")


def test_2d_file_layout(tmp_path):
    src = Fake()
    output.save2D(src, str(tmp_path), ["alpha"], new=True)
    output.save2D(src, str(tmp_path), ["alpha"], new=False)
    rows = _rows(os.path.join(str(tmp_path), "alpha_b0l0.2D"))
    ni, nj = src.Nr + src.g, src.Nz + src.g
    block = 1 + nj * (ni + 1) + 1
    assert rows[0] == "#Time =   1.250000000000000E-01"
    assert rows[block] == rows[0]
    # first frame: full precision, r fastest, z of j = 1-ghost first, var + 1e-30
    first = rows[1:1 + ni]
    vals = np.array([[float(v.replace("D", "E")) for v in (l[:23], l[23:46], l[46:])] for l in first])
    r = (np.arange(1 - src.g, src.Nr + 1) - 0.5) * 0.1
    z0 = (1 - src.g - 0.5 * (src.Nz - src.g + 1)) * 0.1
    assert np.allclose(vals[:, 0], r, rtol=0, atol=1e-15) and np.allclose(vals[:, 1], z0, atol=1e-15)
    assert np.allclose(vals[:, 2], src.a[:, 0] + 1e-30, rtol=1e-15, atol=0)
    assert rows[1 + ni] == ""
    # appended frame: (3ES16.8)
    assert len(rows[block + 1]) == 48
    assert abs(float(rows[block + 1][32:]) - src.a[0, 0]) <= 1e-8 * abs(src.a[0, 0])


@pytest.mark.parametrize("eqsym", [False, True])
def test_1d_cuts(tmp_path, eqsym):
    src = Fake(eqsym)
    output.save1D(src, str(tmp_path), ["trK"], new=True)
    g = src.g
    rl = _rows(os.path.join(str(tmp_path), "trK_b0l0.rl"))
    zl = _rows(os.path.join(str(tmp_path), "trK_b0l0.zl"))
    assert rl[0] == '"Time =   1.250000E-01'
    jeq = 1 if eqsym else (src.Nz - g + 1) // 2     # the row where z = 0 (or the first row with eqsym)
    rv = np.array([float(l[20:]) for l in rl[1:1 + src.Nr + g]])
    zv = np.array([float(l[20:]) for l in zl[1:1 + src.Nz + g]])
    assert np.allclose(rv, src.a[:, jeq + g - 1], rtol=1e-8)
    assert np.allclose(zv, src.a[g, :], rtol=1e-8)
    zcol = np.array([float(l[:20]) for l in zl[1:1 + src.Nz + g]])
    if not eqsym:
        assert abs(zcol[jeq + g - 1]) < 1e-12            # the equator row is z = 0
    assert rl[1 + src.Nr + g] == ""


def test_0d_norm_files(tmp_path):
    src = Fake()
    output.save0D(src, str(tmp_path), ["alpha"], new=True)
    output.save0D(src, str(tmp_path), ["alpha"], new=False)
    rows = _rows(os.path.join(str(tmp_path), "alpha_b0l0_nm2.tl"))
    assert rows[0] == " #alpha_b0l0_nm2.tl"
    assert rows[1] == "     1.25000000E-001     5.00000000E-001" and rows[2] == rows[1]
