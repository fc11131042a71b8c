"""The Fortran side of the boundary: a Fortran main loop (the reference's
host language) drives liboaxis.so through ISO_C_BINDING
(ollinaxis-bib_amd/fortran/oaxis_mod.f90).  Its result must be bit-identical
to the same evolution driven from Python through ctypes, and agree with the
oracle to the parity tolerance."""
import os
import subprocess

import numpy as np
import pytest

from parity_util import TOL

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ollinaxis-bib_amd", "fortran", "build", "oaxis_driver")

PAR = """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
order = four
eqsym = .false.
dr0 = 0.1
dz0 = 0.1
Nrtotal = 32
Nztotal = 64
Nl0 = 1
integrator = rk4
dtfac = 0.25
slicing = 1+log
gauge_f = 2.0
geodiss = 0.1
boundtype = radiative
"""
STEPS = 3


def test_fortran_driver_matches_python_and_oracle(tmp_path, oaxis_mod, oracle_mod):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    par = tmp_path / "schwarz.par"
    par.write_text(PAR)
    out = tmp_path / "out.bin"
    print("fortran driver: Schwarzschild 32x64, 2 levels, %d steps" % STEPS, flush=True)
    res = subprocess.run([DRIVER, str(par), str(STEPS), str(out)], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    p = oaxis_mod.params_from_text(PAR)
    with oaxis_mod.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.evolve(STEPS)
        shape = st.shape(0)
        want = {n: st.download(n, 0) for n in ("alpha", "chi", "trK")}
        ah = st.ahfind(0)
    # the horizon line: the driver's oa_ahfind on the same state
    line = [l for l in res.stdout.splitlines() if l.startswith("horizon found")][0].split()
    assert int(line[2]) == ah["found"] == 1 and int(line[4]) == ah["shots"]
    for key, pos in (("zinit", 6), ("area", 8), ("mass", 10)):
        assert float(line[pos]) == ah[key], (key, line[pos], ah[key])
    got = np.fromfile(out, dtype=np.float64).reshape(3, shape[1], shape[0])
    for k, n in enumerate(("alpha", "chi", "trK")):
        assert np.array_equal(got[k].T, want[n]), n
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    orc.initial_aux()
    for _ in range(STEPS):
        orc.coarse_step()
    for k, n in enumerate(("alpha", "chi", "trK")):
        b = orc.get(n, 0)
        err = np.abs(got[k].T - b).max() / np.abs(b).max()
        assert err <= TOL, (n, err)
    orc.close()
