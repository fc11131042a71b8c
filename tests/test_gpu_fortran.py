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


def test_fortran_driver_restarts_from_a_reference_checkpoint(tmp_path, oaxis_mod):
    """The driver with idata = checkpoint reads the reference's checkpoint
    directory (oa_checkpoint_restart) and continues exactly as the same
    restart driven from Python."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    p = oaxis_mod.params_from_text(PAR)
    ck = tmp_path / "checkpoint_t=0.0750"
    with oaxis_mod.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.evolve(STEPS)
        st.checkpoint_write(str(ck), "schwarz.par")
    text = PAR.replace("idata = schwarzschild", "idata = checkpoint")
    par = tmp_path / "restart.par"
    par.write_text(text)
    out = tmp_path / "out.bin"
    print("fortran driver: restart from %s, %d more steps" % (ck.name, STEPS), flush=True)
    res = subprocess.run([DRIVER, str(par), str(STEPS), str(out), str(ck)], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    with oaxis_mod.Spacetime(oaxis_mod.params_from_text(text)) as st:
        st.restart(str(ck))
        st.initial_aux()
        st.evolve(STEPS)
        shape = st.shape(0)
        want = {n: st.download(n, 0) for n in ("alpha", "chi", "trK")}
    got = np.fromfile(out, dtype=np.float64).reshape(3, shape[1], shape[0])
    for k, n in enumerate(("alpha", "chi", "trK")):
        assert np.array_equal(got[k].T, want[n]), n


ONESTEP = os.path.join(ROOT, "ollinaxis-bib_amd", "fortran", "build", "oaxis_onestep_driver")

PIECEWISE_PAR = """
idata = schwarzschild
BH1mass = 1.0
ilapse = psiminus2
slicing = 1+log
gauge_f = 2.0
shift = Gammadriver2
drivercsi = 0.75
order = four
eqsym = .false.
dr0 = 0.2
dz0 = 0.2
Nrtotal = 24
Nztotal = 48
Nl0 = 2
integrator = icn
dtfac = 0.25
geodiss = 0.1
boundtype = radiative
"""


def test_fortran_onestep_recursion_matches_coarse_steps(tmp_path, oaxis_mod, oracle_mod):
    """The reference's onestep recursion kept in Fortran, its body replaced by
    the piecewise entry points (oaxis_onestep_driver.f90): every evolved
    array of every level equals oa_coarse_step bit for bit, the level clocks
    agree, and the result is within 1e-10 of the oracle (Schwarzschild on 3
    levels, ICN, Gamma-driver)."""
    if not os.path.exists(ONESTEP):
        pytest.fail("Fortran onestep driver not built (run __graft_entry__.build())")
    p = oaxis_mod.params_from_text(PIECEWISE_PAR)
    names = list(oaxis_mod.state_fields(p))
    par = tmp_path / "schwarz_gd2.par"
    par.write_text(PIECEWISE_PAR)
    out = tmp_path / "out.bin"
    print("fortran onestep driver: Schwarzschild 24x48, 3 levels, %d steps" % STEPS, flush=True)
    res = subprocess.run([ONESTEP, str(par), str(STEPS), str(out)] + names, capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    clocks = [l.split() for l in res.stdout.splitlines() if l.startswith("level ")]
    data = np.fromfile(out, dtype=np.float64)
    orc = oracle_mod.Oracle(p)
    orc.initial_data()
    orc.initial_aux()
    for _ in range(STEPS):
        orc.coarse_step()
    with oaxis_mod.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.evolve(STEPS)
        off = 0
        for lev in range(st.nlevels):
            t, t1, t2, s = st.time(lev)
            assert int(clocks[lev][3]) == s and float(clocks[lev][5]) == t, (lev, clocks[lev], t, s)
            assert (t, t1, t2, s) == orc.time(lev)
            shape = st.shape(lev)
            n = shape[0] * shape[1]
            for name in names:
                got = data[off:off + n].reshape(shape[1], shape[0]).T
                off += n
                assert np.array_equal(got, st.download(name, lev)), (lev, name)
                b = orc.get(name, lev)
                scale = max(np.abs(b).max(), 1.0 if name == "phi" else 0.0)
                err = np.abs(got - b).max() / scale if scale > 0 else np.abs(got - b).max()
                assert err <= TOL, (lev, name, err)
        assert off == data.size
    orc.close()
