"""GPU: checkpoints in the reference's format (checkpointsave /
checkpointrestart, src/base/checkpoint.f90:2-690).

A spacetime writes its state as the reference's checkpoint directory
(oa_checkpoint_write) and a handle created with idata = checkpoint reads it
back (oa_checkpoint_restart):
  * the files have the reference's layout: checkparam and checklist, one
    .2D dump per array and level (a "#Time =" header in (A8,ES23.15), then
    (Nr+g)(Nz+g) rows of (3ES16.8), a blank line after each j and at the
    end), and four strip files per evolved array (the clocks in (3ES23.15),
    a blank line, then one (3ES23.15) row of the three time levels per strip
    point);
  * the restarted state is the values the files hold (parsed again here
    with numpy), bit for bit, and the clocks are the strip headers';
  * after oa_initial_aux (what evolve.f90:128-195 runs after initial),
    evolution continues as the uninterrupted run to the 9 significant
    digits the dumps keep.  Two levels: the fine level's boundary time
    levels come from the strip files, so a misplaced strip value would show
    at the fine boundary at O(1) after the first step;
  * with the Gamma-driver the restart sets dtbeta = drivercsi Delta, as the
    reference's evolve.f90:174-186 does after any initial data;
  * oa_initial_data refuses idata = checkpoint; a checkpoint of another grid
    spacing is refused with the reference's message, and one whose checklist
    lacks an evolved array is refused.
"""
import os

import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def _restarted(oaxis, text, directory):
    for key in ("idata = schwarzschild", "idata = scalarpulse"):
        text = text.replace(key, "idata = checkpoint")
    b = oaxis.Spacetime(oaxis.params_from_text(text))
    with pytest.raises(oaxis.OaxisError):
        b.initial_data()                       # initial.f90:157-161: the state comes from the directory
    b.restart(directory)
    return b


def _dump(directory, name, lev, nrt, nzt):
    v = np.loadtxt(os.path.join(directory, "%s_b0l%d.2D" % (name, lev)))   # '#' header skipped
    assert v.shape == (nrt * nzt, 3)
    return v[:, 2].reshape(nzt, nrt).T


def _check_layout(directory, names, info, lev):
    g, nrt, nzt = info["ghost"], info["Nr"] + info["ghost"], info["Nz"] + info["ghost"]
    for name in names:
        lines = open(os.path.join(directory, "%s_b0l%d.2D" % (name, lev))).read().split("\n")
        assert lines[0].startswith("#Time = ") and len(lines[0]) == 8 + 23
        rows = [x for x in lines[1:] if x]
        assert len(rows) == nrt * nzt and all(len(x) == 48 for x in rows)
        assert sum(1 for x in lines[1:-1] if not x) == nzt + 1          # after each j, and at the end
        if name in ("r", "z", "rr"):                                     # coordinates: no strips (:590-592)
            continue
        for piece, count in (("rL", g * nzt), ("rR", g * nzt), ("zL", nrt * g), ("zR", nrt * g)):
            sl = open(os.path.join(directory, "%s_bound_%s_b0l%d" % (name, piece, lev))).read().split("\n")
            assert len(sl[0]) == 69 and sl[1] == ""
            body = [x for x in sl[2:] if x]
            assert len(body) == count and all(len(x) == 69 for x in body)


def test_restart_reads_the_files_and_continues(oaxis_mod, tmp_path):
    oaxis = oaxis_mod
    text = CASES["schwarzschild_2levels_rk4"][0]
    p = oaxis.params_from_text(text)
    d = str(tmp_path / "checkpoint_t=0.2000")
    names = oaxis.state_fields(p)
    with oaxis.Spacetime(p) as a:
        a.initial_data()
        a.initial_aux()
        a.evolve(2)
        a.checkpoint_write(d, "schwarz.par")
        clocks = [a.time(lev)[:3] for lev in range(a.nlevels)]
        listed = [x.strip() for x in open(os.path.join(d, "checklist")).read().split("\n") if x.strip()]
        assert listed[:4] == ["schwarz.par", "r", "z", "rr"] and sorted(listed[4:]) == sorted(names)
        params = dict((x.split("=")[0].strip(), x.split("=")[1].strip())
                      for x in open(os.path.join(d, "checkparam")).read().split("\n") if "=" in x)
        assert params["nproctot"] == "1" and params["eqsym"] == "F" and float(params["dr0"]) == 0.1
        assert int(params["Nl0"]) == 1 and params["mattertype"] == "vacuum"
        b = _restarted(oaxis, text, d)
        for lev in range(a.nlevels):
            info = a.level_info(lev)
            _check_layout(d, ["r", "rr"] + names, info, lev)
            nrt, nzt = info["Nr"] + info["ghost"], info["Nz"] + info["ghost"]
            for name in names:
                assert np.array_equal(b.download(name, lev), _dump(d, name, lev, nrt, nzt)), (lev, name)
            got = b.time(lev)[:3]
            assert np.allclose(got, clocks[lev], rtol=1e-15, atol=0.0), (lev, got, clocks[lev])
            assert b.time(lev)[3] == 0                                   # step counters restart at 0
        b.initial_aux()
        a.evolve(3)
        b.evolve(3)
        worst = 0.0
        for lev in range(a.nlevels):
            for name in names:
                x, y = a.download(name, lev), b.download(name, lev)
                err = np.abs(x - y).max() / max(np.abs(x).max(), 1.0)
                worst = max(worst, err)
                assert err < 1e-7, (lev, name, err)
        print("restart from the reference-format directory: largest relative difference %.2e after 3 steps" % worst)
        b.close()


def test_gamma_driver_restart_sets_dtbeta(oaxis_mod, tmp_path):
    oaxis = oaxis_mod
    text = CASES["scalar_pulse_gd2_levels"][0]
    p = oaxis.params_from_text(text)
    d = str(tmp_path / "ck")
    with oaxis.Spacetime(p) as a:
        a.initial_data()
        a.initial_aux()
        a.evolve(1)
        a.checkpoint_write(d)
        b = _restarted(oaxis, text, d)
        for lev in range(a.nlevels):
            info = a.level_info(lev)
            nrt, nzt = info["Nr"] + info["ghost"], info["Nz"] + info["ghost"]
            for name in ("scalar_phi", "scalar_pi", "dtbeta_r", "beta_z"):
                assert np.array_equal(b.download(name, lev), _dump(d, name, lev, nrt, nzt)), (lev, name)
        b.initial_aux()
        csi = p.drivercsi
        for lev in range(a.nlevels):
            for c in ("r", "z"):
                assert np.array_equal(b.download("dtbeta_" + c, lev), csi * b.download("Delta_" + c, lev))
        b.evolve(1)
        b.close()


def test_other_grids_and_missing_arrays_are_refused(oaxis_mod, tmp_path):
    oaxis = oaxis_mod
    text = CASES["schwarzschild_2levels_rk4"][0]
    d = str(tmp_path / "ck")
    with oaxis.Spacetime(oaxis.params_from_text(text)) as a:
        a.initial_data()
        a.initial_aux()
        a.checkpoint_write(d)
    other = text.replace("dr0 = 0.1", "dr0 = 0.2").replace("idata = schwarzschild", "idata = checkpoint")
    with oaxis.Spacetime(oaxis.params_from_text(other)) as b:
        with pytest.raises(oaxis.OaxisError, match="grid spacing"):
            b.restart(d)
    path = os.path.join(d, "checklist")
    lines = open(path).read().split("\n")
    open(path, "w").write("\n".join(x for x in lines if x.strip() != "KTA") + "\n")
    with oaxis.Spacetime(oaxis.params_from_text(text.replace("idata = schwarzschild", "idata = checkpoint"))) as b:
        with pytest.raises(oaxis.OaxisError, match="lacks"):
            b.restart(d)
