"""Parameter parsing: the reference's par-file keys, defaults and error
behaviour (src/base/param.f90, src/base/parse.f90), through the C ABI."""
import pytest

import oaxis


def get(p, key):
    import ctypes
    buf = ctypes.create_string_buffer(256)
    rc = oaxis.lib().oa_params_get(ctypes.byref(p), key.encode(), buf, 256)
    assert rc == 0, oaxis.lib().oa_last_error().decode()
    return buf.value.decode()


def test_defaults_follow_param_f90():
    p = oaxis.default_params()
    E = oaxis.ENUMS
    assert p.eqsym == 0                              # param.f90:95
    assert p.dr0 == 1.0 and p.Nrtotal == 10          # param.f90:101,106
    assert p.interporder == E["OA_INTERP_BICUBIC"]   # param.f90:144
    assert p.dtfac == 0.5                            # param.f90:215
    assert p.slicing == E["OA_SLICING_HARMONIC"]     # param.f90:307
    assert p.shift == E["OA_SHIFT_NONE"]             # param.f90:344
    assert p.drivercsi == 0.75                       # param.f90:351
    assert p.integrator == E["OA_ICN"]               # param.f90:410
    assert p.order == E["OA_ORDER_TWO"]              # param.f90:411
    assert p.chimethod == 1                          # param.f90:415
    assert p.icniter == 3                            # param.f90:426
    assert p.eta == 2.0                              # param.f90:429
    assert p.geodiss == 0.01                         # param.f90:436
    assert p.boundtype == E["OA_BOUND_RADIATIVE"]    # param.f90:447
    assert p.mattertype == 0                         # param.f90:612 (vacuum)


SCHWARZ = """
# Schwarzschild puncture, fourth order
idata      = schwarzschild
BH1mass    = 1.0
eqsym      = .false.     # no equatorial symmetry
dr0 = 0.05
dz0 = 5.d-2
Nrtotal = 128
Nztotal = 256
order = four
integrator = rk4
slicing = 1+log
gauge_f = 2.0
ilapse = psiminus2
geodiss = 1.d-1
chimethod = T
ahfind = .false.
ELL_solver = wave
"""


def test_parse_reference_style_text():
    p = oaxis.params_from_text(SCHWARZ)
    E = oaxis.ENUMS
    assert p.idata == E["OA_IDATA_SCHWARZSCHILD"]
    assert p.BH1mass == 1.0
    assert p.dz0 == 0.05 and p.geodiss == 0.1       # Fortran d exponents
    assert (p.Nrtotal, p.Nztotal) == (128, 256)
    assert p.order == E["OA_ORDER_FOUR"] and p.integrator == E["OA_RK4"]
    assert p.slicing == E["OA_SLICING_1PLUSLOG"] and p.ilapse == E["OA_ILAPSE_PSIMINUS2"]
    assert p.chimethod == 1 and p.ahfind == 0


def test_get_round_trips():
    p = oaxis.params_from_text(SCHWARZ)
    assert get(p, "slicing") == "1+log"
    assert get(p, "order") == "four"
    assert get(p, "eqsym") == ".false."
    assert float(get(p, "dr0")) == 0.05
    assert int(get(p, "Nztotal")) == 256


def test_multiple_values_only_for_multi_keys():
    p = oaxis.params_from_text("mattertype = scalar\n")
    assert p.mattertype & oaxis.ENUMS["OA_MATTER_SCALAR"]
    with pytest.raises(oaxis.OaxisError, match="Multiple values"):
        oaxis.params_from_text("Nrtotal = 10, 20\n")


@pytest.mark.parametrize("text, msg", [
    ("nosuchkey = 1\n", "non-existent parameter"),
    ("eqsym = maybe\n", "LOGICAL"),
    ("Nrtotal = 1.5\n", "INTEGER"),
    ("dr0 = fast\n", "REAL"),
    ("slicing = geodesic\n", "Out of range"),
    ("Nrtotal 10\n", "missing equal sign"),
    ("dr0 =\n", "missing value"),
])
def test_parse_errors(text, msg):
    with pytest.raises(oaxis.OaxisError, match=msg):
        oaxis.params_from_text(text)


def test_unstored_reference_keys_are_type_checked():
    # keys of the reference outside the evolution path are accepted and ignored
    oaxis.params_from_text("outparallel = fileperproc\nNoutput1D = 5\n")
    with pytest.raises(oaxis.OaxisError):
        oaxis.params_from_text("WE_epsilon = tiny\n")
    with pytest.raises(oaxis.OaxisError):
        oaxis.params_from_text("Noutput2D = often\n")


def test_elliptic_solver_keys_are_stored():
    p = oaxis.params_from_text("ELL_solver = wave\nWE_epsilon = 1.d-9\nWE_method = rk4\nWE_maxiter = 500\n"
                               "WE_eta = 0.5\nWE_diss = 0.2\nWE_dtfac = 0.4\nELL_maxiter = 70\n")
    assert p.ELL_solver == oaxis.ENUMS["OA_ELL_WAVE"] and p.WE_method == oaxis.ENUMS["OA_RK4"]
    assert (p.WE_epsilon, p.WE_maxiter, p.WE_eta, p.WE_diss, p.WE_dtfac, p.ELL_maxiter) == (1e-9, 500, 0.5, 0.2, 0.4, 70)
    d = oaxis.default_params()     # param.f90:463-491
    assert d.ELL_solver == oaxis.ENUMS["OA_ELL_WAVE"] and d.WE_method == oaxis.ENUMS["OA_ICN"]
    assert (d.WE_epsilon, d.WE_maxiter, d.WE_eta, d.WE_diss, d.WE_dtfac) == (1e-8, 100000, 0.0, 0.1, 0.6)


def supported(text):
    import ctypes
    p = oaxis.params_from_text(text)
    buf = ctypes.create_string_buffer(256)
    rc = oaxis.lib().oa_params_check_supported(ctypes.byref(p), buf, 256)
    return rc, buf.value.decode()


def test_supported_configurations():
    size = "Nrtotal = 32\nNztotal = 32\n"
    assert supported("order = four\n" + size)[0] == 0
    assert supported(SCHWARZ)[0] == 0
    assert supported("order = four\nshift = Gammadriver2\nmattertype = scalar\n" + size)[0] == 0
    assert supported("order = four\nidata = BrillWave\n" + size)[0] == 0
    assert supported("order = four\nmattertype = scalar\nidata = scalarpulse\n" + size)[0] == 0
    # a restart from the reference's checkpoint directory (oa_checkpoint_restart)
    assert supported("order = four\nidata = checkpoint\n" + size)[0] == 0
    # the reference's default 10 x 10 grid is below one 16 x 16 tile
    assert supported("order = four\n")[0] != 0


@pytest.mark.parametrize("text, word", [
    ("order = two\n", "order"),
    ("order = four\nshift = Gammadriver1\n", "shift"),
    ("order = four\nformulation = z4c\n", "formulation"),
    ("order = four\nmattertype = complex\n", "scalar field"),
    ("order = four\nNrtotal = 4\n", "Nr"),
    # initial data the library does not build, and solves it does not run
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = kerr\n", "idata"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = testgw\n", "idata"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = complexpulse\n", "idata"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = bosonstar\n", "idata"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = rotboson\n", "idata"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = BrillWave\nELL_solver = none\n", "elliptic solver"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = BrillWave\nELL_solver = sor\nELL_maxiter = 0\n",
     "ELL_maxiter"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nmattertype = scalar\nidata = scalarpulse\nELL_solver = none\n",
     "elliptic solver"),
    ("order = four\nNrtotal = 32\nNztotal = 32\nidata = scalarpulse\n", "mattertype"),
])
def test_unsupported_configurations_are_refused(text, word):
    rc, msg = supported(text)
    assert rc != 0
    assert word.lower() in msg.lower(), msg
