"""The C-ABI library: it loads, exports every function the public header
declares, and agrees with the Python mirror on the parameter block layout.
No compute call is made here (this runs without a GPU)."""
import ctypes
import re

import pytest

import oaxis
from oaxis import _abi


def test_library_loads(oaxis_mod):
    assert oaxis_mod.lib() is not None


def test_every_declared_symbol_is_exported(oaxis_mod):
    names = _abi.declared_symbols()
    assert len(names) >= 25, names
    lib = ctypes.CDLL(_abi.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, "declared in include/oaxis.h but not exported: %s" % missing


def test_params_struct_size_matches_header(oaxis_mod):
    assert ctypes.sizeof(_abi.Params) == oaxis_mod.lib().oa_params_size()


def _c_get(oaxis_mod, p, key):
    buf = ctypes.create_string_buffer(128)
    rc = oaxis_mod.lib().oa_params_get(ctypes.byref(p), key.encode(), buf, 128)
    return rc, buf.value.decode()


def test_params_layout_agrees_with_compiled_code(oaxis_mod):
    """Every field written through the ctypes mirror is read back by the C
    parser at the same offset (the mirror is generated from the header, the
    reader is the compiled library)."""
    checked = 0
    for k, (name, ctype) in enumerate(_abi.Params._fields_):
        p = oaxis_mod.default_params()
        rc, before = _c_get(oaxis_mod, p, name)
        if rc != 0:
            continue                      # a header field that is not a reference key
        if ctype is ctypes.c_double:
            setattr(p, name, 1000.0 + k + 0.25)
            rc, text = _c_get(oaxis_mod, p, name)
            assert float(text) == 1000.0 + k + 0.25, name
        else:
            if re.fullmatch(r"-?\d+", before):
                setattr(p, name, 7000 + k)
                rc, text = _c_get(oaxis_mod, p, name)
                assert int(text) == 7000 + k, name
            elif before in (".true.", ".false."):
                setattr(p, name, 0 if before == ".true." else 1)
                rc, text = _c_get(oaxis_mod, p, name)
                assert text == (".false." if before == ".true." else ".true."), name
            else:
                continue                  # string choice: covered by the parser tests
        checked += 1
    assert checked >= 60, checked


def test_enums_parsed_from_header():
    e = _abi.ENUMS
    assert e["OA_RK4"] == 1 and e["OA_ICN"] == 0
    assert e["OA_ORDER_FOUR"] == 4 and e["OA_ORDER_TWO"] == 2
    assert e["OA_OK"] == 0


def test_create_without_device_fails_cleanly(oaxis_mod):
    """Without a HIP device the library refuses to run (no CPU fallback)."""
    import os
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible; covered by the GPU tests")
    p = oaxis_mod.params_from_text("order = four\nNrtotal = 16\nNztotal = 16\n")
    with pytest.raises(oaxis_mod.OaxisError, match="GPU"):
        oaxis_mod.Spacetime(p)


def test_fortran_module_binds_every_declared_function(oaxis_mod):
    """The ISO_C_BINDING module gives the Fortran main loop every entry point
    of the C header, under the same name."""
    import os
    import re
    from oaxis import _abi
    src = open(os.path.join(os.path.dirname(_abi.HEADER), "..", "ollinaxis-bib_amd", "fortran",
                            "oaxis_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(oa_\w+)"\)', src))
    missing = [n for n in oaxis_mod._abi.declared_symbols() if n not in bound]
    assert not missing, missing


def test_fortran_module_exports_every_binding():
    """Every bound entry point is public in the module."""
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    src = open(os.path.join(here, "..", "ollinaxis-bib_amd", "fortran", "oaxis_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(oa_\w+)"\)', src))
    public = set()
    for block in re.findall(r"public ::((?:[^\n]*&\n)*[^\n]*)", src):
        public |= set(re.findall(r"oa_\w+", block))
    assert bound <= public, sorted(bound - public)
