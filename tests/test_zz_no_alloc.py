"""GPU: oa_norms0d_noalloc, the 0D norms that allocate no device memory.

Called through the C ABI (ctypes) with an output buffer in device memory
that the test allocates itself; every array's five values must equal what
oa_norms0d returns for it (max, min and nm1 exactly; nm2 to 1e-15 relative,
same block partials and the same block-order sum; NaN flag 0).  Two calls in
a row, the second on a smaller level with fewer arrays, so the second reads
only what it wrote into the shared device array.  Over the limit the call is
refused without touching the device."""
import ctypes

import numpy as np
import pytest

from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(oaxis, st, level, names, dev_out, hip):
    lib = oaxis.lib()
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    rc = lib.oa_norms0d_noalloc(st._h, 0, level, len(names), arr, dev_out)
    assert rc == 0, lib.oa_last_error().decode()
    st.sync()
    host = np.zeros(5 * len(names))
    assert hip.hipMemcpy(host.ctypes.data, dev_out, host.nbytes, 2) == 0   # device to host
    return host.reshape(len(names), 5)


def _check(st, level, names, got):
    for k, name in enumerate(names):
        ref = st.norms0d(name, level)
        print("  %-12s level %d: max %.17g min %.17g nm1 %.17g nm2 %.17g" % (name, level, *got[k, :4]),
              flush=True)
        assert got[k, 0] == ref["max"], name
        assert got[k, 1] == ref["min"], name
        assert got[k, 2] == ref["nm1"], name
        assert abs(got[k, 3] - ref["nm2"]) <= 1e-15 * abs(ref["nm2"]), (name, got[k, 3], ref["nm2"])
        assert got[k, 4] == 0.0, name


def test_noalloc_norms_equal_norms0d(oaxis_mod):
    oaxis = oaxis_mod
    hip = _hip()
    dev_out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev_out), 5 * 8 * 32) == 0
    try:
        big = oaxis.params_from_text(CASES["scalar_pulse_gd2_levels"][0])
        small = oaxis.params_from_text(CASES["schwarzschild_rk4_noeqsym"][0])
        with oaxis.Spacetime(big) as st_big, oaxis.Spacetime(small) as st_small:
            for st in (st_big, st_small):
                st.initial_data()
                st.initial_aux()
                st.evolve(2)
            # phi is not kept once a level's clock has left 0 (chi is evolved)
            names_big = [n for n in oaxis.state_fields(big) if n != "phi"]
            print("first call: %d arrays, level 0 of %s" % (len(names_big), st_big.shape(0)), flush=True)
            _check(st_big, 0, names_big, _noalloc(oaxis, st_big, 0, names_big, dev_out, hip))
            names_small = ["alpha", "trK", "Delta_r"]
            print("second call: %d arrays, level 0 of %s" % (len(names_small), st_small.shape(0)), flush=True)
            _check(st_small, 0, names_small, _noalloc(oaxis, st_small, 0, names_small, dev_out, hip))
            # a finer level of the first spacetime, after the smaller call
            _check(st_big, 1, names_small, _noalloc(oaxis, st_big, 1, names_small, dev_out, hip))
            # over the limit: refused before any launch
            lib = oaxis.lib()
            many = (ctypes.c_char_p * 33)(*([b"alpha"] * 33))
            assert lib.oa_norms0d_noalloc(st_big._h, 0, 0, 33, many, dev_out) == oaxis.ENUMS["OA_ERR_ARG"]
            assert lib.oa_norms0d_noalloc(st_big._h, 0, 0, 0, many, dev_out) == oaxis.ENUMS["OA_ERR_ARG"]
    finally:
        hip.hipFree(dev_out)
