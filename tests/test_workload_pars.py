"""The committed par files of the benchmark configurations for seed 0
(ollinaxis-bib_amd/pars/seed0) are what the seeded generators produce, and
every one is a configuration the library accepts."""
import ctypes
import os
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
PARS = os.path.join(ROOT, "ollinaxis-bib_amd", "pars")


def test_committed_pars_match_the_generators(oaxis_mod):
    sys.path.insert(0, PARS)
    import write_pars
    files = write_pars.generate(0)
    assert len(files) == 2 + 64 + 25 + 32
    for rel, text in files.items():
        with open(os.path.join(PARS, "seed0", rel)) as f:
            assert f.read() == text, rel
        if rel.endswith(".par"):
            p = oaxis_mod.params_from_text(text)
            buf = ctypes.create_string_buffer(256)
            rc = oaxis_mod.lib().oa_params_check_supported(ctypes.byref(p), buf, 256)
            assert rc == 0, (rel, buf.value)
