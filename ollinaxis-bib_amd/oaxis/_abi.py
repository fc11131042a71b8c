"""ctypes binding of the C ABI in include/oaxis.h (liboaxis.so).

The library is the HIP implementation; there is no CPU fallback.  Loading
fails with an ImportError naming the missing file when liboaxis.so has not
been built (``python -c "import __graft_entry__; __graft_entry__.build()"``).
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(os.path.dirname(_HERE))
HEADER = os.path.join(_ROOT, "include", "oaxis.h")
LIBPATH = os.path.join(_HERE, "liboaxis.so")

OA_OK = 0


def _parse_params_struct(path=HEADER, name="oa_params"):
    """Field list of a plain struct (``oa_params``, ``oa_horizon``) read from
    the public header, so the ctypes mirror cannot drift from the C
    definition."""
    text = open(path).read()
    body = text[text.index("typedef struct %s {" % name):text.index("} %s;" % name)]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for line in body.split("\n")[1:]:
        line = line.strip()
        if not line:
            continue
        m = re.match(r"(int|double)\s+(.*);", line)
        if not m:
            raise RuntimeError("unexpected line in %s: %r" % (name, line))
        ctype = ctypes.c_int if m.group(1) == "int" else ctypes.c_double
        for field in m.group(2).split(","):
            fields.append((field.strip(), ctype))
    return fields


def _parse_enums(path=HEADER):
    text = open(path).read()
    out = {}
    for block in re.findall(r"enum\s*{([^}]*)}", text):
        value = -1
        for item in block.split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                name, val = [s.strip() for s in item.split("=")]
                value = int(val)
            else:
                name = item
                value += 1
            out[name] = value
    for name, val in re.findall(r"#define\s+(OA_\w+)\s+(\d+)", text):
        out[name] = int(val)
    return out


class Params(ctypes.Structure):
    _fields_ = _parse_params_struct()


class Horizon(ctypes.Structure):
    _fields_ = _parse_params_struct(name="oa_horizon")


ENUMS = _parse_enums()


def load_library(path=LIBPATH):
    if not os.path.exists(path):
        raise ImportError(
            "liboaxis.so not found at %s: build it with __graft_entry__.build() "
            "(the library is HIP-only; there is no CPU fallback)" % path)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER(Params)
    H = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "oa_params_default": (ctypes.c_int, [P]),
        "oa_params_set": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_char_p]),
        "oa_params_parse_file": (ctypes.c_int, [P, ctypes.c_char_p]),
        "oa_params_parse_text": (ctypes.c_int, [P, ctypes.c_char_p]),
        "oa_params_get": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
        "oa_params_check_supported": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int]),
        "oa_params_size": (ctypes.c_size_t, []),
        "oa_last_error": (ctypes.c_char_p, []),
        "oa_create": (ctypes.c_int, [P, ctypes.POINTER(H)]),
        "oa_destroy": (ctypes.c_int, [H]),
        "oa_num_levels": (ctypes.c_int, [H, ip]),
        "oa_level_info": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ip, ip, ip, dp, dp, dp]),
        "oa_upload_field": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, ctypes.c_int]),
        "oa_download_field": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, ctypes.c_int]),
        "oa_get_time": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, dp, dp, dp, ip]),
        "oa_set_time": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int]),
        "oa_saveold": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
        "oa_substep": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double]),
        "oa_finebound": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "oa_auxiliary": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
        "oa_restrict": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
        "oa_advance_time": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
        "oa_initial_aux": (ctypes.c_int, [H]),
        "oa_coarse_step": (ctypes.c_int, [H]),
        "oa_evolve": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_sync": (ctypes.c_int, [H]),
        "oa_initial_data": (ctypes.c_int, [H]),
        "oa_last_elapsed_ms": (ctypes.c_int, [H, dp]),
        "oa_set_kernel_timing": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_rhs_kernel_stats": (ctypes.c_int, [H, dp, ctypes.POINTER(ctypes.c_long)]),
        "oa_point_updates": (ctypes.c_int, [H, dp]),
        "oa_norms0d": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp]),
        "oa_norms0d_noalloc": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_char_p), ctypes.c_void_p]),
        "oa_set_elliptic_exit": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_elliptic_stats": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp]),
        "oa_constraints": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
        "oa_ahfind": (ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(Horizon)]),
        "oa_horizon_shape": (ctypes.c_int, [H, ctypes.c_int, dp, dp, ctypes.c_int, ip]),
        "oa_batch_ahfind": (ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(H), ip, ctypes.POINTER(Horizon)]),
        "oa_ahfind_begin": (ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(H)]),
        "oa_batch_ahfind_begin": (ctypes.c_int, [H, ctypes.c_int, ctypes.POINTER(H), ip, ctypes.POINTER(H)]),
        "oa_ahfind_end": (ctypes.c_int, [H, ctypes.POINTER(Horizon)]),
        "oa_checkpoint_size": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_size_t)]),
        "oa_checkpoint_save": (ctypes.c_int, [H, ctypes.c_void_p, ctypes.c_size_t]),
        "oa_checkpoint_load": (ctypes.c_int, [H, ctypes.c_void_p, ctypes.c_size_t]),
        "oa_checkpoint_restart": (ctypes.c_int, [H, ctypes.c_char_p]),
        "oa_checkpoint_write": (ctypes.c_int, [H, ctypes.c_char_p, ctypes.c_char_p]),
        "oa_params_footprint": (ctypes.c_int, [P, dp]),
        "oa_batch_create": (ctypes.c_int, [ctypes.POINTER(H)]),
        "oa_batch_destroy": (ctypes.c_int, [H]),
        "oa_batch_add": (ctypes.c_int, [H, H]),
        "oa_batch_remove": (ctypes.c_int, [H, H]),
        "oa_batch_size": (ctypes.c_int, [H, ip]),
        "oa_batch_coarse_step": (ctypes.c_int, [H]),
        "oa_batch_evolve": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_batch_sync": (ctypes.c_int, [H]),
        "oa_batch_last_elapsed_ms": (ctypes.c_int, [H, dp]),
        "oa_batch_point_updates": (ctypes.c_int, [H, dp]),
        "oa_batch_launch_stats": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
        "oa_batch_set_kernel_timing": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_batch_set_graphs": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_set_graphs": (ctypes.c_int, [H, ctypes.c_int]),
        "oa_schedule": (ctypes.c_int, [H, ip, ip, ip, dp]),
        "oa_batch_schedule": (ctypes.c_int, [H, ip, ip, ip, dp]),
        "oa_batch_rhs_kernel_stats": (ctypes.c_int, [H, dp, ctypes.POINTER(ctypes.c_long), dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def declared_symbols(path=HEADER):
    """Names of every function the public header declares."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|size_t)\s+(oa_\w+)\s*\(", text, re.M)))
