"""The C-ABI library loads and exports every entry point include/qmcrt.h declares; host-only entry
points work without a GPU and report errors through status codes."""
import ctypes as C
import re
from pathlib import Path

import pytest

from solarenergy_chier_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        text = hdr.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+char\s*\*|qmc_status)\s*(qmc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert len(names) >= 50
    assert {"qmc_trace", "qmc_upload", "qmc_smooth", "qmc_scene_load_file"} <= names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() <= set(N.SIGNATURES)


def test_status_strings_and_errors():
    lib = N.lib()
    assert lib.qmc_status_string(N.QMC_E_PARSE) == b"parse error"
    st, msg = N.status_of("qmc_scene_create", None)
    assert st == N.QMC_E_INVALID_ARG and "null" in msg


def test_null_arguments_are_rejected_not_crashing():
    for name, args in [("qmc_scene_process", [None]), ("qmc_trace", [None, None, 0]), ("qmc_sync", [None]),
                       ("qmc_flux_read", [None, 0, None, 0]), ("qmc_smooth", [None, None, 1, 1, 1, C.c_float(0.1)])]:
        st, _ = N.status_of(name, *args)
        assert st == N.QMC_E_INVALID_ARG, name


def test_trace_before_upload_is_a_state_error_or_no_device():
    # On a machine without a GPU qmc_create fails with QMC_E_HIP; with a GPU, tracing an empty ctx is
    # a call-order error. Both are errors, never a silent CPU fallback.
    h = C.c_void_p()
    st, msg = N.status_of("qmc_create", 0, C.byref(h))
    if st != N.QMC_OK:
        assert st == N.QMC_E_HIP, msg
        return
    try:
        ids = (C.c_int32 * 1)(0)
        st, msg = N.status_of("qmc_trace", h, ids, 1)
        assert st == N.QMC_E_STATE, msg
    finally:
        N.lib().qmc_destroy(h)
