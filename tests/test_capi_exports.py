"""The C-ABI library loads without a GPU and exports every symbol include/h2h.h declares."""
import ctypes
import subprocess

import h2h


def test_header_declares_the_boundary():
    names = h2h.declared_symbols()
    for must in ("h2h_create", "h2h_contract", "h2h_build_labels", "h2h_query", "h2h_query_device",
                 "h2h_labels_upload", "h2h_server_start", "h2h_submit", "h2h_wait"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(h2h.LIB_PATH)
    for name in h2h.declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", h2h.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(h2h.declared_symbols()) <= exported


def test_null_arguments_return_status_not_crash():
    L = h2h.lib()
    assert L.h2h_create(None, 0) == -1           # H2H_ERR_ARG
    assert L.h2h_get_info(None, None) == -1
    assert L.h2h_last_error(None) == b"null context"
