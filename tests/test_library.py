"""The C-ABI library loads on the CPU host and exports every entry point that
include/florence_amd.h declares, with the argument lists the binding uses
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from florence_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    h = open(os.path.join(ROOT, "include", "florence_amd.h")).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    return re.findall(r"^[\w][\w\s\*]*?\b(fl_\w+)\(([^;]*)\);", h, re.M)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 10
    for name, _ in decl:
        assert hasattr(lib, name), name


def test_binding_matches_header_arity():
    for name, args in _declared():
        n = 0 if args.strip() in ("", "void") else len(args.split(","))
        assert name in _lib.SIGNATURES, name
        assert len(_lib.SIGNATURES[name][1]) == n, (name, n, len(_lib.SIGNATURES[name][1]))


def test_host_only_queries():
    lib = _lib.load()
    assert lib.fl_version() >= 100
    assert lib.fl_pattern_workspace_size(1000) > 0
    # workspace grows with elements; unsupported element sizes report 0
    assert lib.fl_df_workspace_size(216, 8, 125) > lib.fl_df_workspace_size(216, 1, 125) > 0
    assert lib.fl_df_workspace_size(8, 1, 200) == 0
    assert lib.fl_df_residual_workspace_size(11, 10, 10) > 0
    assert lib.fl_set_tuning(99, 1) != 0
    assert b"unknown key" in lib.fl_last_error()
