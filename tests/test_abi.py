"""The C ABI: libbsn_hip.so loads on a machine without a GPU and exports every
function include/bsn.h declares (no compute calls here)."""
import ctypes
import pathlib
import re

import binary_supernova_amd as bsn

ROOT = pathlib.Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "bsn.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsn_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(bsn.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(bsn.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    lib = bsn.load()
    assert lib.bsn_version() >= 10000
    assert isinstance(lib.bsn_last_error(), bytes)


def test_bc_masks():
    assert bsn.bc_mask("outflow") == 0 and bsn.bc_mask("periodic") == 7
    assert bsn.bc_mask(["outflow", "periodic", "periodic"]) == 6


def header_arg_counts():
    text = (ROOT / "include" / "bsn.h").read_text() + (ROOT / "include" / "bsn_io.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(bsn_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        args = args.strip()
        out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_julia_ccalls_match_the_header():
    # every ccall of the Julia binding names a declared function with the
    # declared number of arguments (the binding cannot be run here)
    src = (ROOT / "binary-supernova_amd" / "julia" / "BinarySupernovaHIP.jl").read_text()
    decl = header_arg_counts()
    calls = re.findall(r"ccall\(\(:(bsn_[a-z0-9_]+),\s*libbsn(?:_io)?\),\s*\w+,\s*\(([^()]*(?:\([^()]*\)[^()]*)*)\)",
                       src, flags=re.S)
    assert len(calls) == len(re.findall(r"ccall\(", src)) >= 30
    for name, types in calls:
        assert name in decl, name
        t = types.strip().rstrip(",")
        n = 0 if not t else t.count(",") + 1
        assert n == decl[name], (name, n, decl[name])
