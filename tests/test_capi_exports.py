"""The C ABI library loads and exports every entry point include/gofish_capi.h
declares (no GPU needed: nothing is called)."""
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "gofish_capi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void|const char \*)\s*\*?\s*(gf_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    from parallelpopgen_amd import capi
    lib = capi.load()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if getattr(lib, n, None) is None]
    assert not missing, missing
    # and the ctypes table covers every declared symbol
    assert set(names) <= set(capi.SIGNATURES), set(names) - set(capi.SIGNATURES)


def test_error_path_without_device_call():
    from parallelpopgen_amd import capi
    lib = capi.load()
    h = capi.C.c_void_p()
    assert lib.gf_run(None, capi.C.byref(h)) == -1
    assert b"NULL" in lib.gf_last_error()
    assert lib.gf_set_generations_per_launch(-1) == -1
    assert lib.gf_set_generations_per_launch(65) == -1
    assert lib.gf_set_generations_per_launch(0) == 0    # back to the default
