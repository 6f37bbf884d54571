"""The C-ABI library loads and exports every entry point include/ declares;
host-side error paths that need no GPU.  CPU only."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(pj_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    import proteus_amd as pa
    lib = pa.lib()
    declared = _declared()
    assert len(declared) >= 30
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", pa.LIB_PATH]).decode()
    exported = set(re.findall(r" T (pj_[a-z0-9_]+)", out))
    assert declared <= exported


def test_binding_covers_declared_symbols():
    import proteus_amd as pa
    assert _declared() <= set(pa.exported_symbols())


def test_library_is_gfx950_code_object():
    import proteus_amd as pa
    blob = open(pa.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_context_create_fails_cleanly_without_gpu():
    import proteus_amd as pa
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    with pytest.raises(pa.PjError) as e:
        pa.Context(0)
    assert e.value.code in (pa.PJ_E_HIP, pa.PJ_E_ARG, pa.PJ_E_UNSUPPORTED)
    assert pa.lib().pj_last_error()
