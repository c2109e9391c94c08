"""The C-ABI library loads and exports every function include/ibem3d.h declares (no GPU calls)."""
import ctypes
import os
import re

from ibemproject_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ibem3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ibem3d_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_binding():
    assert set(engine.EXPORTS) == set(declared())


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_invalid_problem_is_rejected_without_gpu():
    lib = engine.load()
    assert not lib.ibem3d_create(None, 0)
    assert b"invalid" in lib.ibem3d_last_error(None)


def test_binding_declares_every_signature():
    """Every function of the binding has its argument types set: ctypes would otherwise pass a
    context pointer as a 32-bit int."""
    lib = engine.load()
    for name in engine.EXPORTS:
        f = getattr(lib, name)
        assert f.argtypes is not None, name
