"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/semtex_amd.h declares, and its host-only entry points (sx_gll)
agree with the oracle.  No GPU compute is called."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from semtex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "semtex_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding table out of sync with header"


@pytest.mark.parametrize("np_", [3, 8, 10, 12, 16])
def test_library_gll_equals_oracle(oracle, np_):
    L = _lib.load()
    z, w, D = np.zeros(np_), np.zeros(np_), np.zeros(np_ * np_)
    assert L.sx_gll(np_, z.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                    D.ctypes.data_as(C.c_void_p)) == 0
    zo, wo, Do, _ = oracle.gll(np_)
    assert np.max(np.abs(z - zo)) < 4e-15
    assert np.max(np.abs(w - wo)) < 4e-15
    assert np.max(np.abs(D.reshape(np_, np_) - Do)) < 1e-12


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.SxError):
        _lib.load()


def test_cpp_glue_compiles_and_links():
    """The C++ caller of the ABI (tests/abi_glue.cpp, the calls of the
    drop-in glue) compiles against include/semtex_amd.h alone and links to
    the library; it is run against the oracle under -m gpu."""
    from glue_build import build_glue
    exe = build_glue(force=True)
    assert os.access(exe, os.X_OK)
