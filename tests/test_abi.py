"""CPU checks of the C-ABI library: it loads and exports every symbol that
include/qws_mi355x.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

import qws_amd as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function():
    L = Q.lib()
    names = Q.declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_null_handles_are_refused():
    """Entry points given a null lattice or server handle return 1 (invalid argument)
    with a message, before any device work."""
    L = Q.lib()
    L.qwsg_set_kappa.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.qwsg_volume.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
    L.qwsg_set_sap_fused.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.qwsg_server_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.qwsg_last_error.restype = ctypes.c_char_p
    v = ctypes.c_longlong(-1)
    for rc in (L.qwsg_set_kappa(None, 0.1), L.qwsg_volume(None, ctypes.byref(v)), L.qwsg_set_sap_fused(None, 1)):
        assert rc == 1
        assert b"null lattice handle" in L.qwsg_last_error()
    assert v.value == -1
    assert L.qwsg_server_run(None, None, 0, None) == 1
    assert b"null server handle" in L.qwsg_last_error()
    # null spinor pointers are refused before the handle is used
    L.qwsg_dslash.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_int]
    assert L.qwsg_dslash(None, 0, None, None, 0, -1) == 1
    assert b"qwsg_dslash: null pointer argument" in L.qwsg_last_error()
    best = ctypes.c_int(-1)
    assert L.qwsg_autotune_blocks(None, None, None, None, 1, ctypes.byref(best), None) == 1
    assert b"qwsg_autotune_blocks: null pointer argument" in L.qwsg_last_error()
    assert best.value == -1


def test_library_exports_qws_data_symbols():
    L = Q.lib()
    txt = open(Q.HEADER).read()
    data = re.findall(r"^extern\s+[^(\n\"]*?;", txt, flags=re.M)
    names = set()
    for decl in data:
        body = decl[len("extern"):].rstrip(";")
        body = re.sub(r"\[[^\]]*\]", "", body)
        for part in body.split(",")[0:]:
            tok = part.replace("*", " ").split()
            if tok:
                names.add(tok[-1])
    for n in ["glud", "glus", "clvd", "clvs", "kappa", "mkappa", "kappa2", "nx", "vold", "vols", "npe", "fbc",
              "domain_e", "domain_o", "rank"]:
        assert n in names, n
    for n in names:
        ctypes.c_int.in_dll(L, n)  # raises ValueError if the symbol is absent


def test_no_oracle_in_product():
    """The product sources never reference the oracle."""
    for d, _, files in os.walk(os.path.join(ROOT, "qws_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                txt = open(os.path.join(d, f)).read()
                assert "oracle" not in txt.lower(), f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(Q, "_lib", None)
    monkeypatch.setattr(Q, "LIBPATH", str(tmp_path / "absent.so"))
    with pytest.raises(Q.QwsError):
        Q.lib()
