"""The C-ABI library loads and exports every symbol include/gravitas_hip.h
declares; without a GPU it reports "no device" instead of crashing."""
import ctypes
import os
import re

import pytest

from gravitas_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gravitas_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gx_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_init_reports_status_without_crashing():
    L = _lib.load()
    buf = ctypes.create_string_buffer(256)
    rc = L.gx_init(buf, 256)
    assert rc in (_lib.GX_OK, _lib.GX_ERR_NO_DEVICE, _lib.GX_ERR_CODE_OBJECT)
    if rc != _lib.GX_OK:
        assert buf.value  # a message, like gpu_init's err buffer
        name = ctypes.create_string_buffer(64)
        L.gx_name(name, 64)
        assert name.value == b""
        assert L.gx_solver_new() is None


def test_null_solver_is_an_argument_error():
    L = _lib.load()
    assert L.gx_solver_config(None, 0.5, 0, 0, 0) == _lib.GX_ERR_ARG
    assert L.gx_sync(None) == _lib.GX_ERR_ARG
    assert L.gx_last_seconds(None) == 0.0


def test_product_does_not_import_the_oracle():
    """The product path never routes through the oracle or a CPU fallback."""
    pkg = os.path.join(ROOT, "gravitas_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", src).replace("the oracle", ""), f
