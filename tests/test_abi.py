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


def build_c_smoke(tmp_path):
    """gcc builds tests/c_abi_smoke.c against the library, the way a cgo
    binding links it (plain C, no Python in the loop)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not on PATH")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    lib = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp_path / "c_abi_smoke")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           os.path.join(ROOT, "tests", "c_abi_smoke.c"), "-o", exe,
           "-L", lib, "-lgravitas_hip", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath," + lib, "-Wl,-rpath," + os.path.join(rocm, "lib"), "-lm"]
    subprocess.run(cmd, check=True)
    return exe


def run_c_smoke(exe):
    import subprocess
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_c_harness_builds_and_reports_status(tmp_path):
    out = run_c_smoke(build_c_smoke(tmp_path))
    assert out.startswith(("no device:", "ok:")), out


@pytest.mark.gpu
def test_c_harness_known_answers_on_gpu(gpu, tmp_path):
    out = run_c_smoke(build_c_smoke(tmp_path))
    assert out.startswith("ok:"), out
