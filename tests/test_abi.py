"""The C-ABI library loads on a CPU-only host and exports every symbol the
public header declares (no compute calls here)."""
import ctypes
import os
import re

from soap_gpu_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "soap_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(soap_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_table():
    assert header_symbols() == sorted(E.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(E.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_header_is_plain_c(tmp_path):
    """The boundary is C: the header compiles as C99 with gcc -pedantic (no C++ or
    HIP types), every declared struct usable from C."""
    import shutil
    import subprocess
    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "use_header.c"
    src.write_text('#include "soap_gpu.h"\n'
                   "int main(void) { soap_star s; soap_activity a; soap_region r; soap_diff_params d;\n"
                   "  soap_bis_inject_params b; soap_phoenix_params p;\n"
                   "  (void)s; (void)a; (void)r; (void)d; (void)b; (void)p; return 0; }\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                        os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "use_header.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_version_string_without_gpu():
    L = E.load()
    assert b"gfx950" in L.soap_version()


def test_ctx_create_reports_error_without_device():
    import torch
    if torch.cuda.is_available():
        return
    L = E.load()
    ctx = ctypes.c_void_p()
    assert L.soap_ctx_create(0, ctypes.byref(ctx)) != 0
