"""The C-ABI library loads and exports every symbol include/qasmhd.h
declares (no device calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "qasmhd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qas_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_core_entry_points():
    d = declared()
    for n in ("qas_create", "qas_destroy", "qas_upload", "qas_download", "qas_bound", "qas_step", "qas_cfl_dt",
              "qas_run", "qas_last_error"):
        assert n in d


def test_library_exports_declared_symbols():
    from qasmhd_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libqasmhd_amd.so not built here")
    import ctypes
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= set(declared())


def test_integration_guide_names_every_entry_point():
    """Each declared entry point has a line in the integration guide (what
    it replaces, or that it has no counterpart in the reference)."""
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    missing = [n for n in declared() if not re.search(r"\b%s\b" % n, guide)]
    assert not missing, missing


def test_no_fallback_without_library_or_device():
    """The product path raises rather than computing anything on the host:
    a missing library fails at load, and on a host with no GPU the context
    creation reports HIP's error."""
    import subprocess
    import sys
    from qasmhd_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libqasmhd_amd.so not built here")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    code = ("from qasmhd_amd import capi, problems\n"
            "try:\n    capi.load('/nonexistent/libqasmhd_amd.so')\nexcept capi.QasError as e:\n    print('E1', e)\n"
            "try:\n    capi.solver_for(problems.brio_wu(64))\nexcept capi.QasError as e:\n    print('E2', e)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "E1 libqasmhd_amd.so is not built" in r.stdout
    assert "E2 qas_create:" in r.stdout and "hipErrorNoDevice" in r.stdout
