"""The C-ABI library loads and exports every entry point that
include/cel_gpu.h declares (no compute without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cel_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cel_gpu_\w+)\s*\(", text)))


def test_header_declarations_match_python_front_end():
    import cel_v3_amd as cel

    assert declared_functions() == sorted(cel.EXPORTS)


def test_library_exports_every_declared_symbol():
    import cel_v3_amd as cel

    if not os.path.exists(cel.LIB_PATH):
        pytest.skip("libcel_gpu.so not built")
    L = C.CDLL(cel.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing


def test_front_end_refuses_a_missing_library(tmp_path, monkeypatch):
    import cel_v3_amd as cel

    monkeypatch.setattr(cel, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(cel, "_lib", None)
    with pytest.raises(ImportError):
        cel.lib()


def test_default_parameters_are_the_reference_settings():
    import cel_v3_amd as cel

    if not os.path.exists(cel.LIB_PATH):
        pytest.skip("libcel_gpu.so not built")
    p = cel.params()
    assert p.x0_mode == cel.X0_ONE and p.tol_mode == cel.TOL_ABS
    assert p.iter_max == 1000 and p.tol == 1e-8
    assert p.format == cel.FMT_CSR and p.precision == cel.PREC_FP64


def test_status_without_a_device_is_an_error_code():
    import cel_v3_amd as cel

    if not os.path.exists(cel.LIB_PATH):
        pytest.skip("libcel_gpu.so not built")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    assert cel.lib().cel_gpu_init(0) == -3  # CEL_ERR_NO_DEVICE, no abort
