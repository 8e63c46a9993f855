"""Host-side checks of the C ABI: the library loads and exports every entry
point declared in include/densecrf_ho_amd.h; argument validation and the
error channel work without touching a GPU."""

import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "densecrf_ho_amd.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dcrf_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "dcrf_lattice_create" in names and "dcrf_crf_qp" in names and "dcrf_crf_lp" in names


def test_library_exports_every_declared_symbol():
    from densecrf_ho_amd import _lib
    lib = C.CDLL(_lib.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from densecrf_ho_amd import _lib
    assert set(declared()) == set(_lib.SIGNATURES)


def test_lp_param_defaults_match_reference():
    # LP_inf_params::LP_inf_params(), src/densecrf.cpp:287-302
    from densecrf_ho_amd import _lib
    L = _lib.load()
    p = _lib.LPParams()
    L.dcrf_lp_params_default(C.byref(p))
    assert abs(p.prox_reg_const - 1e-2) < 1e-9 and p.dual_gap_tol == 1e3 and p.prox_energy_tol == 1e3
    assert (p.prox_max_iter, p.fw_max_iter, p.qp_max_iter) == (10, 10, 1000)
    assert p.qp_tol == 1e3 and abs(p.qp_const - 1e-16) < 1e-22 and p.best_int == 1
    assert abs(p.confidence_tol - 0.95) < 1e-7 and p.less_confident_percent == 0


def test_invalid_arguments_report_errors():
    from densecrf_ho_amd import _lib
    L = _lib.load()
    h = C.c_void_p()
    assert L.dcrf_crf_create(0, 21, C.byref(h)) == 1
    assert b"positive" in L.dcrf_last_error()
    assert L.dcrf_crf_create(10, 65, C.byref(h)) == 1
    assert b"64" in L.dcrf_last_error()
    assert L.dcrf_lattice_create(None, 10, 2, 0, C.byref(h)) == 1


def test_python_api_raises_engine_error():
    import densecrf_ho_amd as D
    with pytest.raises(D.EngineError):
        D.DenseCRF(0, 5)


def test_cpp_interface_compiles():
    """include/densecrf_ho_amd.hpp compiles on its own, and against the
    reference's Eigen when that is present (drop-in with Eigen::MatrixXf)."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    src = "#include \"densecrf_ho_amd.hpp\"\nint main(){ dcrf_ho::LP_inf_params p; return p.fw_max_iter == 10 ? 0 : 1; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, "-x", "c++", "-"], input=src, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr
    eigen = "/root/reference/include"
    if os.path.isdir(os.path.join(eigen, "Eigen")):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DDCRF_USE_EIGEN", "-I", inc, "-I", eigen,
                            "-x", "c++", "-"], input=src, text=True, capture_output=True)
        assert r.returncode == 0, r.stderr
