"""The C-ABI libraries load and export every function include/*.h declares.

CPU only: no compute call is made (there may be no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")

LIBS = {
    "spc5_hip.h": os.path.join(ROOT, "spc5_amd", "libspc5_hip.so"),
    "spc5_synth.h": os.path.join(ROOT, "spc5_amd", "libspc5_synth.so"),
}

_DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    names = set()
    for m in _DECL.finditer(text):
        name = m.group(1)
        if name in {"if", "for", "while", "return", "sizeof", "defined"}:
            continue
        names.add(name)
    return sorted(names)


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_symbol(header):
    so = LIBS[header]
    assert os.path.exists(so), f"{so} not built"
    names = declared_functions(header)
    assert len(names) >= 5
    missing = sorted(set(names) - exported(so))
    assert not missing, f"{os.path.basename(so)} lacks {missing}"
    lib = C.CDLL(so)
    for n in names:
        getattr(lib, n)


def test_hip_header_declares_reference_abi():
    names = set(declared_functions("spc5_hip.h"))
    for r in (1, 2, 4, 8):
        for t in ("double", "float"):
            assert f"core_SPC5_{r}rVc_Spmv_{t}" in names
    assert {"SPC5_opti_merge_double", "SPC5_opti_merge_float"} <= names


def test_python_binding_covers_header():
    from spc5_amd._lib import SIGNATURES
    assert set(declared_functions("spc5_hip.h")) == set(SIGNATURES)


def test_python_constants_match_header():
    """Every SPC5_HIP_* constant the header defines has the same value in
    the Python binding (value types, status codes, kernel options)."""
    from spc5_amd import _lib
    text = open(os.path.join(INCLUDE, "spc5_hip.h")).read()
    defs = dict(re.findall(r"^#define\s+(SPC5_HIP_[A-Z0-9_]+)\s+(-?\d+)\b", text, flags=re.M))
    assert {"SPC5_HIP_OPT_ROW_SORT", "SPC5_HIP_OPT_SPMM_COLUMNS", "SPC5_HIP_OPT_TILE_VALUES",
            "SPC5_HIP_OPT_SPMM_X16"} <= set(defs)
    assert len(set(v for k, v in defs.items() if k.startswith("SPC5_HIP_OPT_"))) == \
        len([k for k in defs if k.startswith("SPC5_HIP_OPT_")]), "option numbers must differ"
    for name, value in defs.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == int(value), name
    for name in ("SPC5_HIP_OPT_ROW_SORT", "SPC5_HIP_OPT_SPMM_COLUMNS", "SPC5_HIP_OPT_TILE_VALUES",
                 "SPC5_HIP_OPT_SPMM_X16", "SPC5_HIP_F64", "SPC5_HIP_F32", "SPC5_HIP_BF16", "SPC5_HIP_OK"):
        assert hasattr(_lib, name), name


def test_library_loads_and_reports_version():
    import spc5_amd
    assert spc5_amd.lib().spc5_hip_version().startswith(b"spc5_hip")
    assert spc5_amd.lib().spc5_hip_last_error() is not None


def test_invalid_arguments_are_reported_without_gpu_work():
    from spc5_amd._lib import lib
    h = C.c_void_p()
    rc = lib().spc5_hip_mat_create(0, 8, 3, 10, 10, 0, 0, None, None, None, None, C.byref(h))
    assert rc == 1
    assert b"r must be" in lib().spc5_hip_last_error()
    rc = lib().spc5_hip_mat_create(1, 8, 4, 10, 10, 0, 0, None, None, None, None, None)
    assert rc == 1
