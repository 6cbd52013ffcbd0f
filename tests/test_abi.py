"""The C-ABI library builds, loads and exports every entry point include/*.h
declares (no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+[\s\*]+(cgli_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    for must in ("cgli_create", "cgli_flux_x", "cgli_flux_y", "cgli_flux_z", "cgli_viscous",
                 "cgli_dual_time_stepping", "cgli_dplusgs", "cgli_hyperplane"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from cgli_fu_flow_amd import native
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("libcgli.so not built")
    lib = C.CDLL(native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_matches_header():
    """native.Config mirrors cgli_config field by field."""
    from cgli_fu_flow_amd import native
    text = open(os.path.join(ROOT, "include", "cgli.h")).read()
    body = text[text.index("typedef struct {"): text.index("} cgli_config;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        m = re.match(r"\s*(?:typedef struct \{)?\s*(int|double)\s+(.*)", decl, flags=re.S)
        if m:
            fields += [(n.strip(), m.group(1)) for n in m.group(2).split(",")]
    ours = [(n, "int" if t is C.c_int else "double") for n, t in native.Config._fields_]
    assert ours == fields


def test_reference_signature_shims_compile_as_a_reference_unit(tmp_path):
    """include/cgli_reference_api.hpp defines the reference's routine names
    with its parameter types; a unit that defines Y_m/Z_m (Resolution.h)
    compiles against it, one without them is refused."""
    import subprocess
    hdr = os.path.join(ROOT, "include", "cgli_reference_api.hpp")
    text = open(hdr).read()
    for name in ("Flux_X", "Flux_Y", "Flux_Z", "Viscous_terms", "Dual_time_stepping", "DPLUSGS", "Filter",
                 "Hyperplane"):
        assert re.search(r"inline void %s\(int myid" % name, text), name
    probe = os.path.join(ROOT, "tests", "ref_api", "shim_probe.cpp")
    ok = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                         "-DY_m=20", "-DZ_m=14", probe], capture_output=True, text=True)
    assert ok.returncode == 0, ok.stderr
    src = tmp_path / "no_dims.cpp"
    src.write_text('#include "cgli_reference_api.hpp"\n')
    bad = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                         capture_output=True, text=True)
    assert bad.returncode != 0 and "Resolution.h" in bad.stderr
