"""The C ABI: libmcm.so loads and exports every function include/mcm_abi.h
declares; the ctypes mirror (mcmatlab_amd/abi.py) has the C compiler's
struct sizes and field offsets; entry points fail cleanly without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from mcmatlab_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcm_abi.h")
LIB = os.path.join(ROOT, "mcmatlab_amd", "libmcm.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(mcm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    fns = header_functions()
    for f in ("mcm_mc_run", "mcm_heat_run", "mcm_mc_plan_create", "mcm_heat_plan_advance", "mcm_last_error"):
        assert f in fns


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcm.so not built")
def test_library_exports_every_header_symbol():
    lib = C.CDLL(LIB)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    from mcmatlab_amd import _lib

    assert set(header_functions()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


STRUCTS = ["mcm_beam_distr", "mcm_mc_input", "mcm_mc_output", "mcm_mc_sizes", "mcm_mc_item", "mcm_mc_multi_item", "mcm_mc_grad_input",
           "mcm_mc_grad_output", "mcm_heat_input", "mcm_heat_output", "mcm_heat_segment", "mcm_heat_adjoint_output"]


def test_ctypes_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mcm_abi.h"', "int main(void){"]
    for s in STRUCTS:
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for name, _ in getattr(abi, s)._fields_:
            lines.append(f'printf("{s}.{name} %zu\\n", offsetof({s}, {name}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    for s in STRUCTS:
        st = getattr(abi, s)
        assert int(got[s]) == C.sizeof(st), s
        for name, _ in st._fields_:
            assert int(got[f"{s}.{name}"]) == getattr(st, name).offset, f"{s}.{name}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcm.so not built")
def test_entry_points_report_errors_without_a_device():
    """No GPU in this container: calls return a status and a message (no
    crash, no hang, no CPU fallback)."""
    from mcmatlab_amd import _lib

    L = _lib.lib()
    if L.mcm_device_count() > 0:
        pytest.skip("a device is visible")
    import models

    inp = abi.MCInputs(models.cases()["pencil_bt1"](), seed=1)
    out = abi.MCOutputs(inp)
    rc = L.mcm_mc_run(C.byref(inp.struct), C.byref(out.struct))
    assert rc == abi.MCM_ERR_NO_DEVICE
    assert b"device" in L.mcm_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcm.so not built")
def test_invalid_input_rejected_before_device_use():
    from mcmatlab_amd import _lib
    import models

    L = _lib.lib()
    m = models.cases()["pencil_bt1"]()
    m.MC.M = m.MC.M.copy()
    m.MC.M[0, 0, 0] = 0  # media indices are 1-based
    inp = abi.MCInputs(m, seed=1)
    sz = abi.mcm_mc_sizes()
    assert L.mcm_mc_output_sizes(C.byref(inp.struct), C.byref(sz)) == abi.MCM_ERR_INVALID_ARGUMENT
    m = models.cases()["fiber"]()
    m.MC.LC.res = 3  # fibers have one pixel (runMonteCarlo.m:364-365)
    inp = abi.MCInputs(m, seed=1)
    assert L.mcm_mc_output_sizes(C.byref(inp.struct), C.byref(sz)) == abi.MCM_ERR_INVALID_ARGUMENT


@pytest.mark.skipif(not os.path.exists(LIB), reason="libmcm.so not built")
def test_output_sizes_match_the_reference_layout():
    from mcmatlab_amd import _lib
    import models

    L = _lib.lib()
    for name, dims in (("pencil_bt1", {"NFR": 8000, "NI_xpos": 400, "NI_zneg": 400}), ("pencil_bt2", {"NI_zneg": 10000}),
                       ("lens_timebins", {"image": 8 * 8 * 8}), ("farfield", {"farField": 144}), ("wavelengths3", {"NFR": 24000})):
        inp = abi.MCInputs(models.cases()[name](), seed=1)
        sz = abi.mcm_mc_sizes()
        assert L.mcm_mc_output_sizes(C.byref(inp.struct), C.byref(sz)) == 0
        for k, v in dims.items():
            assert getattr(sz, k) == v, (name, k, getattr(sz, k))
