"""CPU checks of the C ABI: the library loads, exports every entry point that
include/mordred_hip.h declares, and the ctypes structs match the C layout."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from mordred_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mordred_hip.h")
ENGINE_HEADER = os.path.join(ROOT, "include", "mordred_engine.h")
COST_HEADER = os.path.join(ROOT, "include", "mordred_costmodel.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mordred_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 22
    assert set(abi.SIGNATURES) | {"mordred_version"} >= set(names) - {"mordred_version"}


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    for name in declared_functions() + declared_functions(ENGINE_HEADER) + declared_functions(COST_HEADER):
        assert hasattr(lib, name), name
    assert lib.mordred_version().decode().startswith("mordred-hip")


def test_bindings_cover_header():
    assert set(declared_functions()) == set(abi.SIGNATURES) | {"mordred_version"}


STRUCTS = {
    "probeArgsGPU": abi.ProbeArgsGPU,
    "buildArgsGPU": abi.BuildArgsGPU,
    "groupbyArgsGPU": abi.GroupbyArgsGPU,
    "filterArgsGPU": abi.FilterArgsGPU,
    "offsetGPU": abi.OffsetGPU,
}


def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for s, cls in STRUCTS.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{s} {fname} %zu\\n", offsetof({s}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for s, cls in STRUCTS.items():
        assert got[(s, "size")] == C.sizeof(cls), s
        for fname, _ in cls._fields_:
            assert got[(s, fname)] == getattr(cls, fname).offset, (s, fname)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(abi, "_lib", None)
    monkeypatch.setattr(abi, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        abi.load()


def test_library_has_no_unresolved_symbols():
    """Every symbol the library needs resolves (dlopen with RTLD_NOW)."""
    C.CDLL(abi.LIB_PATH, mode=os.RTLD_NOW | os.RTLD_LOCAL)
    out = subprocess.check_output(["nm", "-D", "--undefined-only", abi.LIB_PATH], text=True)
    missing = [ln.split()[-1] for ln in out.splitlines()
               if ln.split() and ln.split()[0] == "U" and "@" not in ln.split()[-1]]
    assert not missing, missing


def test_invalid_arguments_are_rejected_before_launch():
    """Argument checks return hipErrorInvalidValue (1) without touching the
    device, instead of the reference's device-side assert-and-return."""
    lib = abi.load()
    g = abi.GroupbyArgsGPU()
    g.total_val = 0
    assert lib.mordred_probe_group_by_GPU2(None, None, C.pointer(g), 10, None, 0, None, None) == 1
    assert lib.mordred_filter_probe_GPU3(None, None, None, None, None, 5, None, None) == 1
    # NULL segment group needs a segment-aligned start_offset
    g.total_val = 7
    assert lib.mordred_probe_group_by_GPU2(C.c_void_p(16), None, C.pointer(g), 10, C.c_void_p(16), 5, None,
                                           None) == 1
    b = abi.BuildArgsGPU(None, None, 0, 0, 0)
    assert lib.mordred_build_GPU2(C.c_void_p(16), None, C.pointer(b), 10, C.c_void_p(16), 0, None, None) == 1
    # build_GPU3 refuses a dimension filter (GPUProcessing.h:1017)
    b = abi.BuildArgsGPU(C.c_void_p(16), None, 8, 0, 0)
    f = abi.FilterArgsGPU()
    f.filter_idx1 = 16
    assert lib.mordred_build_GPU3(C.c_void_p(16), C.c_void_p(16), C.pointer(f), C.pointer(b), 10,
                                  C.c_void_p(16), None) == 1
    # an empty input is a successful no-op
    assert lib.mordred_probe_group_by_GPU2(None, None, C.pointer(g), 0, C.c_void_p(16), 0, None, None) == 0


def test_engine_and_cost_model_bindings_cover_headers():
    from mordred_amd import costmodel, engine
    assert set(declared_functions(ENGINE_HEADER)) == set(engine._SIGS)
    assert set(declared_functions(COST_HEADER)) == set(costmodel._SIGS)
