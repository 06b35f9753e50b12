"""The C-ABI library loads, exports every symbol the header declares, and its argument
struct matches the Python mirror (no device calls)."""
import os
import re

from gqis import runtime

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "gqis_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gq_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = runtime.load_library()
    names = declared_functions()
    assert "gq_launch_sweep" in names and "gq_module_compile" in names
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(runtime.exported_symbols()) == names


def test_argument_struct_layout_matches():
    import ctypes
    lib = runtime.load_library()
    assert lib.gq_sweep_args_size() == ctypes.sizeof(runtime.SweepArgs)


def test_compile_error_is_reported():
    import pytest
    with pytest.raises(runtime.GqError, match="hiprtc"):
        runtime.compile_check("this is not HIP", ["-O3"])
