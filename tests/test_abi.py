"""The C-ABI library loads on the CPU and exports every function include/hss.h declares
(no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hss.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hss_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["hss_abi_version", "hss_grid_semblance", "hss_build_parameter_array"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from highperformanceseismicstacking_amd import capi
    lib = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_status_strings_and_version_without_gpu():
    from highperformanceseismicstacking_amd import capi
    lib = capi.lib()
    assert lib.hss_abi_version() >= 1
    for code in [0, -1, -2, -3, -4, -5, -6, -7, -8]:
        assert lib.hss_status_string(code)


def test_argument_errors_are_reported_without_touching_the_device():
    """Validation happens on the host before any launch: NULL or malformed arguments
    return negative statuses."""
    from highperformanceseismicstacking_amd import capi
    lib = capi.lib()
    tr = capi.hss_traces(0, 0, 0, 0, 0, 10, 100, 0.004, 2, 5, 0)
    mp = capi.hss_midpoints(1, 0, 0, 0, 0, 0.0, 0.0)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, None, 16, 1, None, None,
                                None, None, None)
    assert rc == -4  # NULL samples
    mp = capi.hss_midpoints(1, 8, 8, 8, 0, 0.0, 0.0)  # never dereferenced: validation fails first
    tr = capi.hss_traces(8, 0, 8, 8, 8, 10, 100, 0.004, 2, 7, 0)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, ctypes.c_void_p(8), 16, 1,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -2  # w != 2*tau+1
    tr = capi.hss_traces(8, 0, 8, 8, 8, 10, 100, 0.004, 2, 5, 0)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 7, ctypes.c_void_p(8), 16, 1,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -1  # bad model
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, ctypes.c_void_p(8), 16, 9,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -6  # bad variant
    tr = capi.hss_traces(10, 0, 8, 8, 8, 10, 100, 0.004, 2, 5, 1)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, ctypes.c_void_p(8), 16, 1,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -5  # 16-bit samples not on a 4-byte boundary
    tr = capi.hss_traces(8, 0, 8, 8, 8, 10, (1 << 24) + 1, 0.004, 2, 5, 0)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, ctypes.c_void_p(8), 16, 1,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -3  # more samples per trace than the exact window bounds allow
    tr = capi.hss_traces(8, 0, 8, 8, 8, 10, 100, 0.004, 16, 33, 0)
    rc = lib.hss_grid_semblance(ctypes.byref(tr), ctypes.byref(mp), 0, ctypes.c_void_p(8), 16, 1,
                                ctypes.c_void_p(8), None, None, None, None)
    assert rc == -2  # w above MAX_WINDOW_SIZE - 1


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """The Python binding's structures have the C header's sizes and field offsets (a C
    program built against include/hss.h prints them)."""
    import shutil
    import subprocess
    from highperformanceseismicstacking_amd import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"hss_traces": capi.hss_traces, "hss_midpoints": capi.hss_midpoints,
               "hss_oct_candidates": capi.hss_oct_candidates}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "hss.h"', 'int main(void) {']
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, key, val = line.split()
        got[(name, key)] = int(val)
    for name, cls in structs.items():
        assert ctypes.sizeof(cls) == got[(name, "size")], name
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == got[(name, field)], (name, field)
