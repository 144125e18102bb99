"""C-ABI boundary: struct layouts, header self-check, exported symbols.

Sizes are the reference's own layout assertions (src/ffi.rs:494-507,
src/metal_bridge.mm:32-39).  No GPU calls.
"""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

from fractalpathtracer_amd import ffi

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("struct,size", list(ffi.EXPECTED_SIZES.items()), ids=lambda v: getattr(v, "__name__", str(v)))
def test_ctypes_mirror_sizes(struct, size):
    assert C.sizeof(struct) == size


def test_config_field_offsets_match_c_header(tmp_path):
    """Compile a C program that prints offsetof() for every config field and
    compare with the ctypes mirror (catches a reordered or retyped field)."""
    fields = [f for f, _ in ffi.FptRenderConfig._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fpt_abi.h"', "int main(void){"]
    for f in fields:
        lines.append(f'printf("%zu\\n", offsetof(struct FptRenderConfig, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "off.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    c_offsets = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    py_offsets = [getattr(ffi.FptRenderConfig, f).offset for f in fields]
    assert c_offsets == py_offsets


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "fpt_abi.h"\nint main(void){return 0;}\n')
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o",
                    str(tmp_path / "a")], check=True)
    subprocess.run(["g++", "-std=c++17", "-x", "c++", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src),
                    "-o", str(tmp_path / "b")], check=True)


def _declared_functions():
    import re
    text = (ROOT / "include" / "fpt_abi.h").read_text()
    return sorted(set(re.findall(r"^(?:int|void|const char \*)\s*\**\s*(fpt_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ffi.load_library()
    declared = _declared_functions()
    assert "fpt_metal_render" in declared and "fpt_hip_trace_primary" in declared
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(ffi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(declared) <= exported


def test_reference_entry_points_keep_their_names():
    # metal_bridge.h:404-486
    for name in ("fpt_metal_render", "fpt_metal_diagnostic_render", "fpt_mandelbulber_sample_field",
                 "fpt_compare_images", "fpt_metal_device_name", "fpt_test_voxel_dda",
                 "fpt_test_async_stitch_context", "fpt_test_typed_soa", "fpt_metal_preview"):
        assert name in _declared_functions()


def test_invalid_configs_fail_with_message_before_touching_the_gpu():
    """metal_bridge.mm:4319-4328: missing config / zero size are errors."""
    lib = ffi.load_library()
    err = ffi.error_buffer()
    rc = lib.fpt_metal_render(None, None, None, None, None, 0, None, None, None, None, None, None, None,
                              None, None, None, None, None, None, 0, err, len(err))
    assert rc == 1 and b"missing render config" in err.value
    cfg = ffi.FptRenderConfig()
    rc = lib.fpt_metal_render(None, None, None, None, None, 0, C.byref(cfg), None, None, None, None, None,
                              None, None, None, None, None, None, None, 0, err, len(err))
    assert rc == 1 and b"invalid render size 0x0" in err.value


def test_linear_output_too_small_is_rejected():
    from fractalpathtracer_amd import scenes
    lib = ffi.load_library()
    cfg = scenes.mandelbulb_scene(8, 8, 1)
    err = ffi.error_buffer()
    buf = (C.c_float * 10)()
    rc = lib.fpt_metal_render(None, None, None, None, None, 0, C.byref(cfg), None, None, None, None, None,
                              None, None, None, None, None, None, buf, 10, err, len(err))
    assert rc == 1 and b"linear output buffer is too small (10 < 256)" in err.value
