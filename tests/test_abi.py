"""The C-ABI libraries load and export every symbol their headers declare;
host-only behaviour that needs no GPU."""

import ctypes
import os
import re

import pytest

import dayflower_amd as dfa
from dayflower_amd import _ffi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[\w ]+?\**\s*\b(dfs?_\w+)\s*\(", text, flags=re.M))


@pytest.mark.parametrize("header,path", [
    ("dayflower_hip.h", _ffi.HIP_LIB_PATH),
    ("dayflower_scene.h", _ffi.SCENE_LIB_PATH),
])
def test_library_exports_every_declared_symbol(header, path):
    names = _declared(header)
    assert len(names) > 10
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_headers():
    assert _declared("dayflower_hip.h") == set(_ffi.HIP_PROTOTYPES)
    assert _declared("dayflower_scene.h") == set(_ffi.SCENE_PROTOTYPES)


def test_create_without_device_fails_loudly():
    if dfa.device_count() > 0:
        pytest.skip("a HIP device is visible")
    ctx = ctypes.c_void_p()
    assert _ffi.hip_lib().df_create(ctypes.byref(ctx), 0) == _ffi.DF_ENODEVICE
    with pytest.raises(dfa.HIPError):
        dfa.HIPRendererKernel(dfa.build_config(1))


def test_null_arguments_are_rejected():
    lib = _ffi.hip_lib()
    assert lib.df_load_scene(None, None) == _ffi.DF_EINVAL
    assert lib.df_execute(None, 1) == _ffi.DF_EINVAL
    assert lib.df_last_error(None) == b"null handle"
    assert _ffi.scene_lib().dfs_build_config(9, ctypes.c_uint64(0), ctypes.byref(ctypes.c_void_p())) == _ffi.DF_EINVAL


def test_pixel_seeds_are_splitmix64():
    # SplitMix64 reference output: state 0 advanced once by the golden gamma
    s = dfa.pixel_seeds(0, 2)
    assert int(s[0]) & ((1 << 64) - 1) == 0xE220A8397B1DCDAF


def test_batch_create_without_device_fails_loudly():
    if dfa.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(dfa.HIPError) as e:
        dfa.HIPBatchRenderer()
    assert e.value.code == -2


def test_struct_layouts_match_the_c_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the sizes (and, for df_schedule, the field offsets) the C
    compiler gives the header's declarations."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    pairs = [("df_schedule", _ffi.Schedule), ("df_sky", _ffi.Sky), ("df_params", _ffi.Params),
             ("df_diff", _ffi.Diff), ("df_counters", _ffi.Counters), ("df_compiled_scene", _ffi.CompiledScene)]
    sched_fields = [f[0] for f in _ffi.Schedule._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "dayflower_types.h"', "int main(void) {"]
    for c_name, _ in pairs:
        src.append('  printf("%%zu\\n", sizeof(%s));' % c_name)
    for f in sched_fields:
        src.append('  printf("%%zu\\n", offsetof(df_schedule, %s));' % f)
    src += ["  return 0;", "}"]
    (tmp_path / "layout.c").write_text("\n".join(src) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.run([cc, "-I", inc, "-o", exe, str(tmp_path / "layout.c")], check=True, capture_output=True)
    out = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    for (c_name, mirror), size in zip(pairs, out):
        assert ctypes.sizeof(mirror) == size, c_name
    for f, off in zip(sched_fields, out[len(pairs):]):
        assert getattr(_ffi.Schedule, f).offset == off, f
