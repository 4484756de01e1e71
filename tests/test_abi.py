"""The C-ABI library loads and exports every function include/tt.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 30
    lib = ctypes.CDLL(os.path.join(ROOT, "trenchanttracer_amd", "libtt.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header(tt):
    from trenchanttracer_amd import _native
    missing = [n for n in declared_functions() if n not in _native.PROTOS]
    assert not missing, missing


def test_struct_sizes(tt):
    from trenchanttracer_amd import _native as N
    assert ctypes.sizeof(N.tt_ray) == 32
    assert ctypes.sizeof(N.tt_hit) == 16
    assert ctypes.sizeof(N.tt_material) == 48
    # the same sizes the library asserts at compile time (api_host.cpp)
    assert ctypes.sizeof(N.tt_geometry) == 48
    assert ctypes.sizeof(N.tt_trace_cfg) == 24
    assert ctypes.sizeof(N.tt_camera) == 60
    assert ctypes.sizeof(N.tt_render_params) == 76
    assert ctypes.sizeof(N.tt_render_stats) == 40
    assert ctypes.sizeof(N.tt_guide_params) == 36
    assert ctypes.sizeof(N.tt_guide_stats) == 32


def test_abi_version_matches_binding(tt):
    from trenchanttracer_amd import _native as N
    assert N.lib.tt_abi_version() == N.ABI_VERSION


def test_no_gpu_entry_points_fail_cleanly_without_device(tt):
    # Host-side objects work without a GPU; on a machine without one, creating
    # a context returns TT_ERR_UNSUPPORTED instead of crashing.
    import pytest
    m = tt.Mesh.soup(10, seed=1)
    tt.BVH.build(m)
    try:
        c = tt.Context(0)
    except tt.TTError as e:
        assert e.status in (3, 5)
    else:
        c.close()
        pytest.skip("a GPU is present")
