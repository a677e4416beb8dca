"""The C ABI library loads and exports every entry point include/althea/althea.h
declares; the C++ facade header compiles; the C++ consumer runs (GPU)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "althea" / "althea.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"ALTHEA_API\s+[\w\s\*]+?\b(althea_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 40, names
    for must in ("althea_ibl_create", "althea_ibl_precompute", "althea_deferred_resolve", "althea_prefilter_cube",
                 "althea_mlp_train_step", "althea_prefilter_batch_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(ROOT / "althea_amd" / "libalthea_amd.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_hides_internal_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "althea_amd" / "libalthea_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    stray = [s for s in exported if not s.startswith("althea_") and not s.startswith("_")]
    assert not stray, stray[:20]


def test_host_queries_without_gpu():
    lib = C.CDLL(str(ROOT / "althea_amd" / "libalthea_amd.so"))
    assert lib.althea_env_levels(512, 256) == 10
    assert lib.althea_env_levels(4096, 2048) == 13
    lib.althea_env_pyramid_texels.restype = C.c_size_t
    lib.althea_cube_texels.restype = C.c_size_t
    assert lib.althea_env_pyramid_texels(4, 2) == 4 * 2 + 2 * 1 + 1
    assert lib.althea_cube_texels(128, 5) == 6 * sum((128 >> m) ** 2 for m in range(5))
    lib.althea_status_string.restype = C.c_char_p
    assert lib.althea_status_string(1) == b"invalid argument"


def test_reflection_queries_and_argument_checks_without_gpu():
    """Reflection-buffer sizes, and the argument checks that run before any
    device work (dummy device pointers are never dereferenced)."""
    import ctypes as C

    from althea_amd import _lib

    lib = _lib.lib()
    assert lib.althea_reflection_texels(40, 24, 3) == 40 * 24 + 20 * 12 + 10 * 6
    assert lib.althea_reflection_texels(37, 5, 5) == 37 * 5 + 18 * 2 + 9 * 1 + 4 * 1 + 2 * 1
    assert lib.althea_reflection_mip_offset(40, 24, 2) == 40 * 24 + 20 * 12
    assert lib.althea_reflection_texels(40, 24, 0) == 0
    dummy = 0x1000
    g = _lib.GBufferViewC(dummy, dummy, dummy, dummy, 64, 32, 0)
    v = _lib.IBLViewsC()
    v.environment, v.env_width, v.env_height = dummy, 64, 32
    v.prefiltered, v.prefiltered_size, v.prefiltered_mips = dummy, 16, 3
    v.irradiance, v.irradiance_size, v.brdf_lut, v.brdf_lut_size = dummy, 8, dummy, 16
    glob = _lib.GlobalsC()
    bad_extent = _lib.ReflectionViewC(dummy, 64, 31, 5, 0)
    # Depth plane in 8 x 4 texel blocks.
    assert lib.althea_ssr_workspace_bytes(40, 24) == 4 * 40 * 24
    assert lib.althea_ssr_workspace_bytes(37, 5) == 4 * (5 * 8) * (2 * 4)
    assert lib.althea_ssr_workspace_bytes(0, 24) == 0
    assert lib.althea_ssr_capture(C.byref(g), C.byref(v), None, 0, C.byref(glob), C.byref(bad_extent), dummy,
                                  None) == 1
    assert b"reflection buffer 64x31" in lib.althea_last_error()
    bad_mips = _lib.ReflectionViewC(dummy, 64, 32, 9, 0)
    assert lib.althea_reflection_convolve(C.byref(bad_mips), None) == 1
    bad_fmt = _lib.ReflectionViewC(dummy, 64, 32, 5, 7)
    assert lib.althea_deferred_resolve_reflect(C.byref(g), C.byref(v), None, 0, C.byref(glob), None,
                                               C.byref(bad_fmt), 0, dummy, None, None) == 1
    assert lib.althea_reflection_convolve(None, None) == 1


def test_shadow_argument_checks_without_gpu():
    """althea_shadow_cubes / althea_deferred_resolve_shadowed reject bad
    arguments before any device work (dummy pointers are never read)."""
    from althea_amd import _lib

    lib = _lib.lib()
    assert lib.althea_shadow_cube_floats(256) == 6 * 258 * 258
    assert lib.althea_shadow_cube_floats(0) == 0
    dummy = 0x1000
    mesh = _lib.MeshViewC(dummy, dummy, 3, 1)
    assert lib.althea_shadow_cubes(C.byref(mesh), dummy, 1, 0, dummy, None) == 1
    assert b"size 0" in lib.althea_last_error()
    assert lib.althea_shadow_cubes(C.byref(mesh), dummy, 1, 4097, dummy, None) == 1
    assert lib.althea_shadow_cubes(C.byref(mesh), dummy, 0, 16, dummy, None) == 1
    assert lib.althea_shadow_cubes(C.byref(mesh), None, 1, 16, dummy, None) == 1
    assert lib.althea_shadow_cubes(None, dummy, 1, 16, dummy, None) == 1
    no_vertices = _lib.MeshViewC(dummy, dummy, 0, 1)
    assert lib.althea_shadow_cubes(C.byref(no_vertices), dummy, 1, 16, dummy, None) == 1
    huge = _lib.MeshViewC(dummy, dummy, 3, 1 << 30)
    assert lib.althea_shadow_cubes(C.byref(huge), dummy, 2, 16, dummy, None) == 1
    # The workspace variant: 8 B offset + 48 B boxes per light-triangle pair,
    # 8 B per 256 pairs, each part 256-byte aligned.
    assert lib.althea_shadow_cubes_workspace_bytes(2, 1000) == 16128 + 96000 + 256
    assert lib.althea_shadow_cubes_workspace_bytes(0, 10) == 0
    assert lib.althea_shadow_cubes_workspace_bytes(2, 1 << 30) == 0
    assert lib.althea_shadow_cubes_ws(C.byref(mesh), dummy, 1, 16, dummy, None, None) == 1
    assert b"null" in lib.althea_last_error()
    assert lib.althea_shadow_cubes_ws(C.byref(mesh), dummy, 1, 0, dummy, dummy, None) == 1
    # Workspaces must be 256-byte aligned (16-byte stores, 64-bit atomics).
    assert lib.althea_shadow_cubes_ws(C.byref(mesh), dummy, 1, 16, dummy, dummy + 8, None) == 1
    assert b"256-byte aligned" in lib.althea_last_error()
    assert lib.althea_shadow_cubes_ws(C.byref(huge), dummy, 2, 16, dummy, dummy, None) == 1
    g = _lib.GBufferViewC(dummy, dummy, dummy, dummy, 64, 32, 0)
    v = _lib.IBLViewsC()
    v.environment, v.env_width, v.env_height = dummy, 64, 32
    v.prefiltered, v.prefiltered_size, v.prefiltered_mips = dummy, 16, 3
    v.irradiance, v.irradiance_size, v.brdf_lut, v.brdf_lut_size = dummy, 8, dummy, 16
    glob = _lib.GlobalsC()
    assert lib.althea_deferred_resolve_shadowed(C.byref(g), C.byref(v), dummy, 1, C.byref(glob), None, None, None,
                                                256, 0, dummy, None, None) == 1
    assert b"shadow cubes" in lib.althea_last_error()
    assert lib.althea_deferred_resolve_shadowed(C.byref(g), C.byref(v), dummy, 1, C.byref(glob), None, None, dummy,
                                                0, 0, dummy, None, None) == 1


def test_gltf_and_texture_argument_checks_without_gpu():
    """Texture and glTF-subpass arguments are checked before any device work."""
    import ctypes as C

    from althea_amd import _lib

    lib = _lib.lib()
    assert lib.althea_texture_texels(4, 2, 3) == 4 * 2 + 2 * 1 + 1 * 1
    assert lib.althea_texture_mip_offset(64, 32, 2) == 64 * 32 + 32 * 16
    assert lib.althea_texture_build_mips(0x1000, 4, 4, 4, 0, None) == 1  # 4x4 has 3 levels
    assert b"levels" in lib.althea_last_error()
    dummy = 0x1000
    tex = (_lib.TextureViewC * 1)(_lib.TextureViewC(dummy, 4, 4, 3, 0, 1, 1, 1, 0, 0))
    mat = (_lib.GltfMaterialC * 1)(_lib.GltfMaterialC((C.c_float * 4)(1, 1, 1, 1), 1, 1, 1, 0.5, 0, 0, 1, 0, 0))
    uv = (C.c_void_p * 4)(dummy, None, None, None)
    prim = (_lib.GltfPrimitiveC * 1)(_lib.GltfPrimitiveC(dummy, dummy, dummy, dummy, uv, dummy, 3, 1, 0))
    g = _lib.GBufferViewC(dummy, dummy, dummy, dummy, 64, 32, 0)
    glob = _lib.GlobalsC()
    glob.projection[10], glob.projection[14] = -1.00001, -0.01
    assert lib.althea_gbuffer_draw_gltf(prim, 1, mat, 1, tex, 1, C.byref(glob), C.byref(g), dummy, None) == 1
    assert b"names a texture outside" in lib.althea_last_error()
    bad_sampler = (_lib.TextureViewC * 2)(tex[0], _lib.TextureViewC(dummy, 4, 4, 1, 0, 7, 1, 1, 0, 0))
    assert lib.althea_gbuffer_draw_gltf(prim, 1, mat, 1, bad_sampler, 2, C.byref(glob), C.byref(g), dummy,
                                        None) == 1
    assert b"sampler state" in lib.althea_last_error()
    assert lib.althea_gbuffer_gltf_workspace_bytes(64, 32, prim, 1, 1, 2) > 64 * 32 * 8


def test_facade_header_compiles():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "facade_demo.cpp")], check=True)


@pytest.mark.gpu
def test_facade_demo_runs():
    exe = ROOT / "tests" / "cpp" / "build" / "facade_demo"
    assert exe.exists(), "built by __graft_entry__.build()"
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "facade ok" in proc.stdout
