"""ctypes binding of ``libalthea_amd.so`` (the C ABI of include/althea/althea.h).

Loading fails loudly when the library has not been built: there is no CPU or
Python fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libalthea_amd.so"

ALTHEA_OK = 0
ALTHEA_ERROR_INVALID_ARGUMENT = 1
PRECOMPUTE_FP32_SUMS = 0x1
RESOLVE_TONEMAP = 0x1
RESOLVE_OUT_FP16 = 0x2
GBUFFER_FP32 = 0
GBUFFER_FP16 = 1
SSAO_MAX_SAMPLES = 64

vp = C.c_void_p
i32 = C.c_int
u32 = C.c_uint32
sz = C.c_size_t


class IBLDescC(C.Structure):
    _fields_ = [("irradiance_size", i32), ("irradiance_theta_samples", i32), ("prefiltered_size", i32),
                ("prefiltered_mips", i32), ("prefilter_samples", i32), ("brdf_lut_size", i32),
                ("brdf_lut_samples", i32), ("flags", C.c_uint)]


class IBLViewsC(C.Structure):
    _fields_ = [("environment", vp), ("env_width", i32), ("env_height", i32), ("env_levels", i32),
                ("prefiltered", vp), ("prefiltered_size", i32), ("prefiltered_mips", i32),
                ("irradiance", vp), ("irradiance_size", i32), ("brdf_lut", vp), ("brdf_lut_size", i32),
                ("prefiltered_bordered", vp), ("irradiance_bordered", vp)]


class GBufferViewC(C.Structure):
    _fields_ = [("position", vp), ("normal", vp), ("albedo", vp), ("mro", vp), ("width", i32), ("height", i32),
                ("format", i32)]


class GlobalsC(C.Structure):
    _fields_ = [("projection", C.c_float * 16), ("inverse_projection", C.c_float * 16), ("view", C.c_float * 16),
                ("prev_view", C.c_float * 16), ("inverse_view", C.c_float * 16),
                ("prev_inverse_view", C.c_float * 16), ("mouse_uv", C.c_float * 2), ("light_count", C.c_int32),
                ("light_buffer_handle", u32), ("time", C.c_float), ("exposure", C.c_float), ("input_mask", u32),
                ("frame_count", u32)]


class SsaoDescC(C.Structure):
    _fields_ = [("enabled", i32), ("samples", i32), ("radius", C.c_float), ("bias", C.c_float),
                ("kernel", (C.c_float * 4) * SSAO_MAX_SAMPLES), ("noise", (C.c_float * 4) * 16)]


class MlpDescC(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("init_output_bias", C.c_float)]


class EnvViewC(C.Structure):
    _fields_ = [("rgba", vp), ("width", i32), ("height", i32)]


class TuningC(C.Structure):
    _fields_ = [("tile_w", i32), ("tile_h", i32), ("waves_per_cu", i32), ("reserved", i32)]


class ReflectionViewC(C.Structure):
    _fields_ = [("data", vp), ("width", i32), ("height", i32), ("mips", i32), ("format", i32)]


class GBufferMeshC(C.Structure):
    _fields_ = [("positions", vp), ("normals", vp), ("colors", vp), ("mro", vp), ("indices", vp),
                ("vertex_count", i32), ("triangle_count", i32)]


class TextureViewC(C.Structure):
    _fields_ = [("texels", vp), ("width", i32), ("height", i32), ("mips", i32), ("srgb", i32), ("mag_filter", i32),
                ("min_filter", i32), ("mipmap_mode", i32), ("wrap_u", i32), ("wrap_v", i32)]


class GltfMaterialC(C.Structure):
    _fields_ = [("base_color_factor", C.c_float * 4), ("metallic_factor", C.c_float), ("roughness_factor", C.c_float),
                ("normal_scale", C.c_float), ("alpha_cutoff", C.c_float), ("base_texture", i32),
                ("normal_texture", i32), ("metallic_roughness_texture", i32), ("base_uv", i32),
                ("metallic_roughness_uv", i32)]


class GltfPrimitiveC(C.Structure):
    _fields_ = [("positions", vp), ("tangents", vp), ("bitangents", vp), ("normals", vp), ("uvs", vp * 4),
                ("indices", vp), ("vertex_count", i32), ("triangle_count", i32), ("material", i32)]


WRAP_REPEAT, WRAP_MIRRORED_REPEAT, WRAP_CLAMP_TO_EDGE = 0, 1, 2  # ALTHEA_WRAP_*
FILTER_NEAREST, FILTER_LINEAR = 0, 1                               # ALTHEA_FILTER_*


class MeshViewC(C.Structure):
    _fields_ = [("positions", vp), ("indices", vp), ("vertex_count", i32), ("triangle_count", i32)]


SHADOW_NEAR = 0.01       # ALTHEA_SHADOW_NEAR
SHADOW_FAR = 1000.0      # ALTHEA_SHADOW_FAR
REFLECTION_MAX_MIPS = 8  # ALTHEA_REFLECTION_MAX_MIPS
STATUS_NOT_FOUND = 4     # ALTHEA_ERROR_NOT_FOUND
REFLECTION_FP16 = 0      # ALTHEA_REFLECTION_FP16
REFLECTION_FP32 = 1      # ALTHEA_REFLECTION_FP32


assert C.sizeof(GlobalsC) == 416, C.sizeof(GlobalsC)

# name -> (restype, argtypes)
MONITOR_CELLS = 64      # ALTHEA_MLP_MONITOR_CELLS
MONITOR_FLOOR = 1e-2    # ALTHEA_MLP_MONITOR_FLOOR

PROTOTYPES = {
    "althea_status_string": (C.c_char_p, [i32]),
    "althea_last_error": (C.c_char_p, []),
    "althea_version": (i32, []),
    "althea_hammersley": (i32, [u32, u32, u32, vp, vp, vp]),
    "althea_env_levels": (i32, [i32, i32]),
    "althea_env_pyramid_texels": (sz, [i32, i32]),
    "althea_env_pyramid_build": (i32, [vp, i32, i32, vp, vp]),
    "althea_cube_texels": (sz, [i32, i32]),
    "althea_cube_bordered_texels": (sz, [i32, i32]),
    "althea_cube_border": (i32, [vp, i32, i32, vp, vp]),
    "althea_cube_sample": (i32, [vp, i32, i32, vp, i32, vp, vp]),
    "althea_env_sample": (i32, [vp, i32, i32, vp, i32, vp, vp]),
    "althea_prefilter_cube": (i32, [vp, i32, i32, i32, i32, i32, C.c_uint, vp, vp]),
    "althea_prefilter_cube_scratch_bytes": (sz, [i32, i32]),
    "althea_prefilter_cube_noalloc": (i32, [vp, i32, i32, i32, i32, i32, C.c_uint, vp, vp]),
    "althea_prefilter_points": (i32, [vp, i32, i32, i32, vp, i32, vp, vp]),
    "althea_irradiance_cube": (i32, [vp, i32, i32, i32, i32, C.c_uint, vp, vp]),
    "althea_brdf_lut": (i32, [i32, i32, vp, vp]),
    "althea_ibl_desc_default": (None, [C.POINTER(IBLDescC)]),
    "althea_ibl_create": (i32, [C.POINTER(IBLDescC), C.POINTER(vp)]),
    "althea_ibl_destroy": (i32, [vp]),
    "althea_ibl_precompute": (i32, [vp, vp, i32, i32, vp]),
    "althea_ibl_get_views": (i32, [vp, C.POINTER(IBLViewsC)]),
    "althea_ibl_precompute_batched": (i32, [C.POINTER(vp), C.POINTER(EnvViewC), i32, C.POINTER(TuningC),
                                            C.POINTER(TuningC), vp]),
    "althea_ibl_state_bytes": (sz, [vp]),
    "althea_ibl_save_cache": (i32, [vp, C.c_char_p, vp]),
    "althea_ibl_load_cache": (i32, [vp, vp, i32, i32, C.c_char_p, vp]),
    "althea_deferred_resolve": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, i32, C.POINTER(GlobalsC),
                                      C.POINTER(SsaoDescC), C.c_uint, vp, vp, vp]),
    "althea_reflection_texels": (sz, [i32, i32, i32]),
    "althea_reflection_mip_offset": (sz, [i32, i32, i32]),
    "althea_ssr_workspace_bytes": (sz, [i32, i32]),
    "althea_ssr_capture": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, i32, C.POINTER(GlobalsC),
                                 C.POINTER(ReflectionViewC), vp, vp]),
    "althea_ssr_capture_shadowed": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, i32,
                                          C.POINTER(GlobalsC), C.POINTER(ReflectionViewC), vp, i32, vp, vp]),
    "althea_reflection_convolve": (i32, [C.POINTER(ReflectionViewC), vp]),
    "althea_deferred_resolve_reflect": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, i32,
                                              C.POINTER(GlobalsC), C.POINTER(SsaoDescC), C.POINTER(ReflectionViewC),
                                              C.c_uint, vp, vp, vp]),
    "althea_gbuffer_workspace_bytes": (sz, [i32, i32, i32]),
    "althea_gbuffer_draw": (i32, [C.POINTER(GBufferMeshC), C.POINTER(GlobalsC), C.POINTER(GBufferViewC), vp, vp]),
    "althea_texture_texels": (sz, [i32, i32, i32]),
    "althea_texture_mip_offset": (sz, [i32, i32, i32]),
    "althea_texture_build_mips": (i32, [vp, i32, i32, i32, i32, vp]),
    "althea_gbuffer_gltf_workspace_bytes": (sz, [i32, i32, C.POINTER(GltfPrimitiveC), i32, i32, i32]),
    "althea_gbuffer_draw_gltf": (i32, [C.POINTER(GltfPrimitiveC), i32, C.POINTER(GltfMaterialC), i32,
                                       C.POINTER(TextureViewC), i32, C.POINTER(GlobalsC), C.POINTER(GBufferViewC), vp,
                                       vp]),
    "althea_shadow_cube_floats": (sz, [i32]),
    "althea_shadow_cubes_workspace_bytes": (sz, [i32, i32]),
    "althea_shadow_cubes_ws": (i32, [C.POINTER(MeshViewC), vp, i32, i32, vp, vp, vp]),
    "althea_shadow_cubes": (i32, [C.POINTER(MeshViewC), vp, i32, i32, vp, vp]),
    "althea_shadow_cubes_gltf_workspace_bytes": (sz, [C.POINTER(GltfPrimitiveC), i32, i32]),
    "althea_shadow_cubes_gltf": (i32, [C.POINTER(GltfPrimitiveC), i32, C.POINTER(GltfMaterialC), i32,
                                       C.POINTER(TextureViewC), i32, vp, i32, i32, vp, vp, vp]),
    "althea_deferred_resolve_shadowed": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, i32,
                                               C.POINTER(GlobalsC), C.POINTER(SsaoDescC), C.POINTER(ReflectionViewC),
                                               vp, i32, C.c_uint, vp, vp, vp]),
    "althea_hdr_info": (i32, [C.c_char_p, C.POINTER(i32), C.POINTER(i32)]),
    "althea_hdr_read": (i32, [C.c_char_p, vp, sz]),
    "althea_hdr_write": (i32, [C.c_char_p, vp, i32, i32]),
    "althea_mlp_desc_default": (None, [C.POINTER(MlpDescC)]),
    "althea_mlp_param_count": (i32, []),
    "althea_mlp_create": (i32, [C.POINTER(MlpDescC), C.POINTER(vp)]),
    "althea_mlp_destroy": (i32, [vp]),
    "althea_mlp_set_params": (i32, [vp, vp, vp]),
    "althea_mlp_get_params": (i32, [vp, vp, vp]),
    "althea_mlp_train_step": (i32, [vp, vp, vp, i32, vp, vp]),
    "althea_mlp_gradients": (i32, [vp, vp, vp, i32, vp, vp, vp]),
    "althea_mlp_infer": (i32, [vp, vp, i32, vp, vp]),
    "althea_mlp_state_bytes": (sz, [vp]),
    "althea_mlp_set_learning_rate": (i32, [vp, C.c_float]),
    "althea_mlp_build_monitor": (i32, [vp, vp, i32, i32, C.c_float, vp, C.POINTER(i32), vp]),
    "althea_deferred_resolve_mlp": (i32, [C.POINTER(GBufferViewC), C.POINTER(IBLViewsC), vp, vp, i32,
                                          C.POINTER(GlobalsC), C.POINTER(SsaoDescC), C.c_uint, vp, vp, vp, vp]),
    "althea_mfma_bf16_probe": (i32, [vp, vp, vp, vp]),
    "althea_prefilter_batch_create": (i32, [C.POINTER(EnvViewC), i32, C.POINTER(i32), i32, i32, C.c_uint,
                                            C.POINTER(vp)]),
    "althea_prefilter_batch_destroy": (i32, [vp]),
    "althea_prefilter_batch_run": (i32, [vp, vp]),
    "althea_prefilter_batch_autotune": (i32, [vp, C.POINTER(TuningC), vp]),
    "althea_prefilter_batch_set_tuning": (i32, [vp, C.POINTER(TuningC)]),
    "althea_prefilter_batch_get_tuning": (i32, [vp, C.POINTER(TuningC)]),
    "althea_prefilter_batch_cube": (i32, [vp, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(i32)]),
    "althea_prefilter_batch_texels": (sz, [vp]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded C ABI library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                               "(there is no fallback path)")
        lib_ = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib_, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib_
    return _lib


class AltheaError(RuntimeError):
    def __init__(self, status: int, where: str):
        l = lib()
        super().__init__(f"{where}: {l.althea_status_string(status).decode()} ({status}): "
                         f"{l.althea_last_error().decode()}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != ALTHEA_OK:
        raise AltheaError(status, where)
