"""Python binding of the MI355X IBL precompute + deferred-resolve library.

The surface mirrors nithinp7/Althea's host API for this path
(Include/Althea/ImageBasedLighting.h:17-54, Include/Althea/DeferredRendering.h:26-102)
with torch device tensors in place of VkImages:

- ``ImageBasedLighting.createResources(env, desc)`` -> ``IBLResources``
  (environment pyramid, prefiltered cube, irradiance cube, BRDF LUT);
  ``IBLResources.getHandles()`` returns the device views in the reference's
  order (environment, prefiltered, irradiance, BRDF LUT).
- ``GBufferResources`` wraps the four G-buffer planes in the order of
  GBufferResources::buildMaterial (Src/DeferredRendering.cpp:23-35).
- ``DeferredRendering.resolve(...)`` runs DeferredPass.frag's lighting.

Every call goes through the C ABI of ``libalthea_amd.so`` onto torch's
current HIP stream.  There is no fallback: without the built library or a
GPU these calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import synth
from ._lib import AltheaError, check, lib

__all__ = ["AltheaError", "IBLDesc", "IBLResources", "ImageBasedLighting", "GBufferResources",
           "DeferredRendering", "ReflectionBuffer", "ScreenSpaceReflection", "PointLightCollection", "GBufferMesh", "Texture",
           "GltfMaterial", "GltfPrimitive",
           "SceneToGBufferPass", "synth",
           "hammersley", "env_pyramid", "prefilter_cube", "irradiance_cube",
           "brdf_lut", "cube_sample", "env_sample", "prefilter_points", "env_levels", "cube_texels"]


def _torch():
    import torch

    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _empty(stream, *args, **kw):
    """torch.empty allocated on the stream the work is launched on, so the
    caching allocator orders reuse of the block after that stream's use of
    it (a block allocated on another stream could be handed out again while
    the launch still reads or writes it)."""
    torch = _torch()
    kw.setdefault("device", "cuda")
    if stream is None:
        return torch.empty(*args, **kw)
    with torch.cuda.stream(stream):
        return torch.empty(*args, **kw)


def _ptr(t) -> int:
    if t is None:
        return None
    assert t.is_cuda, "device tensor required"
    assert t.is_contiguous(), "contiguous tensor required"
    return t.data_ptr()


_hip = None


def _hip_copy_dtod(dst: int, src: int, nbytes: int, stream=None) -> None:
    """hipMemcpyAsync device-to-device on torch's current stream."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMemcpyAsync.restype = C.c_int
    rc = _hip.hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(src), nbytes, 3, C.c_void_p(_stream(stream)))
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed ({rc})")


def env_levels(width: int, height: int) -> int:
    return lib().althea_env_levels(width, height)


def cube_texels(size: int, mips: int) -> int:
    return lib().althea_cube_texels(size, mips)


# ------------------------------------------------------------- primitives
def hammersley(n: int, first: int = 0, count: int | None = None, stream=None):
    """(bits uint32 tensor, xi float32 (count, 2)) from the device Hammersley."""
    torch = _torch()
    count = n - first if count is None else count
    bits = _empty(stream, count, dtype=torch.int32, device="cuda")
    xi = _empty(stream, (count, 2), dtype=torch.float32, device="cuda")
    check(lib().althea_hammersley(n, first, count, _ptr(bits), _ptr(xi), _stream(stream)), "althea_hammersley")
    return bits, xi


def env_pyramid(level0, stream=None):
    """Full pyramid of an equirect RGBA32F map (h, w, 4) -> flat (texels, 4)."""
    torch = _torch()
    h, w = int(level0.shape[0]), int(level0.shape[1])
    out = _empty(stream, (lib().althea_env_pyramid_texels(w, h), 4), dtype=torch.float32, device=level0.device)
    check(lib().althea_env_pyramid_build(_ptr(level0), w, h, _ptr(out), _stream(stream)), "althea_env_pyramid_build")
    return out


def prefilter_cube(pyramid, env_w: int, env_h: int, size: int, mips: int, samples: int, fp32_sums: bool = False,
                   stream=None):
    torch = _torch()
    out = _empty(stream, (cube_texels(size, mips), 4), dtype=torch.float32, device=pyramid.device)
    flags = _lib.PRECOMPUTE_FP32_SUMS if fp32_sums else 0
    check(lib().althea_prefilter_cube(_ptr(pyramid), env_w, env_h, size, mips, samples, flags, _ptr(out),
                                      _stream(stream)), "althea_prefilter_cube")
    return out


def prefilter_points(pyramid, env_w: int, env_h: int, samples: int, dir_rough, stream=None):
    torch = _torch()
    out = _empty(stream, (dir_rough.shape[0], 4), dtype=torch.float32, device=pyramid.device)
    check(lib().althea_prefilter_points(_ptr(pyramid), env_w, env_h, samples, _ptr(dir_rough),
                                        int(dir_rough.shape[0]), _ptr(out), _stream(stream)),
          "althea_prefilter_points")
    return out


def irradiance_cube(pyramid, env_w: int, env_h: int, size: int, theta_samples: int = 300, fp32_sums: bool = False,
                    stream=None):
    torch = _torch()
    out = _empty(stream, (6 * size * size, 4), dtype=torch.float32, device=pyramid.device)
    flags = _lib.PRECOMPUTE_FP32_SUMS if fp32_sums else 0
    check(lib().althea_irradiance_cube(_ptr(pyramid), env_w, env_h, size, theta_samples, flags, _ptr(out),
                                       _stream(stream)), "althea_irradiance_cube")
    return out


def brdf_lut(size: int, samples: int, stream=None):
    torch = _torch()
    out = _empty(stream, (size, size, 2), dtype=torch.float32, device="cuda")
    check(lib().althea_brdf_lut(size, samples, _ptr(out), _stream(stream)), "althea_brdf_lut")
    return out


def cube_sample(cube, size: int, mips: int, dir_lod, stream=None):
    torch = _torch()
    out = _empty(stream, (dir_lod.shape[0], 4), dtype=torch.float32, device=cube.device)
    check(lib().althea_cube_sample(_ptr(cube), size, mips, _ptr(dir_lod), int(dir_lod.shape[0]), _ptr(out),
                                   _stream(stream)), "althea_cube_sample")
    return out


def cube_border(cube, size: int, mips: int, stream=None):
    """Seamless-bordered copy of a cube (althea_cube_border): float32
    (althea_cube_bordered_texels(size, mips), 4)."""
    torch = _torch()
    out = _empty(stream, (lib().althea_cube_bordered_texels(size, mips), 4), dtype=torch.float32, device=cube.device)
    check(lib().althea_cube_border(_ptr(cube), size, mips, _ptr(out), _stream(stream)), "althea_cube_border")
    return out


def env_sample(pyramid, env_w: int, env_h: int, dir_lod, stream=None):
    torch = _torch()
    out = _empty(stream, (dir_lod.shape[0], 4), dtype=torch.float32, device=pyramid.device)
    check(lib().althea_env_sample(_ptr(pyramid), env_w, env_h, _ptr(dir_lod), int(dir_lod.shape[0]), _ptr(out),
                                  _stream(stream)), "althea_env_sample")
    return out


# ------------------------------------------------ ImageBasedLighting mirror
@dataclass
class IBLDesc:
    irradiance_size: int = 32
    irradiance_theta_samples: int = 300
    prefiltered_size: int = 128
    prefiltered_mips: int = 5
    prefilter_samples: int = 1024
    brdf_lut_size: int = 128
    brdf_lut_samples: int = 1024
    fp32_sums: bool = False

    def c(self) -> _lib.IBLDescC:
        return _lib.IBLDescC(self.irradiance_size, self.irradiance_theta_samples, self.prefiltered_size,
                             self.prefiltered_mips, self.prefilter_samples, self.brdf_lut_size,
                             self.brdf_lut_samples, _lib.PRECOMPUTE_FP32_SUMS if self.fp32_sums else 0)


class IBLResources:
    """Owns an ``althea_ibl`` handle (IBLResources, ImageBasedLighting.h:24-43)."""

    def __init__(self, desc: IBLDesc):
        self.desc = desc
        h = C.c_void_p()
        check(lib().althea_ibl_create(C.byref(desc.c()), C.byref(h)), "althea_ibl_create")
        self._h = h
        self._destroy = lib().althea_ibl_destroy

    def __del__(self):
        # The destroy function is bound at creation: module globals may be
        # gone when this runs at interpreter exit.
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def precompute(self, env, stream=None) -> None:
        h, w = int(env.shape[0]), int(env.shape[1])
        self._env_keepalive = env
        check(lib().althea_ibl_precompute(self._h, _ptr(env), w, h, _stream(stream)), "althea_ibl_precompute")

    def save_cache(self, directory, stream=None) -> None:
        """Write the precomputed maps in the reference's cache layout
        (<dir>/Prefiltered<m+1>.hdr, <dir>/IrradianceMap.hdr)."""
        check(lib().althea_ibl_save_cache(self._h, os.fsencode(directory), _stream(stream)), "althea_ibl_save_cache")

    def load_cache(self, env, directory, stream=None) -> bool:
        """Fill the handle from a cache written by save_cache and the source
        map `env` (device tensor (h, w, 4)); False when a file is missing."""
        h, w = int(env.shape[0]), int(env.shape[1])
        self._env_keepalive = env
        st = lib().althea_ibl_load_cache(self._h, _ptr(env), w, h, os.fsencode(directory), _stream(stream))
        if st == _lib.STATUS_NOT_FOUND:
            return False
        check(st, "althea_ibl_load_cache")
        return True

    def getHandles(self) -> _lib.IBLViewsC:
        v = _lib.IBLViewsC()
        check(lib().althea_ibl_get_views(self._h, C.byref(v)), "althea_ibl_get_views")
        return v

    views = getHandles

    def state_bytes(self) -> int:
        return lib().althea_ibl_state_bytes(self._h)

    def _tensor_of(self, ptr: int, count: int, comps: int):
        """Device copy (stream-ordered) of `count` x `comps` floats at `ptr`."""
        torch = _torch()
        out = torch.empty((count, comps), dtype=torch.float32, device="cuda")
        _hip_copy_dtod(out.data_ptr(), ptr, count * comps * 4)
        return out

    def prefiltered_tensor(self):
        v = self.views()
        return self._tensor_of(v.prefiltered, cube_texels(v.prefiltered_size, v.prefiltered_mips), 4)

    def irradiance_tensor(self):
        v = self.views()
        return self._tensor_of(v.irradiance, 6 * v.irradiance_size ** 2, 4)

    def lut_tensor(self):
        v = self.views()
        return self._tensor_of(v.brdf_lut, v.brdf_lut_size ** 2, 2)

    def environment_tensor(self):
        v = self.views()
        return self._tensor_of(v.environment, lib().althea_env_pyramid_texels(v.env_width, v.env_height), 4)

    def views_numpy(self) -> dict:
        v = self.views()
        return {
            "env_w": v.env_width, "env_h": v.env_height,
            "prefiltered": self.prefiltered_tensor().cpu().numpy(), "pref_size": v.prefiltered_size,
            "pref_mips": v.prefiltered_mips,
            "irradiance": self.irradiance_tensor().cpu().numpy(), "irr_size": v.irradiance_size,
            "lut": self.lut_tensor().cpu().numpy(), "lut_size": v.brdf_lut_size,
        }

    def prefiltered_texels(self, ids: np.ndarray) -> np.ndarray:
        """RGB of cube texels ids[:, (mip, face, x, y)] (host numpy)."""
        v = self.views()
        cube = self.prefiltered_tensor().cpu().numpy()
        out = np.empty((ids.shape[0], 3), np.float32)
        for k, (m, f, x, y) in enumerate(ids):
            s = max(1, v.prefiltered_size >> int(m))
            off = cube_texels(v.prefiltered_size, int(m)) + (int(f) * s + int(y)) * s + int(x)
            out[k] = cube[off, :3]
        return out


class ImageBasedLighting:
    """ImageBasedLighting::createResources (ImageBasedLighting.h:45-49): the
    environment comes in as a device tensor (h, w, 4) or as the path of an
    RGBE .hdr file, the reference's form (a missing file raises, as
    ImageBasedLighting.cpp:423-425 does)."""

    @staticmethod
    def createResources(env, desc: IBLDesc | None = None, stream=None) -> IBLResources:
        if isinstance(env, (str, os.PathLike)):
            from . import hdr

            env = _torch().from_numpy(hdr.read(env)).cuda()
        res = IBLResources(desc or IBLDesc())
        res.precompute(env, stream)
        return res

    create_resources = createResources

    @staticmethod
    def createResourcesFromContent(content_dir, env_map_name: str, desc: IBLDesc | None = None,
                                   stream=None) -> IBLResources:
        """The reference's createResources(app, cmd, envMapName)
        (ImageBasedLighting.cpp:415-446): the source is
        <content>/HDRI_Skybox/<name>.hdr (missing: raises); the maps come from
        the cache <content>/PrecomputedMaps/<name>/ when it is complete, else
        they are precomputed and the cache is written."""
        from . import hdr

        src = os.path.join(content_dir, "HDRI_Skybox", env_map_name + ".hdr")
        env = _torch().from_numpy(hdr.read(src)).cuda()
        cache = os.path.join(content_dir, "PrecomputedMaps", env_map_name)
        res = IBLResources(desc or IBLDesc())
        if not res.load_cache(env, cache, stream):
            res.precompute(env, stream)
            res.save_cache(cache, stream)
        return res

    @staticmethod
    def createResourcesBatched(envs, descs, tuning=None, stream=None):
        """IBL resources for many environment maps (device tensors (h, w, 4))
        with one batched prefilter launch (althea_ibl_precompute_batched).
        `descs`: one IBLDesc per map or a single one for all; they must share
        mips, samples and precision.  `tuning` (tile_w, tile_h, waves_per_cu)
        or None for the autotuner.  Returns (resources, tuning used)."""
        if isinstance(descs, IBLDesc):
            descs = [descs] * len(envs)
        res = [IBLResources(d) for d in descs]
        views = (_lib.EnvViewC * len(envs))(*[_lib.EnvViewC(_ptr(e), int(e.shape[1]), int(e.shape[0])) for e in envs])
        handles = (C.c_void_p * len(envs))(*[r.handle.value for r in res])
        t_in = _lib.TuningC(*tuning, 0) if tuning is not None else None
        used = _lib.TuningC()
        check(lib().althea_ibl_precompute_batched(handles, views, len(envs), C.byref(t_in) if t_in else None,
                                                  C.byref(used), _stream(stream)), "althea_ibl_precompute_batched")
        return res, (used.tile_w, used.tile_h, used.waves_per_cu)


# ------------------------------------------------------ G-buffer / resolve
class GBufferResources:
    """The four G-buffer planes (position, normal, albedo, metallic-roughness-
    occlusion) as (h, w, 4) device tensors, FP32 or FP16."""

    def __init__(self, position, normal, albedo, mro):
        self.planes = [position, normal, albedo, mro]
        dt = position.dtype
        torch = _torch()
        assert all(p.dtype == dt and p.shape == position.shape for p in self.planes)
        self.height, self.width = int(position.shape[0]), int(position.shape[1])
        self.format = _lib.GBUFFER_FP16 if dt == torch.float16 else _lib.GBUFFER_FP32

    def getHandles(self) -> _lib.GBufferViewC:
        p = self.planes
        return _lib.GBufferViewC(_ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(p[3]), self.width, self.height,
                                 self.format)


class GBufferMesh:
    """The per-vertex inputs of the pass-through G-buffer shader
    (Shaders/GBuffer/GBufferPassThrough.frag): float32 device tensors
    positions, colors, mro (v, 3), normals (v, 3) or None (all zero), uint32
    or int32 indices (t, 3)."""

    def __init__(self, positions, indices, colors, mro, normals=None):
        torch = _torch()
        for t in (positions, colors, mro) + ((normals,) if normals is not None else ()):
            assert t.dtype == torch.float32 and t.is_contiguous() and t.shape == positions.shape, t.shape
        assert indices.dtype in (torch.int32, torch.uint32) and indices.is_contiguous()
        self.positions, self.indices, self.colors, self.mro, self.normals = positions, indices, colors, mro, normals

    @property
    def triangle_count(self) -> int:
        return self.indices.numel() // 3

    def view(self) -> _lib.GBufferMeshC:
        n = self.triangle_count
        return _lib.GBufferMeshC(_ptr(self.positions), _ptr(self.normals), _ptr(self.colors), _ptr(self.mro),
                                 _ptr(self.indices) if n else None, self.positions.numel() // 3, n)


class Texture:
    """An RGBA8 texture with its mip chain (Src/Texture.cpp:50-100) in one
    uint8 device tensor, levels back to back, and its sampler state
    (Src/Sampler.cpp): `level0` an (h, w, 4) uint8 array or tensor; the other
    `mips` - 1 levels are built by the linear blit of Image::generateMipMaps
    (in linear space for sRGB)."""

    def __init__(self, level0, mips: int = 1, srgb: bool = False, mag_filter: int = _lib.FILTER_LINEAR,
                 min_filter: int = _lib.FILTER_LINEAR, mipmap_mode: int = _lib.FILTER_LINEAR,
                 wrap_u: int = _lib.WRAP_REPEAT, wrap_v: int = _lib.WRAP_REPEAT, stream=None):
        torch = _torch()
        t0 = torch.as_tensor(level0).to(torch.uint8)
        self.height, self.width = int(t0.shape[0]), int(t0.shape[1])
        self.mips, self.srgb = mips, srgb
        self.sampler = (mag_filter, min_filter, mipmap_mode, wrap_u, wrap_v)
        n = lib().althea_texture_texels(self.width, self.height, mips)
        if n == 0:
            raise ValueError(f"invalid texture {self.width}x{self.height}, {mips} levels")
        self.data = torch.zeros((n, 4), dtype=torch.uint8, device="cuda")
        self.data[:self.width * self.height].copy_(t0.reshape(-1, 4))
        check(lib().althea_texture_build_mips(_ptr(self.data), self.width, self.height, mips, 1 if srgb else 0,
                                              _stream(stream)),
              "althea_texture_build_mips")

    def level(self, m: int):
        w, h = max(1, self.width >> m), max(1, self.height >> m)
        off = lib().althea_texture_mip_offset(self.width, self.height, m)
        return self.data[off:off + w * h].view(h, w, 4)

    def view(self) -> _lib.TextureViewC:
        return _lib.TextureViewC(_ptr(self.data), self.width, self.height, self.mips, 1 if self.srgb else 0,
                                 *self.sampler)


class GltfMaterial:
    """MaterialConstants of Src/Material.cpp: factors, texture slots (indices
    into the texture list passed with the draw) and texture-coordinate sets;
    alpha_cutoff 0.5 as glTF's default."""

    def __init__(self, base_texture: int, normal_texture: int, metallic_roughness_texture: int,
                 base_color_factor=(1.0, 1.0, 1.0, 1.0), metallic_factor: float = 1.0, roughness_factor: float = 1.0,
                 normal_scale: float = 1.0, alpha_cutoff: float = 0.5, base_uv: int = 0,
                 metallic_roughness_uv: int = 0):
        self.c = _lib.GltfMaterialC((C.c_float * 4)(*base_color_factor), metallic_factor, roughness_factor,
                                    normal_scale, alpha_cutoff, base_texture, normal_texture,
                                    metallic_roughness_texture, base_uv, metallic_roughness_uv)


class GltfPrimitive:
    """A drawn glTF primitive in world space: float32 device tensors
    positions, tangents, bitangents, normals (v, 3), uvs a list of up to four
    (v, 2) tensors, indices (t, 3) int32 / uint32, and a material index."""

    def __init__(self, positions, tangents, bitangents, normals, uvs, indices, material: int):
        self.positions, self.tangents, self.bitangents, self.normals = positions, tangents, bitangents, normals
        self.uvs, self.indices, self.material = list(uvs), indices, material

    def view(self) -> _lib.GltfPrimitiveC:
        uv = [(_ptr(self.uvs[k]) if k < len(self.uvs) and self.uvs[k] is not None else None) for k in range(4)]
        return _lib.GltfPrimitiveC(_ptr(self.positions), _ptr(self.tangents), _ptr(self.bitangents),
                                   _ptr(self.normals), (C.c_void_p * 4)(*uv), _ptr(self.indices),
                                   self.positions.numel() // 3, self.indices.numel() // 3, self.material)


class SceneToGBufferPass:
    """SceneToGBufferPass (Src/DeferredRendering.cpp:268-332) with the
    pass-through subpass: `begin(mesh, globals)` rasterizes a GBufferMesh with
    the camera of `globals` into this pass's GBufferResources (four planes of
    width x height vec4, FP32 or FP16)."""

    def __init__(self, width: int, height: int, fp16: bool = False):
        torch = _torch()
        dt = torch.float16 if fp16 else torch.float32
        self.gbuffer = GBufferResources(*[torch.empty((height, width, 4), dtype=dt, device="cuda")
                                          for _ in range(4)])
        self.workspace = None

    def begin(self, mesh: GBufferMesh, globals_, stream=None) -> GBufferResources:
        torch = _torch()
        g = self.gbuffer
        need = lib().althea_gbuffer_workspace_bytes(g.width, g.height, mesh.triangle_count)
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = _empty(stream, need, dtype=torch.uint8, device="cuda")
        mv = mesh.view()
        gv = g.getHandles()
        check(lib().althea_gbuffer_draw(C.byref(mv), C.byref(globals_), C.byref(gv), _ptr(self.workspace),
                                        _stream(stream)),
              "althea_gbuffer_draw")
        return g

    def begin_gltf(self, primitives, materials, textures, globals_, stream=None) -> GBufferResources:
        """The glTF subpass (Gltf.frag): GltfPrimitive list drawn in order,
        GltfMaterial list, Texture list."""
        torch = _torch()
        g = self.gbuffer
        prims = (_lib.GltfPrimitiveC * max(1, len(primitives)))(*[p.view() for p in primitives])
        mats = (_lib.GltfMaterialC * max(1, len(materials)))(*[m.c for m in materials])
        texs = (_lib.TextureViewC * max(1, len(textures)))(*[t.view() for t in textures])
        need = lib().althea_gbuffer_gltf_workspace_bytes(g.width, g.height, prims, len(primitives), len(materials),
                                                         len(textures))
        if need == 0:
            raise ValueError("invalid glTF draw")
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = _empty(stream, need, dtype=torch.uint8, device="cuda")
        gv = g.getHandles()
        check(lib().althea_gbuffer_draw_gltf(prims, len(primitives), mats, len(materials), texs, len(textures),
                                             C.byref(globals_), C.byref(gv), _ptr(self.workspace), _stream(stream)),
              "althea_gbuffer_draw_gltf")
        return g


class ReflectionBuffer:
    """The frame-sized reflection buffer of Src/ReflectionBuffer.cpp:20-94:
    `mips` levels (5 in the reference) stored back to back in one device
    tensor of (texels, 4), FP16 (the reference's R16G16B16A16_SFLOAT) or
    FP32."""

    def __init__(self, width: int, height: int, mips: int = 5, fp16: bool = True):
        torch = _torch()
        if not 1 <= mips <= _lib.REFLECTION_MAX_MIPS:
            raise ValueError(f"mips {mips} outside 1..{_lib.REFLECTION_MAX_MIPS}")
        self.width, self.height, self.mips, self.fp16 = width, height, mips, fp16
        n = lib().althea_reflection_texels(width, height, mips)
        self.data = torch.zeros((n, 4), dtype=torch.float16 if fp16 else torch.float32, device="cuda")
        # The capture's depth-plane workspace (the reference's depth attachment).
        ws = lib().althea_ssr_workspace_bytes(width, height)
        self.workspace = torch.empty(max(ws, 4), dtype=torch.uint8, device="cuda")

    def view(self) -> _lib.ReflectionViewC:
        return _lib.ReflectionViewC(_ptr(self.data), self.width, self.height, self.mips,
                                    _lib.REFLECTION_FP16 if self.fp16 else _lib.REFLECTION_FP32)

    def extent(self, level: int) -> tuple:
        return max(1, self.width >> level), max(1, self.height >> level)

    def level(self, level: int):
        """Level `level` as a (h, w, 4) view of the device tensor."""
        w, h = self.extent(level)
        off = lib().althea_reflection_mip_offset(self.width, self.height, level)
        return self.data[off:off + w * h].view(h, w, 4)


class ScreenSpaceReflection:
    """Src/ScreenSpaceReflection.cpp: captureReflection fills level 0 of a
    ReflectionBuffer (Shaders/SSR.frag), convolveReflectionBuffer the glossy
    levels (Shaders/SSRGlossyConvolve.comp)."""

    @staticmethod
    def captureReflection(gbuffer, ibl, lights, globals_, reflection: ReflectionBuffer, shadows=None,
                          stream=None):
        """`lights` an (n, 8) device tensor or a PointLightCollection; with
        `shadows` (a PointLightCollection with shadow maps, or the lights'
        own collection) the hit shading runs pbrMaterial's shadow test."""
        if not isinstance(gbuffer, GBufferResources):
            gbuffer = GBufferResources(*gbuffer)
        if isinstance(lights, PointLightCollection):
            if shadows is None and lights.shadow_storage is not None:
                shadows = lights
            lights = lights.lights
        views = ibl.views() if isinstance(ibl, IBLResources) else ibl
        gv = gbuffer.getHandles()
        rv = reflection.view()
        n = int(lights.shape[0]) if lights is not None else 0
        if shadows is not None:
            if shadows.shadow_storage is None or shadows.getCount() < n:
                raise ValueError("shadows: need one shadow cube per light")
            check(lib().althea_ssr_capture_shadowed(C.byref(gv), C.byref(views), _ptr(lights) if n else None, n,
                                                    C.byref(globals_), C.byref(rv), _ptr(shadows.shadow_storage),
                                                    shadows.shadow_size, _ptr(reflection.workspace),
                                                    _stream(stream)),
                  "althea_ssr_capture_shadowed")
            return reflection
        check(lib().althea_ssr_capture(C.byref(gv), C.byref(views), _ptr(lights) if n else None, n,
                                       C.byref(globals_), C.byref(rv), _ptr(reflection.workspace), _stream(stream)),
              "althea_ssr_capture")
        return reflection

    @staticmethod
    def convolveReflectionBuffer(reflection: ReflectionBuffer, stream=None):
        rv = reflection.view()
        check(lib().althea_reflection_convolve(C.byref(rv), _stream(stream)), "althea_reflection_convolve")
        return reflection


class PointLightCollection:
    """The point lights of Src/PointLight.cpp with their shadow cubes:
    `lights` an (n, 8) float32 device tensor (position xyz, pad, emission
    rgb, pad, as PointLight in Include/Althea/PointLight.h), `shadow_size`
    the cube extent (256 in the reference, PointLight.cpp:35).  drawShadowMaps
    renders the cubes of every light from a triangle mesh (world-space float32
    positions (v, 3), uint32 indices (t, 3), device tensors).  The cubes are
    stored with a one-texel border per face (`shadow_storage`, (n, 6, S + 2,
    S + 2)); getShadowMapImage is the (n, 6, S, S) view of the faces."""

    def __init__(self, light_count: int, create_shadow_map: bool = True, shadow_size: int = 256):
        torch = _torch()
        self.lights = torch.zeros((light_count, 8), dtype=torch.float32, device="cuda")
        self.shadow_size = shadow_size
        P = shadow_size + 2
        self.shadow_storage = (torch.ones((light_count, 6, P, P), dtype=torch.float32, device="cuda")
                               if create_shadow_map else None)
        self.workspace = None

    @property
    def shadow_maps(self):
        return None if self.shadow_storage is None else self.shadow_storage[:, :, 1:-1, 1:-1]

    @classmethod
    def from_lights(cls, lights, create_shadow_map: bool = True, shadow_size: int = 256):
        c = cls(int(lights.shape[0]), create_shadow_map, shadow_size)
        c.lights.copy_(lights.reshape(-1, 8))
        return c

    def getCount(self) -> int:
        return int(self.lights.shape[0])

    def setLight(self, light_id: int, position, emission) -> None:
        torch = _torch()
        row = torch.tensor([*position, 0.0, *emission, 0.0], dtype=torch.float32)
        self.lights[light_id].copy_(row)

    def getShadowMapExtent(self) -> tuple:
        return self.shadow_size, self.shadow_size

    def getShadowMapImage(self):
        return self.shadow_maps

    def drawShadowMaps(self, positions, indices, scheduled: bool = True, stream=None):
        """The cubes of every light from a mesh.  `scheduled` (default) uses
        the workspace path that deals texels out evenly; False the
        workspace-free one (the same cubes)."""
        if self.shadow_maps is None:
            raise ValueError("PointLightCollection was created without shadow maps")
        torch = _torch()
        assert positions.dtype == torch.float32 and positions.is_contiguous() and positions.shape[-1] == 3
        assert indices.dtype in (torch.int32, torch.uint32) and indices.is_contiguous()
        tris = indices.numel() // 3
        mesh = _lib.MeshViewC(_ptr(positions) if tris else None, _ptr(indices) if tris else None,
                              int(positions.numel() // 3), tris)
        assert self.shadow_storage[0].numel() == lib().althea_shadow_cube_floats(self.shadow_size)
        if not scheduled:
            check(lib().althea_shadow_cubes(C.byref(mesh), _ptr(self.lights), self.getCount(), self.shadow_size,
                                            _ptr(self.shadow_storage), _stream(stream)),
                  "althea_shadow_cubes")
            return self.shadow_maps
        need = lib().althea_shadow_cubes_workspace_bytes(self.getCount(), tris)
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = _empty(stream, max(need, 256), dtype=torch.uint8, device="cuda")
        check(lib().althea_shadow_cubes_ws(C.byref(mesh), _ptr(self.lights), self.getCount(), self.shadow_size,
                                           _ptr(self.shadow_storage), _ptr(self.workspace), _stream(stream)),
              "althea_shadow_cubes_ws")
        return self.shadow_maps

    def drawShadowMapsGltf(self, primitives, materials, textures, stream=None):
        """The cubes of every light from glTF primitives (GltfPrimitive list,
        GltfMaterial list, Texture list), with the shadow shaders' alpha test:
        fragments whose base colour alpha is below the material's cut-off
        cast no shadow."""
        if self.shadow_maps is None:
            raise ValueError("PointLightCollection was created without shadow maps")
        torch = _torch()
        prims = (_lib.GltfPrimitiveC * max(1, len(primitives)))(*[p.view() for p in primitives])
        mats = (_lib.GltfMaterialC * max(1, len(materials)))(*[m.c for m in materials])
        texs = (_lib.TextureViewC * max(1, len(textures)))(*[t.view() for t in textures])
        need = lib().althea_shadow_cubes_gltf_workspace_bytes(prims, len(primitives), self.getCount())
        if need == 0:
            raise ValueError("invalid glTF shadow draw")
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = _empty(stream, need, dtype=torch.uint8, device="cuda")
        check(lib().althea_shadow_cubes_gltf(prims, len(primitives), mats, len(materials), texs, len(textures),
                                             _ptr(self.lights), self.getCount(), self.shadow_size,
                                             _ptr(self.shadow_storage), _ptr(self.workspace), _stream(stream)),
              "althea_shadow_cubes_gltf")
        return self.shadow_maps


class DeferredRendering:
    @staticmethod
    def resolve(gbuffer, ibl, lights, globals_, ssao=None, tonemap: bool = False, out_fp16: bool = False,
                ao_out=None, out=None, reflection: ReflectionBuffer | None = None,
                shadows: PointLightCollection | None = None, stream=None):
        """One frame of the deferred resolve.  `gbuffer` is GBufferResources
        or a list of the four planes; `ibl` an IBLResources or IBLViewsC;
        `lights` an (n, 8) float32 device tensor, or a PointLightCollection
        (whose shadow cubes are then read when it has them); `reflection`
        (optional) a ReflectionBuffer blended into the prefiltered term;
        `shadows` (optional) the PointLightCollection whose cubes shadow
        `lights`."""
        if isinstance(lights, PointLightCollection):
            if shadows is None and lights.shadow_maps is not None:
                shadows = lights
            lights = lights.lights
        torch = _torch()
        if not isinstance(gbuffer, GBufferResources):
            gbuffer = GBufferResources(*gbuffer)
        views = ibl.views() if isinstance(ibl, IBLResources) else ibl
        gv = gbuffer.getHandles()
        n = int(lights.shape[0]) if lights is not None else 0
        if out is None:
            dt = torch.float16 if out_fp16 else torch.float32
            out = _empty(stream, (gbuffer.height * gbuffer.width, 4), dtype=dt, device="cuda")
        flags = (_lib.RESOLVE_TONEMAP if tonemap else 0) | (_lib.RESOLVE_OUT_FP16 if out_fp16 else 0)
        ssao_p = C.byref(ssao) if ssao is not None else None
        if shadows is not None:
            if shadows.shadow_storage is None or shadows.getCount() < n:
                raise ValueError("shadows: need one shadow cube per light")
            rv = reflection.view() if reflection is not None else None
            check(lib().althea_deferred_resolve_shadowed(C.byref(gv), C.byref(views), _ptr(lights) if n else None, n,
                                                         C.byref(globals_), ssao_p,
                                                         C.byref(rv) if rv is not None else None,
                                                         _ptr(shadows.shadow_storage), shadows.shadow_size, flags,
                                                         _ptr(out), _ptr(ao_out), _stream(stream)),
                  "althea_deferred_resolve_shadowed")
            return out
        if reflection is not None:
            rv = reflection.view()
            check(lib().althea_deferred_resolve_reflect(C.byref(gv), C.byref(views), _ptr(lights) if n else None, n,
                                                        C.byref(globals_), ssao_p, C.byref(rv), flags, _ptr(out),
                                                        _ptr(ao_out), _stream(stream)),
                  "althea_deferred_resolve_reflect")
            return out
        check(lib().althea_deferred_resolve(C.byref(gv), C.byref(views), _ptr(lights) if n else None, n,
                                            C.byref(globals_), ssao_p, flags, _ptr(out), _ptr(ao_out),
                                            _stream(stream)),
              "althea_deferred_resolve")
        return out

    @staticmethod
    def resolve_mlp(gbuffer, ibl, mlp, lights, globals_, ssao=None, tonemap: bool = False, out_fp16: bool = False,
                    ao_out=None, out=None, mlp_used=None, stream=None):
        """The resolve with the prefiltered term from `mlp` (a PrefilterMLP
        whose error monitor was built for this IBL's cube) where the monitor
        accepts, the cube elsewhere.  `mlp_used` (optional uint8 device tensor
        of width*height) receives 1 where the MLP supplied the term."""
        torch = _torch()
        if not isinstance(gbuffer, GBufferResources):
            gbuffer = GBufferResources(*gbuffer)
        views = ibl.views() if isinstance(ibl, IBLResources) else ibl
        gv = gbuffer.getHandles()
        n = int(lights.shape[0]) if lights is not None else 0
        if out is None:
            dt = torch.float16 if out_fp16 else torch.float32
            out = _empty(stream, (gbuffer.height * gbuffer.width, 4), dtype=dt, device="cuda")
        if mlp_used is not None:
            assert mlp_used.dtype == torch.uint8 and mlp_used.numel() >= gbuffer.height * gbuffer.width
        flags = (_lib.RESOLVE_TONEMAP if tonemap else 0) | (_lib.RESOLVE_OUT_FP16 if out_fp16 else 0)
        check(lib().althea_deferred_resolve_mlp(C.byref(gv), C.byref(views), mlp.handle, _ptr(lights) if n else None,
                                                n, C.byref(globals_), C.byref(ssao) if ssao is not None else None,
                                                flags, _ptr(out), _ptr(ao_out), _ptr(mlp_used), _stream(stream)),
              "althea_deferred_resolve_mlp")
        return out
