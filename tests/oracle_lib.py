"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.  Arrays are numpy; float32 inputs are read
as given and every result is float64.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_PATH = Path(__file__).resolve().parent.parent / "oracle" / "liboracle.so"

vp, i32, u32, dbl = C.c_void_p, C.c_int, C.c_uint32, C.c_double


class PortArgs(C.Structure):
    """port_resolve_args of oracle/port_resolve.cpp."""
    _fields_ = [("position", vp), ("normal", vp), ("albedo", vp), ("mro", vp), ("width", i32), ("height", i32),
                ("env", vp), ("env_w", i32), ("env_h", i32), ("prefiltered", vp), ("pref_size", i32),
                ("pref_mips", i32), ("irradiance", vp), ("irr_size", i32), ("lut", vp), ("lut_size", i32),
                ("lights", vp), ("light_count", i32), ("view", vp), ("proj_view", vp), ("ray_affine", vp),
                ("exposure", C.c_float), ("tonemap", i32), ("ssao_on", i32), ("ssao_samples", i32),
                ("ssao_radius", C.c_float), ("ssao_bias", C.c_float), ("ssao_kernel", vp), ("ssao_noise", vp)]


class ResolveArgs(C.Structure):
    _fields_ = [("position", vp), ("normal", vp), ("albedo", vp), ("mro", vp), ("width", i32), ("height", i32),
                ("env", vp), ("prefiltered", vp), ("pref_size", i32), ("pref_mips", i32), ("irradiance", vp),
                ("irr_size", i32), ("lut", vp), ("lut_size", i32), ("lights", vp), ("light_count", i32),
                ("projection", vp), ("inverse_projection", vp), ("view", vp), ("inverse_view", vp),
                ("exposure", C.c_float), ("tonemap", i32), ("ssao_on", i32), ("ssao_kernel", vp),
                ("ssao_noise", vp), ("ssao_samples", i32), ("ssao_radius", C.c_float), ("ssao_bias", C.c_float),
                ("reflection", vp), ("refl_width", i32), ("refl_height", i32), ("refl_mips", i32),
                ("shadow_cubes", vp), ("shadow_size", i32)]


_L = None


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} missing: run __graft_entry__.build()")
        _L = load(ORACLE_PATH)
    return _L


def load(path) -> C.CDLL:
    """An oracle library at `path` with its argument types declared (the
    mutation tests load rebuilt variants next to the shipped one)."""
    L = C.CDLL(str(path))
    L.oracle_hash_rng.argtypes = [u32, u32, i32, vp, vp]
    L.oracle_hammersley.argtypes = [u32, u32, u32, vp, vp, vp]
    L.oracle_frame.argtypes = [vp, i32, vp]
    L.oracle_ggx_sample.argtypes = [vp, i32, dbl, vp]
    L.oracle_ggx_d.argtypes = [dbl, dbl]
    L.oracle_ggx_d.restype = dbl
    L.oracle_mip_count.argtypes = [i32, i32]
    L.oracle_env_create.argtypes = [vp, i32, i32]
    L.oracle_env_create.restype = vp
    L.oracle_env_destroy.argtypes = [vp]
    L.oracle_env_texels.argtypes = [vp]
    L.oracle_env_texels.restype = C.c_size_t
    L.oracle_env_pyramid.argtypes = [vp, vp]
    L.oracle_env_lookup.argtypes = [vp, vp, vp, i32, i32, vp]
    L.oracle_cube_texel_dirs.argtypes = [i32, vp, i32, vp]
    L.oracle_cube_select.argtypes = [vp, i32, vp, vp]
    L.oracle_cube_fold.argtypes = [i32, i32, i32, i32, vp]
    L.oracle_cube_sample.argtypes = [vp, i32, i32, vp, vp, i32, vp]
    L.oracle_prefilter_texels.argtypes = [vp, i32, i32, i32, vp, i32, vp]
    L.oracle_prefilter_points.argtypes = [vp, i32, vp, i32, vp]
    L.oracle_irradiance_texels.argtypes = [vp, i32, i32, vp, i32, vp]
    L.oracle_prefilter_analytic.argtypes = [i32, vp, vp, i32, dbl, i32, i32, i32, vp]
    L.oracle_irradiance_analytic.argtypes = [i32, vp, vp, i32, i32, i32, i32, vp]
    L.oracle_brdf_lut.argtypes = [i32, i32, vp, i32, vp]
    L.oracle_brdf_lut_point.argtypes = [dbl, dbl, i32, vp]
    L.oracle_resolve.argtypes = [C.POINTER(ResolveArgs), vp, i32, vp, vp, vp]
    L.oracle_resolve_ex.argtypes = [C.POINTER(ResolveArgs), vp, i32, vp, vp, vp, vp, vp]
    L.oracle_pbr_material.argtypes = [vp, dbl, dbl, dbl, vp, i32, vp, i32, vp]
    L.oracle_ssr.argtypes = [C.POINTER(ResolveArgs), vp, i32, vp, vp, vp, vp, vp]
    L.oracle_gbuffer.argtypes = [vp, i32, i32, vp, vp, vp, vp, vp, i32, vp, i32, vp, vp, vp]
    L.oracle_gbuffer_gltf.argtypes = [vp, i32, i32, vp, i32, vp, vp, vp, i32, vp, vp, vp]
    L.oracle_texture_sample.argtypes = [vp, vp, i32, vp, vp]
    L.oracle_texture_mips.argtypes = [vp, i32, i32, i32, i32]
    L.oracle_shadow_cubes.argtypes = [vp, vp, i32, vp, i32, i32, vp]
    L.oracle_shadow_cubes_gltf.argtypes = [vp, i32, vp, vp, vp, i32, i32, vp, vp]
    L.oracle_convolve_level.argtypes = [vp, i32, i32, i32, i32, i32, vp]
    L.oracle_reflection_sample.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp]
    L.oracle_mlp_param_count.restype = i32
    L.oracle_mlp_encode.argtypes = [vp, i32, vp]
    L.oracle_mlp_forward.argtypes = [vp, vp, i32, i32, vp]
    L.oracle_mlp_loss_grad.argtypes = [vp, vp, vp, i32, i32, vp]
    L.oracle_mlp_loss_grad.restype = dbl
    L.oracle_mlp_loss_grad64.argtypes = [vp, vp, vp, i32, vp]
    L.oracle_mlp_loss_grad64.restype = dbl
    L.oracle_adam.argtypes = [vp, vp, vp, vp, i32, i32, dbl, dbl, dbl, dbl]
    L.oracle_bf16_round.argtypes = [C.c_float]
    L.oracle_bf16_round.restype = C.c_float
    L.oracle_max_threads.restype = i32
    L.oracle_prefilter_level_constant.argtypes = [vp, i32, dbl, i32, i32, i32, vp]
    L.oracle_prefilter_crosscheck.argtypes = [vp, vp, i32, i32, i32, dbl, i32, i32, vp, vp, vp]
    L.oracle_ssao_compare.argtypes = [C.POINTER(ResolveArgs), vp, i32, vp, vp]
    if hasattr(L, "port_resolve"):  # the FP32 port (absent from oracle-only rebuilds)
        L.port_resolve.argtypes = [C.POINTER(PortArgs), vp, i32, i32, C.c_uint, vp]
    return L


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous, "contiguous array required"
    return a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------- primitives
def hash_rng(sx: int, sy: int, count: int):
    u = np.empty(count, np.uint32)
    f = np.empty(count, np.float32)
    lib().oracle_hash_rng(sx, sy, count, _p(u), _p(f))
    return u, f


def hammersley(n: int, first: int, count: int):
    bits = np.empty(count, np.uint32)
    x32 = np.empty((count, 2), np.float32)
    x64 = np.empty((count, 2), np.float64)
    lib().oracle_hammersley(n, first, count, _p(bits), _p(x32), _p(x64))
    return bits, x32, x64


def frame(raw_dirs) -> np.ndarray:
    d = _c(raw_dirs, np.float64)
    out = np.empty((d.shape[0], 3, 3), np.float64)
    lib().oracle_frame(_p(d), d.shape[0], _p(out))
    return out  # rows: t, b, n


def ggx_sample(xi, a2: float) -> np.ndarray:
    x = _c(xi, np.float64)
    out = np.empty((x.shape[0], 3), np.float64)
    lib().oracle_ggx_sample(_p(x), x.shape[0], a2, _p(out))
    return out


def ggx_d(n_dot_h: float, a2: float) -> float:
    return lib().oracle_ggx_d(n_dot_h, a2)


def mip_count(w: int, h: int) -> int:
    return lib().oracle_mip_count(w, h)


def cube_texel_dirs(size: int, ids) -> np.ndarray:
    i = _c(ids, np.int32)
    out = np.empty((i.shape[0], 3), np.float64)
    lib().oracle_cube_texel_dirs(size, _p(i), i.shape[0], _p(out))
    return out


def cube_select(dirs):
    d = _c(dirs, np.float64)
    face = np.empty(d.shape[0], np.int32)
    st = np.empty((d.shape[0], 2), np.float64)
    lib().oracle_cube_select(_p(d), d.shape[0], _p(face), _p(st))
    return face, st


def cube_fold(face: int, i: int, j: int, size: int):
    out = np.empty(3, np.int32)
    lib().oracle_cube_fold(face, i, j, size, _p(out))
    return tuple(int(v) for v in out)


def cube_sample(cube, size: int, mips: int, dirs, lods) -> np.ndarray:
    c = _c(cube, np.float32)
    d = _c(dirs, np.float64)
    l = _c(lods, np.float64)
    out = np.empty((d.shape[0], 3), np.float64)
    lib().oracle_cube_sample(_p(c), size, mips, _p(d), _p(l), d.shape[0], _p(out))
    return out


def prefilter_analytic(kind: int, params, raw_dirs, roughness: float, samples: int, env_w: int = 512,
                       env_h: int = 256) -> np.ndarray:
    p = _c(params, np.float64)
    d = _c(raw_dirs, np.float64)
    out = np.empty((d.shape[0], 3), np.float64)
    lib().oracle_prefilter_analytic(kind, _p(p), _p(d), d.shape[0], roughness, samples, env_w, env_h, _p(out))
    return out


def prefilter_level_constant(raw_dirs, roughness: float, samples: int, env_w: int, env_h: int) -> np.ndarray:
    """Prefilter over a pyramid whose level l is the constant 2^l."""
    d = _c(raw_dirs, np.float64)
    out = np.empty((d.shape[0], 3), np.float64)
    lib().oracle_prefilter_level_constant(_p(d), d.shape[0], roughness, samples, env_w, env_h, _p(out))
    return out


def irradiance_analytic(kind: int, params, raw_dirs, env_w: int, env_h: int, theta: int = 300) -> np.ndarray:
    p = _c(params, np.float64)
    d = _c(raw_dirs, np.float64)
    out = np.empty((d.shape[0], 3), np.float64)
    lib().oracle_irradiance_analytic(kind, _p(p), _p(d), d.shape[0], env_w, env_h, theta, _p(out))
    return out


def brdf_lut(size: int, samples: int, ids=None) -> np.ndarray:
    if ids is None:
        yy, xx = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
        ids = np.stack([xx.ravel(), yy.ravel()], 1)
    i = _c(ids, np.int32)
    out = np.empty((i.shape[0], 2), np.float64)
    lib().oracle_brdf_lut(size, samples, _p(i), i.shape[0], _p(out))
    return out


def brdf_lut_point(n_dot_v: float, roughness: float, samples: int):
    out = np.empty(2, np.float64)
    lib().oracle_brdf_lut_point(n_dot_v, roughness, samples, _p(out))
    return out


# ------------------------------------------------------------ environment
class Env:
    """FP64 pyramid of an equirect RGBA32F map (h, w, 4)."""

    def __init__(self, level0: np.ndarray):
        self.level0 = _c(level0, np.float32)
        self.h, self.w = self.level0.shape[:2]
        self._h = lib().oracle_env_create(_p(self.level0), self.w, self.h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_env_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def pyramid(self) -> np.ndarray:
        n = lib().oracle_env_texels(self._h)
        out = np.empty((n, 4), np.float64)
        lib().oracle_env_pyramid(self._h, _p(out))
        return out

    def lookup(self, dirs, lods=None, mode: int = 0) -> np.ndarray:
        d = _c(dirs, np.float64)
        l = _c(np.zeros(d.shape[0]) if lods is None else lods, np.float64)
        out = np.empty((d.shape[0], 3), np.float64)
        lib().oracle_env_lookup(self._h, _p(d), _p(l), d.shape[0], mode, _p(out))
        return out

    def prefilter_texels(self, face_size: int, mips: int, samples: int, ids) -> np.ndarray:
        i = _c(ids, np.int32)
        out = np.empty((i.shape[0], 4), np.float64)
        lib().oracle_prefilter_texels(self._h, face_size, mips, samples, _p(i), i.shape[0], _p(out))
        return out

    def prefilter_points(self, samples: int, dir_rough) -> np.ndarray:
        d = _c(dir_rough, np.float32)
        out = np.empty((d.shape[0], 3), np.float64)
        lib().oracle_prefilter_points(self._h, samples, _p(d), d.shape[0], _p(out))
        return out

    def prefilter_crosscheck(self, texels, w: int, h: int, roughness: float, ref_samples: int = 10000,
                             ham_samples: int = 10000):
        """Reference hash-RNG estimator (value, standard error) and the
        Hammersley estimator at the reference's equirect texels (x, y) of a
        w x h mip."""
        t = _c(texels, np.int32)
        ref, se, ham = (np.empty((t.shape[0], 3), np.float64) for _ in range(3))
        lib().oracle_prefilter_crosscheck(self._h, _p(t), t.shape[0], w, h, roughness, ref_samples, ham_samples,
                                          _p(ref), _p(se), _p(ham))
        return ref, se, ham

    def irradiance_texels(self, face_size: int, ids, theta: int = 300) -> np.ndarray:
        i = _c(ids, np.int32)
        out = np.empty((i.shape[0], 3), np.float64)
        lib().oracle_irradiance_texels(self._h, face_size, theta, _p(i), i.shape[0], _p(out))
        return out


# ----------------------------------------------------------------- resolve
def _resolve_args(scene, views: dict, env, tonemap: bool, gbuffer, reflection, keep: list,
                  shadows=None) -> ResolveArgs:
    """ResolveArgs over numpy copies that `keep` holds alive.  `reflection`:
    (float32 array of every level, (texels, 4), mips) or None; `shadows`:
    float32 cubes (lights, 6, S, S) or None."""
    planes = gbuffer if gbuffer is not None else (scene.position, scene.normal, scene.albedo, scene.material)
    planes = [_c(p, np.float32) for p in planes]
    pref = _c(views["prefiltered"], np.float32)
    irr = _c(views["irradiance"], np.float32)
    lut = _c(views["lut"], np.float32)
    lights = _c(scene.lights, np.float32)
    g = _c(scene.globals_arr, np.float32)
    mats = [_c(g[16 * k:16 * k + 16], np.float32) for k in (0, 1, 2, 4)]
    kern = _c(scene.ssao_kernel, np.float32)
    noise = _c(scene.ssao_noise, np.float32)
    chain, mips = (None, 0) if reflection is None else (_c(reflection[0], np.float32), int(reflection[1]))
    cubes = None if shadows is None else _c(shadows, np.float32)
    keep += [planes, pref, irr, lut, lights, g, mats, kern, noise, chain, cubes]
    return ResolveArgs(_p(planes[0]), _p(planes[1]), _p(planes[2]), _p(planes[3]), scene.width, scene.height,
                       env.handle if env is not None else None, _p(pref), views["pref_size"], views["pref_mips"],
                       _p(irr), views["irr_size"], _p(lut), views["lut_size"], _p(lights), lights.shape[0],
                       _p(mats[0]), _p(mats[1]), _p(mats[2]), _p(mats[3]), float(g[101]), 1 if tonemap else 0,
                       1 if scene.ssao_enabled else 0, _p(kern), _p(noise), kern.shape[0], scene.ssao_radius,
                       scene.ssao_bias, _p(chain) if chain is not None else None, scene.width, scene.height, mips,
                       _p(cubes) if cubes is not None else None, cubes.shape[-1] if cubes is not None else 0)


def resolve(scene, views: dict, env: Env, pixels, tonemap: bool = False, gbuffer=None, refl=None,
            query: bool = False, reflection=None, shadows=None):
    """Oracle resolve of pixels[:, (x, y)] for a synth.Scene and IBL views
    (numpy dict with prefiltered / irradiance / lut arrays and sizes).
    Returns (rgb, ao, decision margin), plus the (R, roughness) query of each
    pixel's prefiltered lookup when `query`.  `refl` (n, 3): prefiltered
    radiance replacing the cube lookup where refl[i, 0] is not NaN.
    `reflection`: (levels as a (texels, 4) float32 array, mips), the
    reflection buffer blended into the prefiltered term.  `shadows`: float32
    shadow cubes (lights, 6, S, S) read by the point-light shadow test."""
    keep = []
    a = _resolve_args(scene, views, env, tonemap, gbuffer, reflection, keep, shadows)
    px = _c(pixels, np.int32)
    rgb = np.empty((px.shape[0], 3), np.float64)
    ao = np.empty(px.shape[0], np.float64)
    margin = np.empty(px.shape[0], np.float64)
    r_in = _c(refl, np.float64) if refl is not None else None
    q = np.empty((px.shape[0], 4), np.float64) if query else None
    lib().oracle_resolve_ex(C.byref(a), _p(px), px.shape[0], _p(r_in) if r_in is not None else None, _p(rgb),
                            _p(ao), _p(margin), _p(q) if q is not None else None)
    return (rgb, ao, margin, q) if query else (rgb, ao, margin)


def ssao_compare(scene, pixels):
    """AO of the reference's 24-ray march (SSAO.glsl:29-85) and of the
    hemisphere kernel at pixels[:, (x, y)]; NaN where uncovered."""
    keep = []
    views = dict(prefiltered=np.zeros((6, 4), np.float32), pref_size=1, pref_mips=1,
                 irradiance=np.zeros((6, 4), np.float32), irr_size=1, lut=np.zeros((1, 2), np.float32), lut_size=1)
    a = _resolve_args(scene, views, None, False, None, None, keep)
    px = _c(pixels, np.int32)
    ref = np.empty(px.shape[0], np.float64)
    ker = np.empty(px.shape[0], np.float64)
    lib().oracle_ssao_compare(C.byref(a), _p(px), px.shape[0], _p(ref), _p(ker))
    return ref, ker


# Steps of the FP32 port that can be put back to plain FP32 addressing
# (error attribution): background, mip-0 reflection tap, half vector.
PORT_FP32_BACKGROUND, PORT_FP32_REFLECTION, PORT_FP32_HALF_VECTOR = 1, 2, 4


def port_resolve(scene, views: dict, env_level0, pixels=None, threads: int = 0, fp32_steps: int = 0,
                 tonemap: bool = False) -> np.ndarray:
    """FP32 CPU port of the resolve (oracle/port_resolve.cpp) for a
    synth.Scene, IBL views (as for resolve()) and the environment's level 0
    (h, w, 4): float32 (n, 4) at pixels[:, (x, y)], or the whole frame in row
    order when pixels is None.  `threads` 0 uses every OpenMP thread."""
    W, H = scene.width, scene.height
    g = scene.globals_arr.astype(np.float64)
    proj, iproj = g[0:16].reshape(4, 4).T, g[16:32].reshape(4, 4).T
    view, iview = g[32:48].reshape(4, 4).T, g[64:80].reshape(4, 4).T

    def e_of(nx, ny):
        e = iproj @ np.array([nx, ny, 0.0, 1.0])
        return iview[:3, :3] @ e[:3]

    o, ax, ay = e_of(-1.0, -1.0), e_of(1.0, -1.0), e_of(-1.0, 1.0)
    aff = _c(np.concatenate([o, (ax - o) / W, (ay - o) / H]), np.float64)
    pv = _c((proj @ view).T.ravel(), np.float32)
    vw = _c(scene.globals_arr[32:48], np.float32)
    planes = [_c(p, np.float32) for p in (scene.position, scene.normal, scene.albedo, scene.material)]
    env = _c(env_level0, np.float32)
    pref, irr, lut = (_c(views[k], np.float32) for k in ("prefiltered", "irradiance", "lut"))
    lights = _c(scene.lights, np.float32)
    kern, noise = _c(scene.ssao_kernel, np.float32), _c(scene.ssao_noise, np.float32)
    a = PortArgs(*[_p(x) for x in planes], W, H, _p(env), env.shape[1], env.shape[0], _p(pref), views["pref_size"],
                 views["pref_mips"], _p(irr), views["irr_size"], _p(lut), views["lut_size"], _p(lights),
                 lights.shape[0], _p(vw), _p(pv), _p(aff), float(scene.globals_arr[101]), 1 if tonemap else 0,
                 1 if scene.ssao_enabled else 0, kern.shape[0], scene.ssao_radius, scene.ssao_bias, _p(kern),
                 _p(noise))
    px = None if pixels is None else _c(pixels, np.int32)
    n = W * H if px is None else px.shape[0]
    out = np.empty((n, 4), np.float32)
    lib().port_resolve(C.byref(a), None if px is None else _p(px), n, threads, fp32_steps, _p(out))
    return out


def gbuffer(scene, positions, indices, colors, mro, normals=None, pixels=None):
    """Oracle G-buffer producer with the camera of a synth.Scene at
    pixels[:, (x, y)] (all pixels by default): (planes (n, 4, 4) position /
    normal / albedo / mro, margin (n,), triangle (n,), -1 uncovered)."""
    g = _c(scene.globals_arr, np.float32)
    mats = [_c(g[16 * k:16 * k + 16], np.float32) for k in (0, 1, 2, 4)]
    cam = (vp * 4)(*[_p(m) for m in mats])
    if pixels is None:
        y, x = np.mgrid[0:scene.height, 0:scene.width]
        pixels = np.stack([x.ravel(), y.ravel()], 1)
    px = _c(pixels, np.int32)
    pos, col, m = _c(positions, np.float32), _c(colors, np.float32), _c(mro, np.float32)
    nrm = _c(normals, np.float32) if normals is not None else None
    idx = _c(indices, np.uint32)
    out = np.empty((px.shape[0], 16), np.float64)
    margin = np.empty(px.shape[0], np.float64)
    tri = np.empty(px.shape[0], np.int32)
    lib().oracle_gbuffer(C.cast(cam, vp), scene.width, scene.height, _p(pos), _p(nrm) if nrm is not None else None, _p(col),
                         _p(m), _p(idx), idx.size // 3, _p(px), px.shape[0], _p(out), _p(margin), _p(tri))
    return out.reshape(-1, 4, 4), margin, tri


class OracleTexC(C.Structure):
    _fields_ = [("texels", vp), ("w", i32), ("h", i32), ("mips", i32), ("srgb", i32), ("mag", i32), ("min", i32),
                ("mipmode", i32), ("wrap_u", i32), ("wrap_v", i32)]


class OracleGltfPrimC(C.Structure):
    _fields_ = [("pos", vp), ("tan", vp), ("bit", vp), ("nrm", vp), ("uv", vp * 4), ("idx", vp),
                ("vertex_count", i32), ("tris", i32), ("material", i32)]


class OracleTex:
    """A host RGBA8 mip chain (texels (n, 4) uint8) and sampler state, as
    the oracle's TexView."""

    def __init__(self, texels, w, h, mips, srgb=0, mag=1, min_=1, mipmode=1, wrap_u=0, wrap_v=0):
        self.texels = _c(texels, np.uint8)
        self.c = OracleTexC(_p(self.texels), w, h, mips, srgb, mag, min_, mipmode, wrap_u, wrap_v)


def texture_mips(level0, mips: int, srgb: bool) -> np.ndarray:
    """The whole mip chain (texels (n, 4) uint8) built by the oracle's blit."""
    h, w = level0.shape[:2]
    n = sum(max(1, w >> m) * max(1, h >> m) for m in range(mips))
    out = np.zeros((n, 4), np.uint8)
    out[:w * h] = level0.reshape(-1, 4)
    lib().oracle_texture_mips(_p(out), w, h, mips, 1 if srgb else 0)
    return out


def texture_sample(tex: OracleTex, q) -> tuple:
    """Samples at q (n, 6): u, v, du/dx, dv/dx, du/dy, dv/dy -> (rgba (n, 4),
    decision margin (n,))."""
    qq = _c(q, np.float64)
    out = np.empty((qq.shape[0], 4), np.float64)
    margin = np.empty(qq.shape[0], np.float64)
    lib().oracle_texture_sample(C.byref(tex.c), _p(qq), qq.shape[0], _p(out), _p(margin))
    return out, margin


def gbuffer_gltf(scene, prims, mats, texs, pixels=None):
    """Oracle glTF subpass: prims a list of dicts (pos, tan, bit, nrm, uvs,
    idx, material) of host arrays, mats a list of GltfMaterialC, texs a list
    of OracleTex; returns (planes (n, 4, 4), margin, triangle)."""
    g = _c(scene.globals_arr, np.float32)
    mats_ = [_c(g[16 * k:16 * k + 16], np.float32) for k in (0, 1, 2, 4)]
    cam = (vp * 4)(*[_p(m) for m in mats_])
    keep = []
    pc = []
    for pr in prims:
        arrs = [_c(pr[k], np.float32) for k in ("pos", "tan", "bit", "nrm")]
        uvs = [_c(u, np.float32) if u is not None else None for u in pr["uvs"]] + [None] * (4 - len(pr["uvs"]))
        idx = _c(pr["idx"], np.uint32)
        keep += arrs + uvs + [idx]
        pc.append(OracleGltfPrimC(*[_p(a) for a in arrs], (vp * 4)(*[_p(u) if u is not None else None for u in uvs]),
                                  _p(idx), arrs[0].shape[0], idx.size // 3, pr["material"]))
    prim_arr = (OracleGltfPrimC * len(pc))(*pc)
    from althea_amd import _lib as L
    mat_arr = (L.GltfMaterialC * len(mats))(*mats)
    tex_arr = (OracleTexC * len(texs))(*[t.c for t in texs])
    if pixels is None:
        y, x = np.mgrid[0:scene.height, 0:scene.width]
        pixels = np.stack([x.ravel(), y.ravel()], 1)
    px = _c(pixels, np.int32)
    out = np.empty((px.shape[0], 16), np.float64)
    margin = np.empty(px.shape[0], np.float64)
    tri = np.empty(px.shape[0], np.int32)
    lib().oracle_gbuffer_gltf(C.cast(cam, vp), scene.width, scene.height, prim_arr, len(pc), mat_arr, tex_arr,
                              _p(px), px.shape[0], _p(out), _p(margin), _p(tri))
    return out.reshape(-1, 4, 4), margin, tri


def _gltf_prims(prims):
    """OracleGltfPrimC array of prims (dicts of host arrays as gbuffer_gltf
    takes them) and the arrays it points into."""
    keep = []
    pc = []
    for pr in prims:
        arrs = [_c(pr[k], np.float32) if pr.get(k) is not None else None for k in ("pos", "tan", "bit", "nrm")]
        uvs = [_c(u, np.float32) if u is not None else None for u in pr["uvs"]] + [None] * (4 - len(pr["uvs"]))
        idx = _c(pr["idx"], np.uint32)
        keep += arrs + uvs + [idx]
        pc.append(OracleGltfPrimC(*[_p(a) if a is not None else None for a in arrs],
                                  (vp * 4)(*[_p(u) if u is not None else None for u in uvs]), _p(idx),
                                  arrs[0].shape[0], idx.size // 3, pr["material"]))
    return (OracleGltfPrimC * len(pc))(*pc), keep


def shadow_cubes_gltf(prims, mats, texs, lights, size: int):
    """Oracle alpha-tested shadow cubes of glTF primitives (dicts as
    gbuffer_gltf takes them; mats GltfMaterialC, texs OracleTex) for (n, 8)
    point lights: (cubes (lights, 6, size, size), alpha-decision margin of the
    same shape, 1e30 where no candidate was tested)."""
    prim_arr, keep = _gltf_prims(prims)
    from althea_amd import _lib as L
    mat_arr = (L.GltfMaterialC * len(mats))(*mats)
    tex_arr = (OracleTexC * len(texs))(*[t.c for t in texs])
    lt = _c(lights, np.float32).reshape(-1, 8)
    out = np.empty((lt.shape[0], 6, size, size), np.float64)
    margin = np.empty_like(out)
    lib().oracle_shadow_cubes_gltf(prim_arr, len(prims), mat_arr, tex_arr, _p(lt), lt.shape[0], size, _p(out),
                                   _p(margin))
    del keep
    return out, margin


def shadow_cubes(positions, indices, lights, size: int) -> np.ndarray:
    """Oracle shadow cubes (lights, 6, size, size), FP64, of a mesh
    (positions (V, 3), indices (T, 3)) for (n, 8) point lights."""
    pos = _c(positions, np.float32)
    idx = _c(indices, np.uint32)
    lt = _c(lights, np.float32).reshape(-1, 8)
    out = np.empty((lt.shape[0], 6, size, size), np.float64)
    lib().oracle_shadow_cubes(_p(pos), _p(idx), idx.size // 3, _p(lt), lt.shape[0], size, _p(out))
    return out


def ssr(scene, views: dict, pixels, gbuffer=None, shadows=None):
    """Oracle screen-space reflection of pixels[:, (x, y)]: (rgba, decision
    margin, hit step (-1 miss, -2 uncovered), hit uv, hit position)."""
    keep = []
    a = _resolve_args(scene, views, None, False, gbuffer, None, keep, shadows)
    px = _c(pixels, np.int32)
    n = px.shape[0]
    rgba = np.empty((n, 4), np.float64)
    margin = np.empty(n, np.float64)
    step = np.empty(n, np.int32)
    uv = np.empty((n, 2), np.float64)
    pos = np.empty((n, 3), np.float64)
    lib().oracle_ssr(C.byref(a), _p(px), n, _p(rgba), _p(margin), _p(step), _p(uv), _p(pos))
    return rgba, margin, step, uv, pos


def convolve_level(src, dw: int, dh: int, vertical: bool) -> np.ndarray:
    """One glossy-convolution level (dh, dw, 4) from src (sh, sw, 4)."""
    s = _c(src, np.float32)
    sh, sw = s.shape[:2]
    out = np.empty((dh, dw, 4), np.float64)
    lib().oracle_convolve_level(_p(s), sw, sh, dw, dh, 1 if vertical else 0, _p(out))
    return out


def reflection_sample(chain, width: int, height: int, mips: int, pixels, lods) -> np.ndarray:
    """textureLod of the reflection buffer (all levels, (texels, 4)) at pixel centres."""
    c = _c(chain, np.float32)
    px = _c(pixels, np.int32)
    lo = _c(lods, np.float64)
    out = np.empty((px.shape[0], 4), np.float64)
    lib().oracle_reflection_sample(_p(c), width, height, mips, _p(px), _p(lo), px.shape[0], _p(out))
    return out


def pbr_material(pos, V, N, base, refl, irr, metallic: float, roughness: float, ao: float, lut, lut_size: int,
                 lights) -> np.ndarray:
    vec = _c(np.concatenate([pos, V, N, base, refl, irr]), np.float64)
    l = _c(lut, np.float32)
    lt = _c(lights, np.float32).reshape(-1, 8)
    out = np.empty(3, np.float64)
    lib().oracle_pbr_material(_p(vec), metallic, roughness, ao, _p(l), lut_size, _p(lt), lt.shape[0], _p(out))
    return out


# --------------------------------------------------------------------- mlp
def mlp_param_count() -> int:
    return lib().oracle_mlp_param_count()


def mlp_encode(dir_rough) -> np.ndarray:
    d = _c(dir_rough, np.float64)
    out = np.empty((d.shape[0], 28), np.float64)
    lib().oracle_mlp_encode(_p(d), d.shape[0], _p(out))
    return out


def mlp_forward(params, dir_rough, bf16: bool) -> np.ndarray:
    p = _c(params, np.float32)
    d = _c(dir_rough, np.float32)
    out = np.empty((d.shape[0], 3), np.float64)
    lib().oracle_mlp_forward(_p(p), _p(d), d.shape[0], 1 if bf16 else 0, _p(out))
    return out


def mlp_loss_grad(params, dir_rough, targets, bf16: bool):
    p = _c(params, np.float32)
    d = _c(dir_rough, np.float32)
    t = _c(targets, np.float32)
    g = np.empty(mlp_param_count(), np.float64)
    loss = lib().oracle_mlp_loss_grad(_p(p), _p(d), _p(t), d.shape[0], 1 if bf16 else 0, _p(g))
    return loss, g


def mlp_loss_grad64(params, dir_rough, targets):
    """Exact-mode loss and gradient with FP64 parameters."""
    p = _c(params, np.float64)
    d = _c(dir_rough, np.float32)
    t = _c(targets, np.float32)
    g = np.empty(mlp_param_count(), np.float64)
    loss = lib().oracle_mlp_loss_grad64(_p(p), _p(d), _p(t), d.shape[0], _p(g))
    return loss, g


def adam(params, m, v, grad, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    p = _c(params, np.float64).copy()
    m = _c(m, np.float64).copy()
    v = _c(v, np.float64).copy()
    g = _c(grad, np.float64)
    lib().oracle_adam(_p(p), _p(m), _p(v), _p(g), p.size, step, lr, b1, b2, eps)
    return p, m, v


def bf16_round(x: float) -> float:
    return lib().oracle_bf16_round(x)


def max_threads() -> int:
    return lib().oracle_max_threads()
