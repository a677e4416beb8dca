"""Seeded synthetic inputs (environment maps, fBm-heightfield G-buffers,
point lights, SSAO kernels) from ``libalthea_synth.so``.  Host-side numpy
arrays; the same bytes feed the GPU path and the CPU oracle."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib

SYNTH_PATH = Path(__file__).resolve().parent / "libalthea_synth.so"
_synth = None


def _s() -> C.CDLL:
    global _synth
    if _synth is None:
        if not SYNTH_PATH.exists():
            raise RuntimeError(f"{SYNTH_PATH} is missing: run __graft_entry__.build() first")
        s = C.CDLL(str(SYNTH_PATH))
        u64, i32, d, vp = C.c_uint64, C.c_int, C.c_double, C.c_void_p
        s.althea_synth_env.argtypes = [u64, i32, i32, d, i32, d, d, d, vp, vp]
        s.althea_synth_camera.argtypes = [vp, vp, d, i32, i32, i32, C.c_float, vp]
        s.althea_synth_gbuffer.argtypes = [u64, vp, i32, i32, d, vp, vp, vp, vp]
        s.althea_synth_lights.argtypes = [u64, i32, d, d, d, d, d, vp]
        s.althea_synth_heightfield_mesh.argtypes = [u64, d, i32, d, vp, vp]
        s.althea_synth_ssao.argtypes = [u64, i32, vp, vp]
        s.althea_synth_dir_rough.argtypes = [u64, C.c_int64, d, d, vp]
        _synth = s
    return _synth


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def env(seed: int, width: int, height: int, sigma: float = 1.5, suns: int = 3, peak_min: float = 1e3,
        peak_max: float = 1e3, sun_sigma_deg: float = 2.0, return_suns: bool = False):
    """Equirect RGBA32F map (height, width, 4): log-normal(0, sigma) per channel
    plus `suns` Gaussian lobes of angular sigma `sun_sigma_deg`."""
    out = np.empty((height, width, 4), np.float32)
    suns_out = np.zeros((max(suns, 1), 4), np.float32)
    _s().althea_synth_env(seed, width, height, sigma, suns, peak_min, peak_max, sun_sigma_deg, _p(out),
                          _p(suns_out))
    return (out, suns_out[:suns]) if return_suns else out


def dir_rough(seed: int, n: int, rmin: float = 0.0, rmax: float = 1.0) -> np.ndarray:
    out = np.empty((n, 4), np.float32)
    _s().althea_synth_dir_rough(seed, n, rmin, rmax, _p(out))
    return out


@dataclass
class Scene:
    width: int
    height: int
    position: np.ndarray
    normal: np.ndarray
    albedo: np.ndarray
    material: np.ndarray  # metallic, roughness, occlusion
    lights: np.ndarray      # (n, 8)
    globals_arr: np.ndarray  # 104 floats, GlobalUniforms layout
    ssao_kernel: np.ndarray  # (K, 4)
    ssao_noise: np.ndarray   # (16, 4)
    ssao_radius: float = 0.5
    ssao_bias: float = 0.025
    ssao_enabled: bool = True

    @property
    def globals(self) -> _lib.GlobalsC:
        g = _lib.GlobalsC()
        C.memmove(C.addressof(g), self.globals_arr.ctypes.data, C.sizeof(g))
        g.light_count = int(self.lights.shape[0])
        return g

    @property
    def ssao(self) -> _lib.SsaoDescC:
        d = _lib.SsaoDescC()
        d.enabled = 1 if self.ssao_enabled else 0
        d.samples = int(self.ssao_kernel.shape[0])
        d.radius = self.ssao_radius
        d.bias = self.ssao_bias
        for k in range(self.ssao_kernel.shape[0]):
            for c in range(4):
                d.kernel[k][c] = float(self.ssao_kernel[k, c])
        for k in range(16):
            for c in range(4):
                d.noise[k][c] = float(self.ssao_noise[k, c])
        return d

    def mat(self, index: int) -> np.ndarray:
        """Column-major mat4 `index` of GlobalUniforms (0 projection, 1
        inverse projection, 2 view, 4 inverse view)."""
        return self.globals_arr[16 * index:16 * index + 16]

    def gbuffer_tensors(self, device, fp16: bool = False):
        import torch

        dt = torch.float16 if fp16 else torch.float32
        return [torch.from_numpy(a).to(device=device, dtype=dt).contiguous()
                for a in (self.position, self.normal, self.albedo, self.material)]

    def lights_tensor(self, device):
        import torch

        return torch.from_numpy(self.lights).to(device)

    @property
    def pixels(self) -> int:
        return self.width * self.height


def scene(seed: int, width: int, height: int, lights: int, ssao_samples: int = 16, exposure: float = 0.5,
          eye=(0.0, 3.5, 12.0), center=(0.0, 0.0, 0.0), fovy_deg: float = 60.0, amplitude: float = 1.5) -> Scene:
    """Pinhole camera (60 deg vertical FOV) over a seeded 4-octave fBm
    heightfield (amplitude 0: the plane y = 0), lights uniform in the scene
    box, 16-sample SSAO kernel."""
    s = _s()
    g = np.zeros(104, np.float32)
    eye_a = np.array(eye, np.float64)
    cen_a = np.array(center, np.float64)
    s.althea_synth_camera(_p(eye_a), _p(cen_a), fovy_deg, width, height, lights, exposure, _p(g))
    planes = [np.empty((height, width, 4), np.float32) for _ in range(4)]
    s.althea_synth_gbuffer(seed, _p(g), width, height, amplitude, *[_p(a) for a in planes])
    lt = np.empty((lights, 8), np.float32)
    s.althea_synth_lights(seed, lights, 10.0, -1.5, 3.5, 1.0, 50.0, _p(lt))
    k = np.empty((ssao_samples, 4), np.float32)
    n = np.empty((16, 4), np.float32)
    s.althea_synth_ssao(seed, ssao_samples, _p(k), _p(n))
    return Scene(width, height, *planes, lt, g, k, n)


def mesh_attributes(seed: int, positions: np.ndarray, indices: np.ndarray):
    """Per-vertex inputs of the pass-through G-buffer shader for a mesh:
    normals (the normalised sum of the adjacent faces' (b - a) x (c - a)),
    seeded colours in [0.05, 0.95] and metallic / roughness / occlusion in
    [0, 1] x [0.05, 1] x [0.5, 1], float32 (v, 3) each."""
    p = positions.astype(np.float64)
    t = indices.astype(np.int64)
    fn = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, t[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    rng = np.random.default_rng(seed)
    col = rng.uniform(0.05, 0.95, p.shape)
    mro = np.stack([rng.uniform(0.0, 1.0, len(p)), rng.uniform(0.05, 1.0, len(p)), rng.uniform(0.5, 1.0, len(p))], 1)
    return vn.astype(np.float32), col.astype(np.float32), mro.astype(np.float32)


def heightfield_mesh(seed: int, n: int = 128, half: float = 10.0, amplitude: float = 1.5):
    """The scene's heightfield as a triangle mesh: (positions (V, 3) float32,
    indices (T, 3) uint32), an n x n grid over [-half, half]^2, two triangles
    per cell, vertices at the terrain height of scene(seed, ...)."""
    pos = np.empty(((n + 1) * (n + 1), 3), np.float32)
    idx = np.empty((2 * n * n, 3), np.uint32)
    _s().althea_synth_heightfield_mesh(seed, amplitude, n, half, _p(pos), _p(idx))
    return pos, idx
