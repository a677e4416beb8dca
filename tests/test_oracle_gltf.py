"""CPU pins of the oracle's glTF subpass: texture sampling as Vulkan defines
it, the mip blit, and fetchMaterial / Gltf.frag (no GPU).

Sampling (Src/Sampler.cpp, the Vulkan texel filtering rules):
- a constant texture samples to its colour at any coordinate, footprint,
  filter and wrap mode;
- REPEAT is periodic in 1, MIRRORED_REPEAT symmetric about 0, CLAMP_TO_EDGE
  constant beyond the edge;
- bilinear filtering of a linear ramp gives the ramp (away from the wrap);
- the mip level: a footprint of 2 texels per pixel (lambda = 1) reads level 1
  exactly under LINEAR mipmapping, lambda = 1.5 the mean of levels 1 and 2;
  NEAREST mipmapping picks level ceil(lambda + 0.5) - 1;
- sRGB texels decode by the sRGB transfer function (0 -> 0, 255 -> 1, 188 ->
  0.5029).
Mip blit (Image::generateMipMaps): power-of-two levels are the rounded 2x2
means (of the decoded values for sRGB), exact ties rounding either way.
fetchMaterial: a flat normal map (128, 128, 255) tilts N by 1/255 along T and
B as the shader's arithmetic gives; base colour and metallic-roughness are
texture x factor with metallic from blue and roughness from green; fragments
whose alpha is under the cut-off are discarded.
"""
import ctypes

import numpy as np
import pytest

from althea_amd import _lib as L
from althea_amd import synth
from tests import oracle_lib as ol


def _tex(level0, mips=1, srgb=0, **kw):
    h, w = level0.shape[:2]
    chain = ol.texture_mips(level0, mips, bool(srgb))
    return ol.OracleTex(chain, w, h, mips, srgb, **kw)


def _q(u, v, dudx=0.0, dvdx=0.0, dudy=0.0, dvdy=0.0):
    u = np.atleast_1d(np.asarray(u, np.float64))
    v = np.broadcast_to(np.asarray(v, np.float64), u.shape)
    z = np.zeros_like(u)
    return np.stack([u, v, z + dudx, z + dvdx, z + dudy, z + dvdy], 1)


@pytest.mark.parametrize("wrap", [0, 1, 2])
@pytest.mark.parametrize("filt", [0, 1])
def test_constant_texture(wrap, filt):
    lvl = np.tile(np.array([10, 200, 77, 128], np.uint8), (8, 16, 1))
    t = _tex(lvl, mips=5, mag=filt, min_=filt, mipmode=filt, wrap_u=wrap, wrap_v=wrap)
    rng = np.random.default_rng(1)
    q = np.concatenate([rng.uniform(-3, 3, (300, 2)), rng.uniform(-0.3, 0.3, (300, 4))], 1)
    out, _ = ol.texture_sample(t, q)
    np.testing.assert_allclose(out, np.tile(np.array([10, 200, 77, 128]) / 255.0, (300, 1)), atol=1e-12)


def _random_tex(seed, w=16, h=8, **kw):
    rng = np.random.default_rng(seed)
    return _tex(rng.integers(0, 256, (h, w, 4), dtype=np.uint8), **kw)


def test_wrap_modes():
    rng = np.random.default_rng(2)
    u, v = rng.uniform(0.05, 0.95, 200), rng.uniform(0.05, 0.95, 200)
    rep = _random_tex(3, wrap_u=0, wrap_v=0)
    a, _ = ol.texture_sample(rep, _q(u, v))
    b, _ = ol.texture_sample(rep, _q(u + 2.0, v - 1.0))
    np.testing.assert_allclose(a, b, atol=1e-12)
    mir = _random_tex(3, wrap_u=1, wrap_v=1)
    a, _ = ol.texture_sample(mir, _q(u, v))
    b, _ = ol.texture_sample(mir, _q(-u, -v))
    np.testing.assert_allclose(a, b, atol=1e-12)
    clamp = _random_tex(3, wrap_u=2, wrap_v=2, mag=0, min_=0)
    a, _ = ol.texture_sample(clamp, _q(np.full(5, 0.99), np.full(5, 0.5)))
    b, _ = ol.texture_sample(clamp, _q(np.linspace(1.2, 7.0, 5), np.full(5, 0.5)))
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_bilinear_ramp():
    w, h = 32, 4
    lvl = np.zeros((h, w, 4), np.uint8)
    lvl[..., 0] = (np.arange(w) * 8)[None, :]
    lvl[..., 3] = 255
    t = _tex(lvl, wrap_u=2, wrap_v=2)
    u = np.linspace(1.0 / w, 1.0 - 1.0 / w, 50)
    out, _ = ol.texture_sample(t, _q(u, 0.5))
    want = (u * w - 0.5) * 8 / 255.0
    np.testing.assert_allclose(out[:, 0], want, atol=1e-12)


def test_mip_level_selection():
    w = h = 16
    rng = np.random.default_rng(4)
    lvl = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    u, v = rng.uniform(0.1, 0.9, 64), rng.uniform(0.1, 0.9, 64)
    lin = _tex(lvl, mips=5)
    near = _tex(lvl, mips=5, mipmode=0)
    one_level = lambda m: ol.OracleTex(lin.texels[ol_offset(w, h, m):], max(1, w >> m), max(1, h >> m), 1)

    def at(t, lam):
        return ol.texture_sample(t, _q(u, v, dudx=2.0 ** lam / w))[0]

    l1 = ol.texture_sample(one_level(1), _q(u, v))[0]
    l2 = ol.texture_sample(one_level(2), _q(u, v))[0]
    np.testing.assert_allclose(at(lin, 1.0), l1, atol=1e-12)
    np.testing.assert_allclose(at(lin, 1.5), 0.5 * (l1 + l2), atol=1e-12)
    np.testing.assert_allclose(at(near, 1.4), l1, atol=1e-12)
    np.testing.assert_allclose(at(near, 1.6), l2, atol=1e-12)
    # Magnification (lambda <= 0) reads level 0 with the mag filter.
    mag_nearest = _tex(lvl, mips=5, mag=0)
    l0n = ol.texture_sample(_tex(lvl, mag=0, min_=0), _q(u, v))[0]
    np.testing.assert_allclose(ol.texture_sample(mag_nearest, _q(u, v, dudx=0.5 / w))[0], l0n, atol=1e-12)


def ol_offset(w, h, m):
    return sum(max(1, w >> k) * max(1, h >> k) for k in range(m))


def test_srgb_decode_points():
    lvl = np.array([[[0, 255, 188, 188]]], np.uint8)
    t = _tex(lvl, srgb=1)
    out, _ = ol.texture_sample(t, _q(0.5, 0.5))
    c = 188 / 255.0
    np.testing.assert_allclose(out[0], [0.0, 1.0, ((c + 0.055) / 1.055) ** 2.4, c], atol=1e-12)
    assert abs(out[0, 2] - 0.5029) < 1e-4


@pytest.mark.parametrize("srgb", [0, 1])
def test_mip_blit_power_of_two_is_the_rounded_mean(srgb):
    rng = np.random.default_rng(5)
    lvl = rng.integers(0, 256, (8, 16, 4), dtype=np.uint8)
    chain = ol.texture_mips(lvl, 3, bool(srgb))
    l1 = chain[16 * 8:16 * 8 + 8 * 4].reshape(4, 8, 4)
    x = lvl.astype(np.float64) / 255.0
    if srgb:
        x[..., :3] = np.where(x[..., :3] <= 0.04045, x[..., :3] / 12.92, ((x[..., :3] + 0.055) / 1.055) ** 2.4)
    m = 0.25 * (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2])
    if srgb:
        m[..., :3] = np.where(m[..., :3] <= 0.0031308, 12.92 * m[..., :3], 1.055 * m[..., :3] ** (1 / 2.4) - 0.055)
    want = np.floor(m * 255.0 + 0.5)
    # Exact ties (the mean of four 8-bit values lands on .5) may round either
    # way; every other texel is the rounded mean.
    tie = np.abs(m * 255.0 - np.floor(m * 255.0) - 0.5) < 1e-9
    diff = l1.astype(np.int64) - want.astype(np.int64)
    assert np.all(diff[~tie] == 0) and np.all(np.abs(diff[tie]) <= 1)


# ------------------------------------------------------------ fetchMaterial
def _quad_prim(material, z=-5.0, s=3.0):
    """A square at z facing +z (towards a camera at the origin looking down
    -z), TBN = (x, y, z), uv = the square's (x, y) mapped to [0, 1]."""
    pos = np.array([[-s, -s, z], [s, -s, z], [s, s, z], [-s, s, z]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    rep = lambda v: np.tile(np.array(v, np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return {"pos": pos, "tan": rep([1, 0, 0]), "bit": rep([0, 1, 0]), "nrm": rep([0, 0, 1]), "uvs": [uv], "idx": idx,
            "material": material}


@pytest.fixture(scope="module")
def cam():
    return synth.scene(seed=7, width=48, height=32, lights=1, eye=(0.0, 0.0, 0.0), center=(0.0, 0.0, -1.0))


def _solid(rgba):
    return _tex(np.array([[rgba]], np.uint8))


def _mat(**kw):
    m = dict(base_color_factor=(1.0, 1.0, 1.0, 1.0), metallic_factor=1.0, roughness_factor=1.0, normal_scale=1.0,
             alpha_cutoff=0.5, base_texture=0, normal_texture=1, metallic_roughness_texture=2, base_uv=0,
             metallic_roughness_uv=0)
    m.update(kw)
    return L.GltfMaterialC((ctypes.c_float * 4)(*m["base_color_factor"]), m["metallic_factor"], m["roughness_factor"],
                           m["normal_scale"], m["alpha_cutoff"], m["base_texture"], m["normal_texture"],
                           m["metallic_roughness_texture"], m["base_uv"], m["metallic_roughness_uv"])


def test_flat_normal_map_factors_and_channels(cam):
    texs = [_solid([255, 255, 255, 255]), _solid([128, 128, 255, 255]), _solid([10, 200, 50, 255])]
    mat = _mat(base_color_factor=(0.8, 0.4, 0.2, 1.0), metallic_factor=0.5, roughness_factor=0.9, normal_scale=2.0)
    out, margin, tri = ol.gbuffer_gltf(cam, [_quad_prim(0)], [mat], texs)
    sel = tri >= 0
    assert sel.mean() > 0.5
    np.testing.assert_allclose(out[sel, 2], np.tile(np.float32([0.8, 0.4, 0.2, 1.0]), (sel.sum(), 1)), atol=1e-7)
    e = (2 * 128 / 255.0 - 1) * 2.0
    n = np.array([e, e, 1.0]) / np.linalg.norm([e, e, 1.0])
    np.testing.assert_allclose(out[sel, 1, :3], np.tile(n, (sel.sum(), 1)), atol=1e-12)
    np.testing.assert_allclose(out[sel, 3, :2], np.tile([50 / 255 * 0.5, 200 / 255 * np.float32(0.9)], (sel.sum(), 1)),
                               atol=1e-7)
    np.testing.assert_array_equal(out[sel, 3, 2], 0.0)
    np.testing.assert_allclose(out[sel, 0, 2], -5.0, atol=1e-9)


@pytest.mark.parametrize("alpha,covered", [(0.4, False), (0.6, True)])
def test_alpha_cutoff(cam, alpha, covered):
    texs = [_solid([255, 255, 255, 255]), _solid([128, 128, 255, 255]), _solid([255, 255, 255, 255])]
    mat = _mat(base_color_factor=(1.0, 1.0, 1.0, alpha))
    # A second square behind shows through where the first is discarded.
    back = _quad_prim(1, z=-8.0, s=8.0)
    _, _, tri = ol.gbuffer_gltf(cam, [_quad_prim(0), back], [mat, _mat()], texs)
    assert ((tri == 0) | (tri == 1)).any() == covered
    assert (tri >= 2).any()
