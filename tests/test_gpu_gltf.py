"""GPU parity of the glTF subpass of the G-buffer producer
(althea_gbuffer_draw_gltf: Gltf.frag with fetchMaterial) and of the texture
mip blit (althea_texture_build_mips) against the FP64 oracle, through the C
ABI.

The scene: the synthetic heightfield as a textured glTF primitive (sRGB base
colour with an alpha channel that the 0.5 cut-off partly discards, a normal
map, a metallic-roughness map sampled NEAREST with MIRRORED_REPEAT) over a
second, plain primitive below it that shows through the discarded fragments.
Texture coordinates repeat four times across the terrain, so the mip level
varies over the frame.

Decisions (triangle edges, the alpha test, NEAREST texel boundaries and the
nearest-level rule) are compared only where the oracle's margin clears
MARGIN; the sampled values carry the FP32 texture coordinate (about 1e-6 of
the coordinate, times the texture extent, times the texel contrast).
"""
import ctypes

import numpy as np
import pytest

import althea_amd as aa
from althea_amd import _lib as L
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

SEED0 = 0x414C0000
MARGIN = 2e-4


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


# Texture extents (w, h) of the base colour, normal and metallic-roughness
# maps: powers of two (wrapped by a mask) and others (by an integer modulo).
EXTENTS = {"pow2": [(64, 64), (64, 64), (32, 32)], "odd": [(48, 40), (40, 24), (24, 20)]}


def _levels0(seed, extents=EXTENTS["pow2"]):
    rng = np.random.default_rng(seed)
    (bw, bh), (nw, nh), (mw, mh) = extents
    base = rng.integers(0, 256, (bh, bw, 4), dtype=np.uint8)
    base[..., 3] = np.where(rng.uniform(size=(bh, bw)) < 0.25, 60, 230)  # alpha below / above the cut-off
    y, x = np.mgrid[0:nh, 0:nw]
    y, x = y / nh, x / nw
    nrm = np.stack([128 + 60 * np.sin(6.28 * x), 128 + 60 * np.cos(6.28 * y), np.full_like(x, 235),
                    np.full_like(x, 255)], -1).astype(np.uint8)
    mr = rng.integers(0, 256, (mh, mw, 4), dtype=np.uint8)
    return base, nrm, mr


SPECS = [dict(srgb=1, mips=7, mag=1, min_=1, mipmode=1, wrap_u=0, wrap_v=0),
         dict(srgb=0, mips=7, mag=1, min_=1, mipmode=1, wrap_u=0, wrap_v=0),
         dict(srgb=0, mips=6, mag=0, min_=0, mipmode=0, wrap_u=1, wrap_v=1)]


@pytest.mark.parametrize("w,h,mips,srgb", [(64, 64, 7, True), (64, 32, 7, False), (37, 23, 6, True),
                                           (5, 3, 3, False)])
def test_texture_mips_match_oracle(torch, w, h, mips, srgb):
    """Each level against the oracle's blit of the GPU's own level above it.
    A value within rounding of a .5 step may round either way between FP32
    and FP64: for the 2x2 means of an 8-bit UNORM level that is exactly the
    texels whose four sources sum to 2 mod 4; otherwise only off-by-one
    differences, rarely."""
    rng = np.random.default_rng(w * 100 + h)
    lvl = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    t = aa.Texture(torch.from_numpy(lvl), mips=mips, srgb=srgb)
    torch.cuda.synchronize()
    for m in range(1, mips):
        src = t.level(m - 1).cpu().numpy()
        got = t.level(m).cpu().numpy().astype(np.int64)
        dh, dw = got.shape[:2]
        want = ol.texture_mips(src, 2, srgb)[src.shape[0] * src.shape[1]:].reshape(dh, dw, 4).astype(np.int64)
        diff = np.abs(got - want)
        assert diff.max() <= 1, (m, diff.max())
        if not srgb and src.shape[0] == 2 * dh and src.shape[1] == 2 * dw:
            s4 = (src[0::2, 0::2].astype(np.int64) + src[1::2, 0::2] + src[0::2, 1::2] + src[1::2, 1::2])
            assert np.all(diff[s4 % 4 != 2] == 0), m
        else:
            assert (diff == 1).sum() <= max(1, 1e-2 * diff.size), (m, int((diff == 1).sum()), diff.size)


def _scene(torch, seed):
    sc = aa.synth.scene(seed=seed, width=320, height=180, lights=1)
    pos, idx = aa.synth.heightfield_mesh(seed, n=40)
    nrm, _, _ = aa.synth.mesh_attributes(seed, pos, idx)
    tan = np.tile(np.float32([1, 0, 0]), (len(pos), 1))
    bit = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    uv0 = np.stack([(pos[:, 0] / 20.0 + 0.5) * 4.0, (pos[:, 2] / 20.0 + 0.5) * 4.0], 1).astype(np.float32)
    uv1 = np.stack([pos[:, 0] / 7.0, pos[:, 2] / 5.0], 1).astype(np.float32)
    floor = {"pos": np.array([[-30, -2, -30], [30, -2, 30], [30, -2, -30], [-30, -2, 30]], np.float32),
             "idx": np.array([[0, 1, 2], [0, 3, 1]], np.uint32)}
    floor.update(tan=np.tile(np.float32([1, 0, 0]), (4, 1)), bit=np.tile(np.float32([0, 0, 1]), (4, 1)),
                 nrm=np.tile(np.float32([0, 1, 0]), (4, 1)),
                 uvs=[np.array([[0, 0], [1, 1], [1, 0], [0, 1]], np.float32)], material=1)
    terrain = {"pos": pos, "idx": idx, "tan": tan, "bit": bit, "nrm": nrm, "uvs": [uv0, uv1], "material": 0}
    return sc, [terrain, floor]


def _materials():
    def mat(**kw):
        m = dict(f=(1.0, 1.0, 1.0, 1.0), mf=1.0, rf=1.0, ns=1.0, cut=0.5, bt=0, nt=1, mt=2, bu=0, mu=0)
        m.update(kw)
        return L.GltfMaterialC((ctypes.c_float * 4)(*m["f"]), m["mf"], m["rf"], m["ns"], m["cut"], m["bt"], m["nt"],
                               m["mt"], m["bu"], m["mu"])
    return [mat(f=(0.9, 0.8, 0.7, 1.0), mf=0.7, rf=0.9, ns=1.5, mu=1), mat(bt=3, nt=4, mt=4)]


def _mips(k, level0):
    """SPECS[k]'s level count, at most the texture's full chain."""
    return min(SPECS[k]["mips"], int(np.log2(max(level0.shape[:2]))) + 1)


def _gpu(torch, sc, prims, mats, levels, fp16=False):
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    texs = [aa.Texture(f(levels[k]), mips=_mips(k, levels[k]), srgb=bool(SPECS[k]["srgb"]), mag_filter=SPECS[k]["mag"],
                       min_filter=SPECS[k]["min_"], mipmap_mode=SPECS[k]["mipmode"], wrap_u=SPECS[k]["wrap_u"],
                       wrap_v=SPECS[k]["wrap_v"]) for k in range(3)]
    texs.append(aa.Texture(f(np.full((1, 1, 4), 255, np.uint8))))
    texs.append(aa.Texture(f(np.array([[[128, 128, 255, 255]]], np.uint8))))
    gp = [aa.GltfPrimitive(f(p["pos"]), f(p["tan"]), f(p["bit"]), f(p["nrm"]), [f(u) for u in p["uvs"]],
                           f(p["idx"].view(np.int32)), p["material"]) for p in prims]
    mobj = []
    for m in mats:
        o = aa.GltfMaterial.__new__(aa.GltfMaterial)
        o.c = m
        mobj.append(o)
    pas = aa.SceneToGBufferPass(sc.width, sc.height, fp16=fp16)
    g = pas.begin_gltf(gp, mobj, texs, sc.globals)
    torch.cuda.synchronize()
    planes = np.stack([p.float().cpu().numpy().reshape(-1, 4) for p in g.planes], 1).astype(np.float64)
    chains = [t.data.cpu().numpy() for t in texs]
    return planes, chains


@pytest.mark.parametrize("extents", ["pow2", "odd"])
def test_textured_heightfield_matches_oracle(torch, extents):
    sc, prims = _scene(torch, SEED0 + 51)
    levels = _levels0(SEED0 + 52, EXTENTS[extents])
    mats = _materials()
    got, chains = _gpu(torch, sc, prims, mats, levels)
    # The oracle samples the GPU's own mip chains (their blit is tested above).
    specs = SPECS + [dict(srgb=0, mips=1, mag=1, min_=1, mipmode=1, wrap_u=0, wrap_v=0)] * 2
    sizes = EXTENTS[extents] + [(1, 1), (1, 1)]
    mips = [_mips(k, levels[k]) for k in range(3)] + [1, 1]
    otex = [ol.OracleTex(chains[k], sizes[k][0], sizes[k][1], mips[k], specs[k]["srgb"], specs[k]["mag"],
                         specs[k]["min_"], specs[k]["mipmode"], specs[k]["wrap_u"], specs[k]["wrap_v"])
            for k in range(5)]
    ref, margin, tri = ol.gbuffer_gltf(sc, prims, mats, otex)
    clear = margin > MARGIN
    cov_g, cov_r = got[:, 0, 3] == 1.0, tri >= 0
    terrain_r = (tri >= 0) & (tri < prims[0]["idx"].shape[0])
    # Which primitive won: the albedo of the floor is the white texture.
    r = {"pixels": len(tri), "clear": float(clear.mean()), "terrain": int(terrain_r.sum()),
         "floor": int((cov_r & ~terrain_r).sum()),
         "coverage_disagree_clear": int(((cov_g != cov_r) & clear).sum())}
    both = cov_g & cov_r & clear
    err = np.abs(got[both] - ref[both]) / np.maximum(np.abs(ref[both]), 1.0)
    r["max_by_plane"] = [float(err[:, k].max()) for k in range(4)]
    r["mean_by_plane"] = [float(err[:, k].mean()) for k in range(4)]
    print("gltf", extents, r)
    assert r["terrain"] > 0.2 * r["pixels"] and r["floor"] > 0.05 * r["pixels"], r
    assert r["clear"] > 0.9, r
    assert r["coverage_disagree_clear"] == 0, r
    assert max(r["mean_by_plane"]) < 2e-5, r
    assert max(r["max_by_plane"]) < 5e-3, r


def test_invalid_material_texture_index(torch):
    sc, prims = _scene(torch, SEED0 + 53)
    levels = _levels0(SEED0 + 54)
    mats = _materials()
    mats[0].normal_texture = 9
    with pytest.raises(aa.AltheaError, match="names a texture outside"):
        _gpu(torch, sc, prims, mats, levels)


def test_fp16_planes_are_the_rounded_fp32_planes(torch):
    sc, prims = _scene(torch, SEED0 + 55)
    levels = _levels0(SEED0 + 56)
    mats = _materials()
    f32, _ = _gpu(torch, sc, prims, mats, levels)
    f16, _ = _gpu(torch, sc, prims, mats, levels, fp16=True)
    np.testing.assert_array_equal(f16, f32.astype(np.float32).astype(np.float16).astype(np.float64))

