"""GPU parity of the point-light shadow path against the FP64 oracle, through
the C ABI:

- althea_shadow_cubes (Src/PointLight.cpp:235-282, ShadowMapBindless.*): the
  cubes of the synthetic heightfield mesh for several lights, texel by texel;
  the clear value without triangles; culled back faces; odd cube sizes; a
  large occluder whose boxes span whole faces (the wave-cooperative walk);
- the one-texel border of every stored face: the seamless filter's reads
  beyond the edge (cube_fold onto the adjacent face, corners the mean of
  three), bit for bit;
- althea_deferred_resolve_shadowed (PBRMaterial.glsl:119-140): against the
  oracle resolve reading the same cubes, with and without a reflection
  buffer; cubes of 1 give the plain resolve.

A texel-centre ray that passes within rounding of a triangle edge may be
claimed by the triangle in FP32 and not in FP64 (or by its neighbour): such
texels are counted, and must stay a small fraction.
"""
import numpy as np
import pytest

import althea_amd as aa
from tests import oracle_lib as ol
from tests.parity import RADIANCE_FLOOR, rel_err, summary

pytestmark = pytest.mark.gpu

SEED0 = 0x414C0000


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def ibl(torch):
    env = aa.synth.env(seed=SEED0 + 100, width=512, height=256, sigma=0.0, suns=4, peak_min=100.0,
                       peak_max=100.0, sun_sigma_deg=8.0)
    desc = aa.IBLDesc(prefiltered_size=128, prefiltered_mips=5)
    res = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    torch.cuda.synchronize()
    return res, ol.Env(env)


def _mesh(torch, pos, idx):
    return (torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda(),
            torch.from_numpy(np.ascontiguousarray(idx, np.uint32).view(np.int32)).cuda())


def _gpu_cubes(torch, pos, idx, lights, size, scheduled=True):
    coll = aa.PointLightCollection.from_lights(torch.from_numpy(lights).cuda(), shadow_size=size)
    coll.drawShadowMaps(*_mesh(torch, pos, idx), scheduled=scheduled)
    torch.cuda.synchronize()
    return coll, coll.getShadowMapImage().cpu().numpy()


def _box(size):
    rng = np.random.default_rng(size)
    lights = np.concatenate([rng.uniform(-0.5, 0.5, (3, 3)), np.zeros((3, 1)), np.full((3, 3), 4.0),
                             np.zeros((3, 1))], 1).astype(np.float32)
    corners = np.array([[x, y, z] for x in (-2.0, 2.5) for y in (-1.5, 3.0) for z in (-2.2, 1.8)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    # (a, c, b) winding: every normal points into the box.
    idx = np.array([t for a, b, c, d in quads for t in ((a, c, b), (a, d, c))], np.uint32)
    return corners, idx, lights


def _compare_cubes(got, ref):
    got = got.astype(np.float64)
    hit_g, hit_r = got < 1.0, ref < 1.0
    both = hit_g & hit_r
    err = np.abs(got[both] - ref[both]) / ref[both]
    return {"texels": got.size, "hits": int(hit_r.sum()), "hit_disagree": int((hit_g != hit_r).sum()),
            "depth_off": int((err > 1e-4).sum()), "err": summary(err) if both.any() else None}


@pytest.mark.parametrize("size", [128, 37])
def test_heightfield_cubes_match_oracle(torch, size):
    sc = aa.synth.scene(seed=SEED0 + 31, width=8, height=8, lights=6)
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 31, n=40)
    _, got = _gpu_cubes(torch, pos, idx, sc.lights, size)
    ref = ol.shadow_cubes(pos, idx, sc.lights, size)
    r = _compare_cubes(got, ref)
    print("shadow cubes", size, r)
    assert r["hits"] > 0.2 * r["texels"], r
    # Texels whose ray grazes an edge: a small fraction of the hits.
    assert r["hit_disagree"] <= 2e-3 * r["hits"] + 4, r
    # Elsewhere the FP32 depth is the FP64 one to FP32 arithmetic: the
    # ray-plane division a.n / d.n loses digits only for grazing triangles;
    # edge texels that both claim but from different triangles aside.
    assert r["depth_off"] <= 2e-3 * r["hits"] + 4, r
    assert r["err"]["mean"] < 1e-6 and r["err"]["p99"] < 5e-6, r


def test_no_triangles_and_back_faces(torch):
    lights = np.array([[0.0, 3.0, 0.0, 0.0, 5.0, 5.0, 5.0, 0.0]], np.float32)
    empty = (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    _, got = _gpu_cubes(torch, *empty, lights, 16)
    assert np.all(got == 1.0)
    # A floor facing down, away from the light above it: culled.
    pos = np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)  # normal (b - a) x (c - a) = -y
    _, got = _gpu_cubes(torch, pos, idx, lights, 16)
    assert np.all(got == 1.0)
    _, got = _gpu_cubes(torch, pos, idx[:, ::-1].copy(), lights, 16)
    ref = ol.shadow_cubes(pos, idx[:, ::-1].copy(), lights, 16)
    r = _compare_cubes(got, ref)
    assert r["hits"] > 0 and r["hit_disagree"] == 0 and r["err"]["max"] < 1e-6, r


@pytest.mark.parametrize("size", [1, 2, 256])
def test_large_occluders_span_whole_faces(torch, size):
    """A box of six inward-facing squares around the light: every face of
    every cube is covered, through the wave-cooperative walk (boxes larger
    than a lane's share) and for triangles that cross a face's plane."""
    corners, idx, lights = _box(size)
    _, got = _gpu_cubes(torch, corners, idx, lights, size)
    ref = ol.shadow_cubes(corners, idx, lights, size)
    assert np.all(ref < 1.0), "the box must enclose every light"
    r = _compare_cubes(got, ref)
    print("box", size, r)
    assert r["hit_disagree"] <= 2e-3 * r["hits"] + 2, r
    assert r["depth_off"] <= 2e-3 * r["hits"] + 2, r


@pytest.mark.parametrize("case", ["heightfield", "box256", "box3", "mixed", "full"])
def test_scheduled_path_is_bit_identical(torch, case):
    """althea_shadow_cubes_ws (texels dealt out evenly over the GPU) writes
    the cubes of althea_shadow_cubes bit for bit: the same texel tests and an
    order-free atomic minimum.  Cases: many small boxes, whole-face boxes,
    boxes of one or two texels, both mixed with culled and degenerate
    triangles (zero slots), and the full benchmark size (64 lights, 256^2
    cubes, the 128 x 128-cell heightfield: 2 097 152 light-triangle pairs)."""
    if case == "full":
        sc = aa.synth.scene(seed=SEED0 + 37, width=8, height=8, lights=64)
        pos, idx = aa.synth.heightfield_mesh(SEED0 + 37, n=128)
        lights, size = sc.lights, 256
    elif case == "heightfield":
        sc = aa.synth.scene(seed=SEED0 + 35, width=8, height=8, lights=9)
        pos, idx = aa.synth.heightfield_mesh(SEED0 + 35, n=64)
        lights, size = sc.lights, 96
    elif case.startswith("box"):
        size = int(case[3:])
        pos, idx, lights = _box(size)
    else:
        sc = aa.synth.scene(seed=SEED0 + 36, width=8, height=8, lights=5)
        hp, hi = aa.synth.heightfield_mesh(SEED0 + 36, n=24)
        bp, bi, _ = _box(7)
        pos = np.concatenate([hp, bp * 6.0, np.float32([[0, 1, 0], [1, 1, 0], [2, 1, 0]])])
        extra = np.array([[len(pos) - 3, len(pos) - 2, len(pos) - 1], [4, 4, 4]], np.uint32)
        idx = np.concatenate([hi, extra, bi + len(hp), hi[::-1, ::-1]])
        lights, size = sc.lights, 61
    _, a = _gpu_cubes(torch, pos, idx, lights, size, scheduled=False)
    coll, b = _gpu_cubes(torch, pos, idx, lights, size, scheduled=True)
    assert (a < 1.0).any()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    # A second, smaller draw through the same (kept) workspace.
    coll.drawShadowMaps(*_mesh(torch, pos, idx[: len(idx) // 3]))
    _, c = _gpu_cubes(torch, pos, idx[: len(idx) // 3], lights, size, scheduled=False)
    np.testing.assert_array_equal(coll.getShadowMapImage().cpu().numpy().view(np.uint32), c.view(np.uint32))


def _expected_border(faces, S):
    """The bordered storage of one cube's faces (6, S, S) float32: each border
    texel is the seamless filter's tap beyond the edge, in FP32 as the GPU
    sums it."""
    P = S + 2
    out = np.zeros((6, P, P), np.float32)
    out[:, 1:-1, 1:-1] = faces
    third = np.float32(1.0 / 3.0)
    ring = [(i, -1) for i in range(-1, S + 1)] + [(i, S) for i in range(-1, S + 1)]
    ring += [(-1, j) for j in range(S)] + [(S, j) for j in range(S)]
    for f in range(6):
        for i, j in ring:
            in_i, in_j = 0 <= i < S, 0 <= j < S
            if in_i or in_j:
                g, a, b = ol.cube_fold(f, i, j, S)
                v = faces[g, b, a]
            else:
                ic, jc = min(max(i, 0), S - 1), min(max(j, 0), S - 1)
                g1, a1, b1 = ol.cube_fold(f, i, jc, S)
                g2, a2, b2 = ol.cube_fold(f, ic, j, S)
                v = ((faces[f, jc, ic] + faces[g1, b1, a1]) + faces[g2, b2, a2]) * third
            out[f, j + 1, i + 1] = v
    return out


@pytest.mark.parametrize("size", [1, 2, 37])
def test_face_borders_hold_the_seamless_taps(torch, size):
    sc = aa.synth.scene(seed=SEED0 + 35, width=8, height=8, lights=3)
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 35, n=24)
    coll, faces = _gpu_cubes(torch, pos, idx, sc.lights, size)
    stored = coll.shadow_storage.cpu().numpy()
    assert stored.shape == (3, 6, size + 2, size + 2)
    for l in range(3):
        np.testing.assert_array_equal(stored[l], _expected_border(faces[l], size))


def _shadow_scene(torch, W, H, lights, seed):
    sc = aa.synth.scene(seed=seed, width=W, height=H, lights=lights)
    pos, idx = aa.synth.heightfield_mesh(seed, n=96)
    coll, cubes = _gpu_cubes(torch, pos, idx, sc.lights, 128)
    return sc, coll, cubes


def test_all_ones_cubes_are_the_plain_resolve(torch, ibl):
    ibl_res, _ = ibl
    sc = aa.synth.scene(seed=SEED0 + 32, width=320, height=180, lights=8)
    g = sc.gbuffer_tensors("cuda")
    lights = sc.lights_tensor("cuda")
    coll = aa.PointLightCollection.from_lights(lights, shadow_size=8)  # cleared to 1
    plain = aa.DeferredRendering.resolve(g, ibl_res, lights, sc.globals, sc.ssao)
    got = aa.DeferredRendering.resolve(g, ibl_res, lights, sc.globals, sc.ssao, shadows=coll)
    torch.cuda.synchronize()
    np.testing.assert_allclose(got.cpu().numpy(), plain.cpu().numpy(), rtol=2e-6, atol=0)


@pytest.mark.parametrize("fp16_out", [False, True])
def test_resolve_shadowed_matches_oracle(torch, ibl, fp16_out):
    ibl_res, oenv = ibl
    sc, coll, cubes = _shadow_scene(torch, 480, 270, 8, SEED0 + 33)
    g = sc.gbuffer_tensors("cuda")
    out = aa.DeferredRendering.resolve(g, ibl_res, coll, sc.globals, sc.ssao, out_fp16=fp16_out)
    torch.cuda.synchronize()
    pix = np.stack(np.meshgrid(np.arange(sc.width), np.arange(sc.height)), -1).reshape(-1, 2)[::5]
    pix = pix.astype(np.int32)
    ref, _, margin = ol.resolve(sc, ibl_res.views_numpy(), oenv, pix, shadows=cubes)
    plain, _, _ = ol.resolve(sc, ibl_res.views_numpy(), oenv, pix)
    got = out.float().cpu().numpy()[pix[:, 1] * sc.width + pix[:, 0], :3].astype(np.float64)
    clear = margin > 2e-3
    err = rel_err(got, ref, RADIANCE_FLOOR).max(1)
    s = summary(err[clear])
    shadowed = (np.abs(ref - plain).max(1) > 1e-3 * np.abs(plain).max(1)) & clear
    print("resolve shadowed", s, "clear", clear.mean(), "shadowed", shadowed.mean())
    assert clear.mean() > 0.6
    assert shadowed.mean() > 0.05
    if fp16_out:
        assert s["max"] < 2.0 ** -10, s
    else:
        assert s["mean"] < 1e-5 and s["max"] < 1e-3, s


def test_resolve_shadowed_with_reflection_buffer(torch, ibl):
    ibl_res, oenv = ibl
    sc, coll, cubes = _shadow_scene(torch, 320, 180, 6, SEED0 + 34)
    g = sc.gbuffer_tensors("cuda")
    lights = sc.lights_tensor("cuda")
    refl = aa.ReflectionBuffer(sc.width, sc.height, mips=5, fp16=True)
    aa.ScreenSpaceReflection.captureReflection(g, ibl_res, lights, sc.globals, refl)
    aa.ScreenSpaceReflection.convolveReflectionBuffer(refl)
    out = aa.DeferredRendering.resolve(g, ibl_res, lights, sc.globals, sc.ssao, reflection=refl, shadows=coll)
    torch.cuda.synchronize()
    chain = refl.data.float().cpu().numpy()
    pix = np.stack(np.meshgrid(np.arange(sc.width), np.arange(sc.height)), -1).reshape(-1, 2)[::3]
    pix = pix.astype(np.int32)
    ref, _, margin = ol.resolve(sc, ibl_res.views_numpy(), oenv, pix, reflection=(chain, refl.mips),
                                shadows=cubes)
    got = out.cpu().numpy()[pix[:, 1] * sc.width + pix[:, 0], :3].astype(np.float64)
    clear = margin > 2e-3
    s = summary(rel_err(got, ref, RADIANCE_FLOOR).max(1)[clear])
    print("resolve shadowed + reflection", s, "clear", clear.mean())
    assert clear.mean() > 0.6
    assert s["mean"] < 1e-5 and s["max"] < 1e-3, s


def test_ssr_capture_shadowed_matches_oracle(torch, ibl):
    """The SSR hit shading with the shadow test (pbrMaterial reads the cubes
    for SSR.frag too): against the oracle with the same cubes; cubes of 1 give
    the unshadowed capture."""
    ibl_res, _ = ibl
    sc, coll, cubes = _shadow_scene(torch, 320, 180, 8, SEED0 + 36)
    g = sc.gbuffer_tensors("cuda")
    refl = aa.ReflectionBuffer(sc.width, sc.height, mips=1, fp16=False)
    aa.ScreenSpaceReflection.captureReflection(g, ibl_res, coll, sc.globals, refl)
    torch.cuda.synchronize()
    got = refl.level(0).cpu().numpy().reshape(-1, 4).astype(np.float64)
    pix = np.stack(np.meshgrid(np.arange(sc.width), np.arange(sc.height)), -1).reshape(-1, 2).astype(np.int32)
    rgba, margin, step, _, _ = ol.ssr(sc, ibl_res.views_numpy(), pix, shadows=cubes)
    plain, _, _, _, _ = ol.ssr(sc, ibl_res.views_numpy(), pix)
    hit = (step >= 0) & (got[:, 3] == 1.0) & (margin > 2e-5)
    err = rel_err(got[hit, :3], rgba[hit, :3], RADIANCE_FLOOR).max(1)
    shadowed = np.abs(rgba[hit, :3] - plain[hit, :3]).max(1) > 1e-6
    s = summary(err)
    print("ssr shadowed", s, "hits", int(hit.sum()), "shadowed hits", int(shadowed.sum()))
    assert hit.sum() > 300 and shadowed.sum() > 30
    assert s["mean"] < 5e-5 and s["max"] < 2e-3, s
    ones = aa.PointLightCollection.from_lights(sc.lights_tensor("cuda"), shadow_size=8)
    r1 = aa.ReflectionBuffer(sc.width, sc.height, mips=1, fp16=False)
    r0 = aa.ReflectionBuffer(sc.width, sc.height, mips=1, fp16=False)
    aa.ScreenSpaceReflection.captureReflection(g, ibl_res, ones, sc.globals, r1)
    aa.ScreenSpaceReflection.captureReflection(g, ibl_res, sc.lights_tensor("cuda"), sc.globals, r0)
    torch.cuda.synchronize()
    np.testing.assert_allclose(r1.level(0).cpu().numpy(), r0.level(0).cpu().numpy(), rtol=2e-6, atol=0)



# ------------------------------------------------- glTF (alpha-tested) pass
def _gltf_shadow_scene(seed, n=40):
    """The heightfield as a glTF primitive whose base colour alpha cuts about
    a quarter of its texture (sRGB, LINEAR with 7 mips, coordinates
    repeating four times across the terrain, so the mip level varies over
    the cubes), over an opaque floor primitive (a 1 x 1 white texture)."""
    pos, idx = aa.synth.heightfield_mesh(seed, n=n)
    uv = np.stack([(pos[:, 0] / 20.0 + 0.5) * 4.0, (pos[:, 2] / 20.0 + 0.5) * 4.0], 1).astype(np.float32)
    floor_pos = np.array([[-30, -3, -30], [30, -3, 30], [30, -3, -30], [-30, -3, 30]], np.float32)
    floor_idx = np.array([[0, 1, 2], [0, 3, 1]], np.uint32)
    floor_uv = np.array([[0, 0], [1, 1], [1, 0], [0, 1]], np.float32)
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    base[..., 3] = np.where(rng.uniform(size=(64, 64)) < 0.25, 60, 230)
    return [(pos, idx, uv, 0), (floor_pos, floor_idx, floor_uv, 1)], base


def _gltf_shadow_gpu(torch, prims, base, lights, size, opaque=False):
    import ctypes

    f = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    lvl = base.copy()
    if opaque:
        lvl[..., 3] = 255
    texs = [aa.Texture(f(lvl), mips=7, srgb=True), aa.Texture(f(np.full((1, 1, 4), 255, np.uint8)))]
    mats = [aa.GltfMaterial(0, 0, 0), aa.GltfMaterial(1, 1, 1)]
    gp = [aa.GltfPrimitive(f(p), None, None, None, [f(u)], f(i.view(np.int32)), m) for p, i, u, m in prims]
    coll = aa.PointLightCollection.from_lights(torch.from_numpy(lights).cuda(), shadow_size=size)
    coll.drawShadowMapsGltf(gp, mats, texs)
    torch.cuda.synchronize()
    chains = [t.data.cpu().numpy() for t in texs]
    return coll, coll.getShadowMapImage().cpu().numpy(), mats, chains


@pytest.mark.parametrize("size", [96, 37])
def test_gltf_alpha_tested_cubes_match_oracle(torch, size):
    """althea_shadow_cubes_gltf against the oracle's alpha-tested pass (which
    samples the GPU's own mip chain).  Texels whose alpha decision (or a
    texture decision) lies within MARGIN of its threshold may go either way
    and are only counted."""
    sc = aa.synth.scene(seed=SEED0 + 61, width=8, height=8, lights=5)
    prims, base = _gltf_shadow_scene(SEED0 + 61)
    _, got, mats, chains = _gltf_shadow_gpu(torch, prims, base, sc.lights, size)
    otex = [ol.OracleTex(chains[0], 64, 64, 7, 1, 1, 1, 1, 0, 0), ol.OracleTex(chains[1], 1, 1, 1, 0, 1, 1, 1, 0, 0)]
    oprims = [{"pos": p, "uvs": [u], "idx": i, "material": m} for p, i, u, m in prims]
    ref, margin = ol.shadow_cubes_gltf(oprims, [m.c for m in mats], otex, sc.lights, size)
    plain = ol.shadow_cubes(np.concatenate([prims[0][0], prims[1][0]]),
                            np.concatenate([prims[0][1], prims[1][1] + len(prims[0][0])]), sc.lights, size)
    clear = margin > 2e-4
    r = _compare_cubes(got[clear], ref[clear])
    r["clear"] = float(clear.mean())
    r["cut"] = int(((plain < 1.0) & (ref != plain)).sum())
    print("gltf shadow cubes", size, r)
    assert r["cut"] > 0.02 * r["hits"], r  # the alpha test removed texels
    assert r["clear"] > 0.97, r
    assert r["hit_disagree"] <= 2e-3 * r["hits"] + 4, r
    assert r["depth_off"] <= 2e-3 * r["hits"] + 4, r
    assert r["err"]["mean"] < 1e-6, r


def test_gltf_opaque_material_is_the_mesh_pass(torch):
    """With every base colour alpha 1 the alpha test keeps every fragment:
    the glTF pass writes the mesh pass's cubes bit for bit (the same
    light-relative vertices and texel tests)."""
    sc = aa.synth.scene(seed=SEED0 + 62, width=8, height=8, lights=4)
    prims, base = _gltf_shadow_scene(SEED0 + 62, n=32)
    _, got, _, _ = _gltf_shadow_gpu(torch, prims, base, sc.lights, 64, opaque=True)
    pos = np.concatenate([prims[0][0], prims[1][0]])
    idx = np.concatenate([prims[0][1], prims[1][1] + len(prims[0][0])])
    _, want = _gpu_cubes(torch, pos, idx, sc.lights, 64)
    assert (want < 1.0).any()
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
