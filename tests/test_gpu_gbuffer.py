"""GPU parity of the G-buffer producer (althea_gbuffer_draw: SceneToGBufferPass
with GBufferPassThrough.frag) against the FP64 oracle, through the C ABI:

- the synthetic heightfield as a mesh with per-vertex normals, colours and
  MRO: coverage and all four planes, FP32 and FP16 planes;
- the reference's floor (Floor.vert) at 4K: two triangles that cover most of
  the frame, every pixel;
- overlapping random triangles (depth order, culling, triangles crossing the
  near plane) with zero normals (the derivative normal);
- an empty mesh (every pixel uncovered) and a G-buffer that feeds the resolve.

A pixel centre within rounding of a triangle edge may go to the neighbour in
FP32: the oracle reports each pixel's smallest margin (the ray's angular
distance to the edge planes, relative depth gaps) and pixels within MARGIN
are only counted.
"""
import numpy as np
import pytest

import althea_amd as aa
from tests import oracle_lib as ol
from tests.test_oracle_gbuffer import _floor

pytestmark = pytest.mark.gpu

SEED0 = 0x414C0000
MARGIN = 2e-5


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _mesh(torch, pos, idx, col, mro, nrm):
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    return aa.GBufferMesh(f(pos), torch.from_numpy(np.ascontiguousarray(idx, np.uint32).view(np.int32)).cuda(),
                          f(col), f(mro), f(nrm) if nrm is not None else None)


def _draw(torch, sc, mesh_arrays, fp16=False):
    pas = aa.SceneToGBufferPass(sc.width, sc.height, fp16=fp16)
    g = pas.begin(_mesh(torch, *mesh_arrays), sc.globals)
    torch.cuda.synchronize()
    return np.stack([p.float().cpu().numpy().reshape(-1, 4) for p in g.planes], 1).astype(np.float64)


def _compare(got, ref, margin, tri):
    clear = margin > MARGIN
    cov_g, cov_r = got[:, 0, 3] == 1.0, tri >= 0
    r = {"pixels": len(tri), "covered": int(cov_r.sum()), "clear": float(clear.mean()),
         "coverage_disagree_clear": int(((cov_g != cov_r) & clear).sum()),
         "coverage_disagree_near": int(((cov_g != cov_r) & ~clear).sum())}
    both = cov_g & cov_r & clear
    scale = np.maximum(np.abs(ref[both]), 1.0)
    err = np.abs(got[both] - ref[both]) / scale
    r["err_by_plane"] = [float(err[:, k].max()) if both.any() else 0.0 for k in range(4)]
    r["mean_by_plane"] = [float(err[:, k].mean()) if both.any() else 0.0 for k in range(4)]
    unc = ~cov_g & ~cov_r
    r["uncovered_exact"] = bool(np.all(got[unc] == ref[unc]))
    return r


@pytest.mark.parametrize("fp16", [False, True])
def test_heightfield_mesh_matches_oracle(torch, fp16):
    sc = aa.synth.scene(seed=SEED0 + 41, width=480, height=270, lights=1)
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 41, n=48)
    nrm, col, mro = aa.synth.mesh_attributes(SEED0 + 41, pos, idx)
    got = _draw(torch, sc, (pos, idx, col, mro, nrm), fp16=fp16)
    ref, margin, tri = ol.gbuffer(sc, pos, idx, col, mro, nrm)
    if fp16:
        ref_q = ref.astype(np.float16).astype(np.float64)
        r = _compare(got, ref_q, margin, tri)
        print("heightfield fp16", r)
        assert max(r["err_by_plane"]) < 2.0 ** -10, r
    else:
        r = _compare(got, ref, margin, tri)
        print("heightfield", r)
        # FP32 barycentrics: the edge functions of a triangle 0.4 units wide
        # seen from 12 units at a grazing angle (the terrain under a camera
        # 3.5 high) carry about 1e-5 of rounding relative to their sum, which
        # moves an interpolated attribute by up to 1e-4 of its spread across
        # the triangle (random colours and MRO spread by up to 0.9).
        assert max(r["err_by_plane"]) < 1e-4 and max(r["mean_by_plane"]) < 2e-6, r
    assert r["covered"] > 0.5 * r["pixels"] and r["clear"] > 0.97, r
    assert r["coverage_disagree_clear"] == 0, r
    assert r["coverage_disagree_near"] <= 1e-3 * r["pixels"], r
    assert r["uncovered_exact"], r


def test_floor_4k_every_pixel(torch):
    sc = aa.synth.scene(seed=SEED0 + 42, width=3840, height=2160, lights=1)
    pos, idx, nrm, col, mro = _floor(h=-0.25, w=8.0)
    got = _draw(torch, sc, (pos, idx, col, mro, nrm))
    ref, margin, tri = ol.gbuffer(sc, pos, idx, col, mro, nrm)
    r = _compare(got, ref, margin, tri)
    print("floor 4k", r)
    assert r["covered"] > 0.3 * r["pixels"], r
    assert r["coverage_disagree_clear"] == 0 and r["coverage_disagree_near"] <= 200, r
    assert max(r["err_by_plane"]) < 1e-5, r


@pytest.mark.parametrize("seed", [1, 2])
def test_overlapping_triangles_zero_normals(torch, seed):
    """Random overlapping triangles around the camera, some crossing its near
    plane or passing behind it; zero normals take the derivative normal."""
    sc = aa.synth.scene(seed=SEED0 + 43, width=320, height=180, lights=1)
    rng = np.random.default_rng(seed)
    T = 400
    eye = np.array([0.0, 3.5, 12.0])
    centres = eye + rng.normal(0.0, 4.0, (T, 1, 3)) - np.array([0.0, 0.0, 6.0])
    pos = (centres + rng.normal(0.0, 2.0, (T, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    col = rng.uniform(0, 1, pos.shape).astype(np.float32)
    mro = rng.uniform(0, 1, pos.shape).astype(np.float32)
    got = _draw(torch, sc, (pos, idx, col, mro, None))
    ref, margin, tri = ol.gbuffer(sc, pos, idx, col, mro, None)
    r = _compare(got, ref, margin, tri)
    print("overlapping", seed, r)
    assert r["covered"] > 0.3 * r["pixels"], r
    assert r["coverage_disagree_clear"] == 0, r
    # Depth order and attributes on clear pixels; the derivative normal is the
    # plane's unit normal from FP32 differences of neighbouring hits.
    assert max(r["err_by_plane"][k] for k in (0, 2, 3)) < 1e-4, r
    assert max(r["mean_by_plane"][k] for k in (0, 2, 3)) < 2e-6, r
    assert r["err_by_plane"][1] < 1e-3, r


@pytest.mark.parametrize("W,H", [(1, 1), (7, 3), (1, 40)])
def test_small_extents_and_degenerate_triangles(torch, W, H):
    """Extents down to one pixel; the heightfield mixed with degenerate
    triangles (collinear vertices, a repeated vertex) that must cover
    nothing."""
    sc = aa.synth.scene(seed=SEED0 + 46, width=W, height=H, lights=1)
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 46, n=16)
    nrm, col, mro = aa.synth.mesh_attributes(SEED0 + 46, pos, idx)
    v = len(pos)
    pos = np.concatenate([pos, np.float32([[0, 1, 0], [1, 1, 0], [2, 1, 0]])])  # exactly collinear
    nrm, col, mro = (np.concatenate([a, a[:3]]) for a in (nrm, col, mro))
    extra = np.array([[v, v + 1, v + 2], [5, 5, 9], [3, 3, 3]], np.uint32)
    idx2 = np.concatenate([extra, idx])
    got = _draw(torch, sc, (pos, idx2, col, mro, nrm))
    ref, margin, tri = ol.gbuffer(sc, pos, idx2, col, mro, nrm)
    r = _compare(got, ref, margin, tri)
    print("small", W, H, r)
    assert not np.any((tri >= 0) & (tri < 3))
    assert r["coverage_disagree_clear"] == 0, r
    assert r["uncovered_exact"], r
    if r["covered"]:
        assert max(r["err_by_plane"]) < 1e-4, r


def test_empty_mesh_and_resolve(torch):
    sc = aa.synth.scene(seed=SEED0 + 44, width=200, height=120, lights=4)
    z = np.zeros((0, 3), np.float32)
    got = _draw(torch, sc, (z, np.zeros((0, 3), np.uint32), z, z, None))
    assert np.all(got[:, 0] == 0.0) and np.all(got[:, 1:, :3] == 0.0) and np.all(got[:, 1:, 3] == 1.0)
    # A produced G-buffer is a resolve input: shaded where covered, the
    # background elsewhere, and finite.
    env = aa.synth.env(seed=SEED0 + 45, width=256, height=128)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(),
                                                aa.IBLDesc(prefiltered_size=64, prefiltered_mips=5))
    pos, idx = aa.synth.heightfield_mesh(SEED0 + 44, n=32)
    nrm, col, mro = aa.synth.mesh_attributes(SEED0 + 44, pos, idx)
    pas = aa.SceneToGBufferPass(sc.width, sc.height)
    g = pas.begin(_mesh(torch, pos, idx, col, mro, nrm), sc.globals)
    out = aa.DeferredRendering.resolve(g, ibl, sc.lights_tensor("cuda"), sc.globals, sc.ssao)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.all(np.isfinite(o))
    covered = g.planes[0][..., 3].cpu().numpy().ravel() == 1.0
    assert covered.mean() > 0.3
