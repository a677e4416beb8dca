"""GPU parity of the deferred resolve (DeferredPass.frag + pbrMaterial + SSAO)
against the FP64 oracle, through the C ABI, at the 1e-4 max-relative contract
on the smooth and on the configured (per-texel noise) sources.  The lookups
that mirror the source's texel contrast (the background, the mip-0
reflection tap) and the half vector near V = -L are what keep the noisy
configurations inside it (df32 addressing in the kernel).

The oracle shades the same FP32 G-buffer with the IBL state the GPU produced,
so this isolates the resolve from the precompute.  SSAO makes discrete
decisions (nearest-texel fetch, depth compare); the oracle reports each
pixel's smallest decision margin and pixels within MARGIN of a threshold are
counted separately as possible FP32/FP64 decision flips.
"""
import numpy as np
import pytest

import althea_amd as aa
from tests import oracle_lib as ol
from tests.parity import MAX_REL, RADIANCE_FLOOR, rel_err, summary

pytestmark = pytest.mark.gpu

SEED0 = 0x414C0000
MARGIN = 2e-3


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _ibl(torch, kind: str, size=128, mips=5):
    if kind == "smooth":
        env = aa.synth.env(seed=SEED0 + 100, width=512, height=256, sigma=0.0, suns=4, peak_min=100.0,
                           peak_max=100.0, sun_sigma_deg=8.0)
    else:
        env = aa.synth.env(seed=SEED0, width=512, height=256)
    desc = aa.IBLDesc(prefiltered_size=size, prefiltered_mips=mips)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    torch.cuda.synchronize()
    return ibl, ol.Env(env)


def _pixels(scene, n, seed):
    rng = np.random.default_rng(seed)
    W, H = scene.width, scene.height
    pts = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], 1)
    border = np.array([[0, 0], [W - 1, 0], [0, H - 1], [W - 1, H - 1], [W // 2, 0], [0, H // 2]])
    return np.concatenate([pts, border]).astype(np.int32)


def _run(torch, scene, ibl, fp16=False, tonemap=False, out_fp16=False):
    g = scene.gbuffer_tensors("cuda", fp16=fp16)
    ao = torch.empty(scene.pixels, dtype=torch.float32, device="cuda")
    out = aa.DeferredRendering.resolve(g, ibl, scene.lights_tensor("cuda"), scene.globals, scene.ssao,
                                       tonemap=tonemap, out_fp16=out_fp16, ao_out=ao)
    torch.cuda.synchronize()
    return out.float().cpu().numpy().reshape(scene.height, scene.width, 4), ao.cpu().numpy().reshape(scene.height,
                                                                                                        scene.width)


def _check(scene, ibl, oenv, img, ao, pix, tonemap=False, gbuffer=None, min_clear=0.6):
    ref, ref_ao, margin = ol.resolve(scene, ibl.views_numpy(), oenv, pix, tonemap=tonemap, gbuffer=gbuffer)
    got = img[pix[:, 1], pix[:, 0], :3]
    got_ao = ao[pix[:, 1], pix[:, 0]]
    clear = margin > MARGIN
    err = rel_err(got, ref, RADIANCE_FLOOR).max(1)
    ao_err = np.abs(got_ao - ref_ao)
    s = summary(err[clear]) if clear.any() else {"max": 0.0}
    flips = int(np.sum(ao_err[~clear] > 1e-4))
    # With 16 taps per pixel, ~15% of pixels have a tap within MARGIN of a
    # texel boundary or depth threshold; they are checked for flips only.
    assert clear.mean() >= min_clear, clear.mean()
    assert not clear.any() or ao_err[clear].max() < 1e-5, summary(ao_err[clear])
    assert s["max"] < MAX_REL, s
    assert flips <= max(2, 0.01 * (~clear).sum()), flips
    return s


@pytest.mark.parametrize("kind", ["smooth", "config0"])
def test_resolve_config0_matches_oracle(torch, kind):
    ibl, oenv = _ibl(torch, kind)
    scene = aa.synth.scene(seed=SEED0, width=1280, height=720, lights=8)
    img, ao = _run(torch, scene, ibl)
    _check(scene, ibl, oenv, img, ao, _pixels(scene, 6000, 1))
    assert np.isfinite(img).all()
    assert np.all(img[..., 3] == 1.0)


def test_resolve_tonemap_and_fp16_output(torch):
    ibl, oenv = _ibl(torch, "smooth")
    scene = aa.synth.scene(seed=SEED0 + 1, width=320, height=200, lights=8)
    img, ao = _run(torch, scene, ibl, tonemap=True)
    _check(scene, ibl, oenv, img, ao, _pixels(scene, 3000, 2), tonemap=True)
    img16, _ = _run(torch, scene, ibl, tonemap=True, out_fp16=True)
    np.testing.assert_allclose(img16, img, rtol=1e-3, atol=1e-4)


def test_resolve_fp16_gbuffer(torch):
    ibl, oenv = _ibl(torch, "smooth")
    scene = aa.synth.scene(seed=SEED0 + 2, width=640, height=360, lights=8)
    img, ao = _run(torch, scene, ibl, fp16=True)
    planes16 = [p.astype(np.float16).astype(np.float32) for p in
                (scene.position, scene.normal, scene.albedo, scene.material)]
    _check(scene, ibl, oenv, img, ao, _pixels(scene, 3000, 3), gbuffer=planes16)


def test_resolve_config1_full_size(torch):
    """3840x2160, 64 lights, 512^2 x 6 mips prefiltered cube: full-size run,
    a sample of pixels against the oracle."""
    env = aa.synth.env(seed=SEED0 + 1, width=2048, height=1024)
    desc = aa.IBLDesc(prefiltered_size=512, prefiltered_mips=6, prefilter_samples=256)
    ibl = aa.ImageBasedLighting.createResources(torch.from_numpy(env).cuda(), desc)
    scene = aa.synth.scene(seed=SEED0 + 1, width=3840, height=2160, lights=64)
    img, ao = _run(torch, scene, ibl)
    assert np.isfinite(img).all()
    _check(scene, ibl, ol.Env(env), img, ao, _pixels(scene, 4000, 4))


def test_resolve_general_projection_path(torch):
    """Under a perspective projection whose w row is (0, 0, -1, 0) the SSAO
    taps take their view depth as clip w (bit for bit the same value).  The
    projection scaled by 2 (the same projective map, w row (0, 0, -2, 0))
    takes the general path, which forms w from the matrix: checked against
    the oracle, and close to the frame of the unscaled projection."""
    ibl, oenv = _ibl(torch, "smooth")
    scene = aa.synth.scene(seed=SEED0 + 6, width=640, height=360, lights=8)
    base, base_ao = _run(torch, scene, ibl)
    scene.globals_arr[0:16] *= 2.0
    scene.globals_arr[16:32] *= 0.5
    img, ao = _run(torch, scene, ibl)
    _check(scene, ibl, oenv, img, ao, _pixels(scene, 3000, 7))
    assert np.mean(np.abs(ao - base_ao) > 1e-4) < 0.01
    np.testing.assert_allclose(img[ao == base_ao], base[ao == base_ao], rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("W,H", [(1, 1), (7, 3), (63, 5), (65, 2), (1, 40), (130, 7)])
def test_resolve_ragged_frames_every_pixel(torch, W, H):
    """Frames smaller than one 64 x 4 tile, one pixel wide, or one column or
    row past a tile: every pixel against the oracle (the lanes outside the
    frame load a clamped pixel and store nothing, and SSAO taps that land
    off screen are rejected)."""
    ibl, oenv = _ibl(torch, "config0")
    scene = aa.synth.scene(seed=SEED0 + 7 + W, width=W, height=H, lights=8)
    img, ao = _run(torch, scene, ibl)
    assert np.isfinite(img).all()
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    pix = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    # A frame of a few pixels can have every pixel near an SSAO decision.
    _check(scene, ibl, oenv, img, ao, pix, min_clear=0.0)
    assert np.all(img[..., 3] == 1.0)


def test_resolve_no_lights_no_ssao(torch):
    ibl, oenv = _ibl(torch, "smooth")
    scene = aa.synth.scene(seed=SEED0 + 3, width=64, height=48, lights=0)
    scene.ssao_enabled = False
    img, ao = _run(torch, scene, ibl)
    pix = _pixels(scene, 500, 5)
    ref, ref_ao, _ = ol.resolve(scene, ibl.views_numpy(), oenv, pix)
    err = rel_err(img[pix[:, 1], pix[:, 0], :3], ref, RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)
    np.testing.assert_array_equal(ao[pix[:, 1], pix[:, 0]], ref_ao)


def test_views_without_bordered_cubes_build_them_per_call(torch):
    """The resolve reads seamless-bordered copies of the prefiltered and
    irradiance cubes.  With the handle's copies and with views that carry none
    (the resolve builds them in its own allocation), the frames are the same
    bits."""
    ibl, _ = _ibl(torch, "config0")
    v = ibl.getHandles()
    assert v.prefiltered_bordered and v.irradiance_bordered
    bare = aa._lib.IBLViewsC()
    for name, _t in aa._lib.IBLViewsC._fields_:
        setattr(bare, name, getattr(v, name))
    bare.prefiltered_bordered = None
    bare.irradiance_bordered = None
    scene = aa.synth.scene(seed=SEED0 + 4, width=640, height=360, lights=8)
    a, ao_a = _run(torch, scene, v)
    b, ao_b = _run(torch, scene, bare)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    np.testing.assert_array_equal(ao_a.view(np.uint32), ao_b.view(np.uint32))
