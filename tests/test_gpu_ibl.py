"""GPU parity of the IBL precompute kernels against the FP64 oracle, through
the C ABI.  Sizes are the benchmark's small configuration (512x256 source,
128^2 x 6 x 5 prefiltered mips at 1024 samples, 32^2 x 6 irradiance, 128^2
BRDF LUT) plus edge cases; the oracle evaluates a subset of texels.

Two sources, both held to the 1e-4 max-relative contract:
  smooth  background 1 + four 8-degree sun lobes (peak 100);
  config0 per-texel log-normal(0, 1.5) noise + three 2-degree suns of 1e3:
          neighbouring texels differ by up to 1000x, so a mirror-like lookup
          (the roughness-0 mip, the env_sample probe) carries its texel
          coordinate in df32; rough mips average 1000+ FP32 lookups whose
          addressing errors cancel.
"""
import numpy as np
import pytest

import althea_amd as aa
from tests import oracle_lib as ol
from tests.parity import (LUT_FLOOR, MAX_REL, RADIANCE_FLOOR, cube_gather, cube_texel_ids, rel_err, summary)

pytestmark = pytest.mark.gpu

SEED0 = 0x414C0000


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _env(kind: str):
    if kind == "smooth":
        return aa.synth.env(seed=SEED0 + 100, width=512, height=256, sigma=0.0, suns=4, peak_min=100.0,
                            peak_max=100.0, sun_sigma_deg=8.0)
    return aa.synth.env(seed=SEED0, width=512, height=256)


@pytest.fixture(scope="module", params=["smooth", "config0"])
def env_case(request, torch):
    env = _env(request.param)
    dev = torch.from_numpy(env).cuda()
    pyr = aa.env_pyramid(dev)
    torch.cuda.synchronize()
    return request.param, env, dev, pyr, ol.Env(env)


# --------------------------------------------------------------- Hammersley
@pytest.mark.parametrize("n", [1024, 4096, 1000, 1 << 16])
def test_hammersley_bit_exact(torch, n):
    bits, xi = aa.hammersley(n)
    torch.cuda.synchronize()
    rbits, rx32, _ = ol.hammersley(n, 0, n)
    np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint32), rbits)
    np.testing.assert_array_equal(xi.cpu().numpy().view(np.uint32), rx32.view(np.uint32))


def test_hammersley_offset_window(torch):
    bits, xi = aa.hammersley(1 << 20, first=(1 << 20) - 5000, count=5000)
    torch.cuda.synchronize()
    rbits, rx32, _ = ol.hammersley(1 << 20, (1 << 20) - 5000, 5000)
    np.testing.assert_array_equal(bits.cpu().numpy().view(np.uint32), rbits)
    np.testing.assert_array_equal(xi.cpu().numpy().view(np.uint32), rx32.view(np.uint32))


# ------------------------------------------------------------------ pyramid
def test_pyramid_matches_oracle(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    ref = oenv.pyramid()
    got = pyr.cpu().numpy().astype(np.float64)
    assert got.shape == ref.shape
    err = rel_err(got, ref, 1e-30)
    assert err.max() < 1e-6, summary(err)


def test_pyramid_4096_levels(torch):
    env = aa.synth.env(seed=SEED0 + 2, width=4096, height=2048, sigma=1.5, suns=8, peak_min=1e3, peak_max=1e4)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda()).cpu().numpy()
    assert pyr.shape[0] == aa.lib().althea_env_pyramid_texels(4096, 2048)
    # Last level (1x1) is the mean of the whole map: a checksum of checksums.
    ref = env.astype(np.float64).reshape(-1, 4).mean(0)
    np.testing.assert_allclose(pyr[-1], ref, rtol=2e-6)


# ---------------------------------------------------------------- samplers
def test_env_sample_matches_oracle(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    rng = np.random.default_rng(1)
    q = np.concatenate([rng.normal(size=(20000, 3)), rng.uniform(-0.5, 9.5, size=(20000, 1))], 1).astype(np.float32)
    q[:8, :3] = [[0, 1, 0], [0, -1, 0], [0, 1, 0.0009], [1, 0, 0], [-1, 0, 1e-7], [-1, 0, -1e-7], [0, 0, 1], [0, 0, -1]]
    got = aa.env_sample(pyr, 512, 256, torch.from_numpy(q).cuda()).cpu().numpy()[:, :3]
    ref = oenv.lookup(q[:, :3].astype(np.float64), q[:, 3].astype(np.float64))
    err = rel_err(got, ref, RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


def test_cube_sample_matches_oracle(torch):
    rng = np.random.default_rng(2)
    size, mips = 16, 5
    n = aa.cube_texels(size, mips)
    cube = rng.uniform(0.2, 3.0, size=(n, 4)).astype(np.float32)
    q = np.concatenate([rng.normal(size=(30000, 3)), rng.uniform(-0.5, 4.5, size=(30000, 1))], 1)
    # Directions near face edges and corners exercise the seamless folds.
    q[:10000, :3] = np.sign(rng.normal(size=(10000, 3))) * rng.uniform(0.97, 1.0, size=(10000, 3))
    q = q.astype(np.float32)
    got = aa.cube_sample(torch.from_numpy(cube).cuda(), size, mips, torch.from_numpy(q).cuda()).cpu().numpy()
    ref = ol.cube_sample(cube, size, mips, q[:, :3].astype(np.float64), q[:, 3].astype(np.float64))
    err = rel_err(got[:, :3], ref, RADIANCE_FLOOR)
    assert err.max() < 2e-5, summary(err)


def _bordered_sample(bc, size, mips, dirs, lods):
    """Trilinear lookup of a seamless-bordered cube in FP64: four plain taps
    per mip at (floor + 1, floor + 2) of the bordered face, no folding."""
    face, st = ol.cube_select(dirs)
    offs, o = [], 0
    for m in range(mips):
        offs.append(o)
        o += 6 * (max(1, size >> m) + 2) ** 2

    def bil(m):
        S = max(1, size >> m)
        W = S + 2
        x, y = st[:, 0] * S - 0.5, st[:, 1] * S - 0.5  # the oracle's (s, t) are in [0, 1]
        fx, fy = np.floor(x), np.floor(y)
        ax, ay = (x - fx)[:, None], (y - fy)[:, None]
        base = offs[m] + (face * W + fy.astype(np.int64) + 1) * W + fx.astype(np.int64) + 1
        t = lambda k: bc[k, :3].astype(np.float64)
        top = t(base) * (1 - ax) + t(base + 1) * ax
        bot = t(base + W) * (1 - ax) + t(base + W + 1) * ax
        return top * (1 - ay) + bot * ay

    d = np.clip(lods, 0, mips - 1)
    fl = np.floor(d)
    l0 = fl.astype(np.int64)
    delta = (d - fl)[:, None]
    out = np.zeros((dirs.shape[0], 3))
    for m in range(mips):
        sel = l0 == m
        if not sel.any():
            continue
        a = bil(m)[sel]
        b = bil(min(m + 1, mips - 1))[sel]
        out[sel] = a * (1 - delta[sel]) + b * delta[sel]
    return out


@pytest.mark.parametrize("size,mips", [(16, 5), (1, 1), (5, 1)])
def test_cube_border_matches_seamless_oracle(torch, size, mips):
    """The bordered copy the resolve reads: interior texels are the cube's,
    and lookups near edges and corners (four plain taps, no folding) equal the
    oracle's seamless trilinear lookup; 5 is an irradiance-like odd size."""
    rng = np.random.default_rng(3 + size)
    n = aa.cube_texels(size, mips)
    cube = rng.uniform(0.2, 3.0, size=(n, 4)).astype(np.float32)
    bc = aa.cube_border(torch.from_numpy(cube).cuda(), size, mips).cpu().numpy()
    assert bc.shape[0] == aa.lib().althea_cube_bordered_texels(size, mips)
    o, co = 0, 0
    for m in range(mips):
        S = max(1, size >> m)
        faces = bc[o:o + 6 * (S + 2) ** 2].reshape(6, S + 2, S + 2, 4)
        np.testing.assert_array_equal(faces[:, 1:S + 1, 1:S + 1, :3], cube[co:co + 6 * S * S].reshape(6, S, S, 4)[..., :3])
        o += 6 * (S + 2) ** 2
        co += 6 * S * S
    q = np.concatenate([rng.normal(size=(20000, 3)), rng.uniform(-0.5, mips - 0.5, size=(20000, 1))], 1)
    q[:8000, :3] = np.sign(rng.normal(size=(8000, 3))) * rng.uniform(0.97, 1.0, size=(8000, 3))
    got = _bordered_sample(bc, size, mips, q[:, :3], q[:, 3])
    ref = ol.cube_sample(cube, size, mips, q[:, :3], q[:, 3])
    err = rel_err(got, ref, RADIANCE_FLOOR)
    assert err.max() < 2e-6, summary(err)


# ---------------------------------------------------------------- BRDF LUT
def test_brdf_lut_matches_oracle(torch):
    got = aa.brdf_lut(128, 1024).cpu().numpy().reshape(-1, 2)
    ref = ol.brdf_lut(128, 1024)
    err = rel_err(got, ref, LUT_FLOOR)
    assert err.max() < MAX_REL, summary(err)


@pytest.mark.parametrize("size,samples", [(1, 1), (7, 33), (64, 4096)])
def test_brdf_lut_edge_sizes(torch, size, samples):
    got = aa.brdf_lut(size, samples).cpu().numpy().reshape(-1, 2)
    ref = ol.brdf_lut(size, samples)
    err = rel_err(got, ref, LUT_FLOOR)
    assert err.max() < MAX_REL, summary(err)


# --------------------------------------------------------------- prefilter
def test_prefilter_cube_matches_oracle(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    cube = aa.prefilter_cube(pyr, 512, 256, 128, 5, 1024).cpu().numpy()
    ids = cube_texel_ids(128, 5, per_face=40, seed=3)
    got = cube_gather(cube, 128, ids)
    ref = oenv.prefilter_texels(128, 5, 1024, ids)
    err = rel_err(got, ref[:, :3], RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)
    mirror = ids[:, 0] == 0
    assert err[mirror].max() < 1e-5, summary(err[mirror])  # one df32-addressed lookup per texel
    # Alpha channel and the untouched layout.
    assert np.all(cube[:, 3] == 1.0)


def test_prefilter_fp32_sums_variant(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    c64 = aa.prefilter_cube(pyr, 512, 256, 32, 3, 1024).cpu().numpy()
    c32 = aa.prefilter_cube(pyr, 512, 256, 32, 3, 1024, fp32_sums=True).cpu().numpy()
    ids = cube_texel_ids(32, 3, per_face=20, seed=4)
    ref = oenv.prefilter_texels(32, 3, 1024, ids)[:, :3]
    e64 = rel_err(cube_gather(c64, 32, ids), ref, RADIANCE_FLOOR)
    e32 = rel_err(cube_gather(c32, 32, ids), ref, RADIANCE_FLOOR)
    assert e64.mean() <= e32.mean() * 1.5 + 1e-7, (summary(e64), summary(e32))
    if name == "smooth":
        assert e32.max() < 1e-3, summary(e32)


@pytest.mark.parametrize("size,mips,samples", [(1, 1, 16), (8, 4, 7), (64, 7, 64)])
def test_prefilter_edge_shapes(torch, size, mips, samples):
    # Shape coverage on the smooth source, where the max-relative contract holds.
    env = _env("smooth")
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    cube = aa.prefilter_cube(pyr, 512, 256, size, mips, samples).cpu().numpy()
    assert cube.shape[0] == aa.cube_texels(size, mips)
    ids = cube_texel_ids(size, mips, per_face=4, seed=5)
    err = rel_err(cube_gather(cube, size, ids), ol.Env(env).prefilter_texels(size, mips, samples, ids)[:, :3],
                  RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


@pytest.mark.parametrize("size,mips,samples", [(64, 5, 128), (256, 5, 1024), (128, 5, 1000), (64, 4, 100)])
def test_prefilter_sample_split_shapes(torch, size, mips, samples):
    """Cubes whose rough mips have few 8x8 tiles split each tile's sample
    table over a block of waves (chunks of a multiple of 32 entries, FP64
    partials added in order), as long as every chunk keeps 64 entries.
    These shapes take 2 (132 rough tiles, 128 samples), 4 (2 040 tiles),
    8 (510 tiles; 1 000 samples, so the last chunk is short) and 1 wave
    (100 samples) per tile."""
    env = _env("smooth")
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    cube = aa.prefilter_cube(pyr, 512, 256, size, mips, samples).cpu().numpy()
    ids = cube_texel_ids(size, mips, per_face=6, seed=6)
    err = rel_err(cube_gather(cube, size, ids), ol.Env(env).prefilter_texels(size, mips, samples, ids)[:, :3],
                  RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)
    # The allocation-free entry point runs the same kernel: the same bits.
    from althea_amd import _lib
    again = torch.full((aa.cube_texels(size, mips), 4), float("nan"), dtype=torch.float32, device="cuda")
    st = aa.lib().althea_prefilter_cube_noalloc(pyr.data_ptr(), 512, 256, size, mips, samples, 0, again.data_ptr(),
                                                torch.cuda.current_stream().cuda_stream)
    assert st == _lib.ALTHEA_OK, aa.lib().althea_last_error()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(again.cpu().numpy().view(np.uint32), cube.view(np.uint32))


def test_prefilter_points_match_oracle(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    q = aa.synth.dir_rough(seed=9, n=3000, rmin=0.05, rmax=1.0)
    got = aa.prefilter_points(pyr, 512, 256, 1024, torch.from_numpy(q).cuda()).cpu().numpy()[:, :3]
    ref = oenv.prefilter_points(1024, q)
    err = rel_err(got, ref, RADIANCE_FLOOR)
    s = summary(err)
    assert s["max"] < (MAX_REL if name == "smooth" else 1e-3) and s["mean"] < 1e-5, s


def test_prefilter_points_agree_with_cube_texels(torch):
    # The point kernel at a cube texel's direction and roughness m/(mips-1)
    # estimates the same quantity as the cube kernel (same samples).
    env = _env("smooth")
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    cube = aa.prefilter_cube(pyr, 512, 256, 16, 5, 256).cpu().numpy()
    ids = cube_texel_ids(16, 5, per_face=3, seed=6)
    raw = []
    for m, f, x, y in ids:
        s = max(1, 16 >> int(m))
        d = ol.cube_texel_dirs(s, [[f, x, y]])[0]
        raw.append([*d, m / 4.0])
    q = np.array(raw, np.float32)
    got = aa.prefilter_points(pyr, 512, 256, 256, torch.from_numpy(q).cuda()).cpu().numpy()[:, :3]
    err = rel_err(got, cube_gather(cube, 16, ids), RADIANCE_FLOOR)
    assert err.max() < 2e-4, summary(err)


@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 4), (4, 16), (16, 4), (64, 8)])
def test_gather_copy_edge_sources(torch, w, h):
    """Sources down to one texel and of either aspect.  The prefilter and the
    irradiance read a padded copy of each level (its REPEAT neighbours stored
    around it); both match the oracle, and the allocation-free prefilter, which
    reads the pyramid itself, writes the same bits."""
    env = aa.synth.env(seed=SEED0 + 131 * w + h, width=w, height=h, sigma=0.5, suns=1, peak_min=10.0,
                       peak_max=10.0, sun_sigma_deg=20.0)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    oenv = ol.Env(env)
    size, mips, samples = 8, 4, 64
    n = aa.cube_texels(size, mips)
    ref = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    got = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    L = aa.lib()
    print(f"gather edge source {w}x{h}", flush=True)
    assert L.althea_prefilter_cube(pyr.data_ptr(), w, h, size, mips, samples, 0, ref.data_ptr(), stream) == 0, \
        L.althea_last_error()
    assert L.althea_prefilter_cube_noalloc(pyr.data_ptr(), w, h, size, mips, samples, 0, got.data_ptr(),
                                           stream) == 0, L.althea_last_error()
    torch.cuda.synchronize()
    cube = ref.cpu().numpy()
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint32), cube.view(np.uint32))
    ids = cube_texel_ids(size, mips, per_face=4, seed=11)
    err = rel_err(cube_gather(cube, size, ids), oenv.prefilter_texels(size, mips, samples, ids)[:, :3],
                  RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)
    theta = 16
    irr = aa.irradiance_cube(pyr, w, h, 4, theta).cpu().numpy()
    ids = cube_texel_ids(4, 1, per_face=3, seed=12)
    err = rel_err(cube_gather(irr, 4, ids), oenv.irradiance_texels(4, ids[:, 1:], theta), RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


@pytest.mark.parametrize("w,h,theta", [(64, 64, 24), (32, 128, 12), (1024, 4, 300), (256, 128, 100)])
def test_irradiance_blend_grid_aspects(torch, w, h, theta):
    """Two source levels (a fractional LOD log2(W / theta)) read through their
    blend grid, on square, tall and 256:1 sources and at a 1.36 LOD: the
    grid's nodes follow each axis of the fine level separately."""
    env = aa.synth.env(seed=SEED0 + 7 * w + h, width=w, height=h, sigma=0.5, suns=2, peak_min=10.0, peak_max=20.0,
                       sun_sigma_deg=15.0)
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    lod = np.log2(w / theta)
    assert lod % 1.0 != 0.0
    irr = aa.irradiance_cube(pyr, w, h, 8, theta).cpu().numpy()
    ids = cube_texel_ids(8, 1, per_face=6, seed=13)
    err = rel_err(cube_gather(irr, 8, ids), ol.Env(env).irradiance_texels(8, ids[:, 1:], theta), RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


# -------------------------------------------------------------- irradiance
def test_irradiance_cube_matches_oracle(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    cube = aa.irradiance_cube(pyr, 512, 256, 32).cpu().numpy()
    ids = cube_texel_ids(32, 1, per_face=30, seed=7)
    got = cube_gather(cube, 32, ids)
    ref = oenv.irradiance_texels(32, ids[:, 1:])
    err = rel_err(got, ref, RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


def test_irradiance_fp32_sums_variant(env_case, torch):
    name, env, dev, pyr, oenv = env_case
    c32 = aa.irradiance_cube(pyr, 512, 256, 8, fp32_sums=True).cpu().numpy()
    ids = cube_texel_ids(8, 1, per_face=8, seed=8)
    err = rel_err(cube_gather(c32, 8, ids), oenv.irradiance_texels(8, ids[:, 1:]), RADIANCE_FLOOR)
    assert err.max() < 1e-3, summary(err)


@pytest.mark.parametrize("size,theta", [(1, 300), (5, 17), (8, 100), (32, 300), (1, 4096), (1, 4097)])
def test_irradiance_edge_shapes(torch, size, theta):
    """Cube sizes and azimuth counts; up to 4096 azimuth rows the kernel
    reads its azimuth table from LDS, above it from global memory."""
    env = _env("smooth")
    pyr = aa.env_pyramid(torch.from_numpy(env).cuda())
    cube = aa.irradiance_cube(pyr, 512, 256, size, theta).cpu().numpy()
    ids = cube_texel_ids(size, 1, per_face=3, seed=9)
    err = rel_err(cube_gather(cube, size, ids), ol.Env(env).irradiance_texels(size, ids[:, 1:], theta),
                  RADIANCE_FLOOR)
    assert err.max() < MAX_REL, summary(err)


# ------------------------------------------------------------- IBL handle
def test_ibl_handle_matches_low_level_calls(torch):
    env = _env("config0")
    dev = torch.from_numpy(env).cuda()
    desc = aa.IBLDesc(irradiance_size=16, prefiltered_size=32, prefiltered_mips=4, prefilter_samples=256,
                      brdf_lut_size=32, brdf_lut_samples=256)
    ibl = aa.ImageBasedLighting.createResources(dev, desc)
    pyr = aa.env_pyramid(dev)
    torch.testing.assert_close(ibl.environment_tensor(), pyr, rtol=0, atol=0)
    torch.testing.assert_close(ibl.prefiltered_tensor(), aa.prefilter_cube(pyr, 512, 256, 32, 4, 256), rtol=0, atol=0)
    torch.testing.assert_close(ibl.irradiance_tensor(), aa.irradiance_cube(pyr, 512, 256, 16), rtol=0, atol=0)
    torch.testing.assert_close(ibl.lut_tensor(), aa.brdf_lut(32, 256).reshape(-1, 2), rtol=0, atol=0)
    v = ibl.getHandles()
    assert (v.env_width, v.env_height, v.env_levels) == (512, 256, 10)
    assert ibl.state_bytes() == 16 * aa.cube_texels(32, 4) + 16 * 6 * 16 * 16 + 8 * 32 * 32


def test_invalid_arguments_return_status(torch):
    dev = torch.zeros((256, 500, 4), device="cuda")
    with pytest.raises(aa.AltheaError, match="power-of-two"):
        aa.env_pyramid(dev)
    pyr = aa.env_pyramid(torch.zeros((16, 32, 4), device="cuda"))
    with pytest.raises(aa.AltheaError, match="samples"):
        aa.prefilter_cube(pyr, 32, 16, 8, 2, 0)
    with pytest.raises(aa.AltheaError, match="mips"):
        aa.prefilter_cube(pyr, 32, 16, 8, 5, 16)
