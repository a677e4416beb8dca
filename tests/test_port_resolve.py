"""The FP32 CPU port of the resolve (oracle/port_resolve.cpp, the kernel's
formulation on the host: the CPU baseline) against the FP64 oracle, and what
its df32 addressing steps buy.

On the config-0 frame (1280x720 fBm heightfield, 8 lights, SSAO 16) with
IBL from the config-0 source (512x256 log-normal noise + 2-degree suns; the
cube's mip 0 is the roughness-0 mirror of that noise), the port with df32
addressing stays within the 1e-4 max relative contract on every pixel clear
of an SSAO decision threshold.  Putting the background lookup back to FP32
addressing breaks it (the FP32 texel coordinate times the source's texel
contrast), which is the measured reason for that step."""
import numpy as np
import pytest

import althea_amd as aa
from tests import oracle_lib as ol
from tests.parity import MAX_REL, RADIANCE_FLOOR, rel_err, summary

SEED0 = 0x414C0000


def _oracle_views(env, size=128, mips=5, spp=32):
    e = ol.Env(env)
    ids = []
    for m in range(mips):
        s = max(1, size >> m)
        f, y, x = np.meshgrid(np.arange(6), np.arange(s), np.arange(s), indexing="ij")
        ids.append(np.stack([np.full(f.size, m), f.ravel(), x.ravel(), y.ravel()], 1))
    ids = np.concatenate(ids).astype(np.int32)
    pref = np.ones((ids.shape[0], 4), np.float32)
    pref[:, :3] = e.prefilter_texels(size, mips, spp, ids)[:, :3]
    irr_ids = np.array([[f, x, y] for f in range(6) for y in range(32) for x in range(32)], np.int32)
    irr = np.ones((irr_ids.shape[0], 4), np.float32)
    irr[:, :3] = e.irradiance_texels(32, irr_ids, 30)
    lut = np.ascontiguousarray(ol.brdf_lut(64, 128).astype(np.float32).reshape(-1, 2))
    return e, dict(prefiltered=pref, pref_size=size, pref_mips=mips, irradiance=irr, irr_size=32, lut=lut,
                   lut_size=64)


@pytest.fixture(scope="module")
def frame():
    env = aa.synth.env(seed=SEED0, width=512, height=256)
    oenv, views = _oracle_views(env)
    scene = aa.synth.scene(seed=SEED0, width=1280, height=720, lights=8)
    rng = np.random.default_rng(17)
    pix = np.stack([rng.integers(0, 1280, 8000), rng.integers(0, 720, 8000)], 1).astype(np.int32)
    ref, _, margin = ol.resolve(scene, views, oenv, pix)
    return env, views, scene, pix, ref, margin > 2e-3


def test_port_meets_the_contract_on_config0(frame):
    env, views, scene, pix, ref, clear = frame
    got = ol.port_resolve(scene, views, env, pix)
    err = rel_err(got[:, :3], ref, RADIANCE_FLOOR).max(1)
    assert clear.mean() > 0.6
    assert err[clear].max() < MAX_REL, summary(err[clear])
    assert np.all(got[:, 3] == 1.0)


def test_fp32_background_addressing_breaks_the_contract(frame):
    env, views, scene, pix, ref, clear = frame
    got = ol.port_resolve(scene, views, env, pix, fp32_steps=ol.PORT_FP32_BACKGROUND)
    err = rel_err(got[:, :3], ref, RADIANCE_FLOOR).max(1)
    assert err[clear].max() > MAX_REL, summary(err[clear])
    sky = scene.position[pix[:, 1], pix[:, 0], 3] == 0
    # the excess is on background pixels only
    assert err[clear & ~sky].max() < MAX_REL


def test_port_whole_frame_matches_pixel_calls(frame):
    env, views, scene, pix, ref, clear = frame
    small = aa.synth.scene(seed=SEED0 + 5, width=96, height=64, lights=4)
    whole = ol.port_resolve(small, views, env)
    ys, xs = np.mgrid[0:64, 0:96]
    some = np.stack([xs.ravel(), ys.ravel()], 1)[::7].astype(np.int32)
    np.testing.assert_array_equal(ol.port_resolve(small, views, env, some, threads=1),
                                  whole[some[:, 1] * 96 + some[:, 0]])
