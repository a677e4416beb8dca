"""RGBE (.hdr) reader and writer of the C ABI (host code, no GPU).

Checks:
- The reference's own cache file (tests/golden/neoclassical_prefiltered5.hdr,
  written by stb_image_write) decodes identically to an independent
  pure-Python RGBE decoder below: the run-length parser and the
  2^(E - 136) conversion.
- write -> read round trips within the RGBE quantum (max component * 2^-7
  per channel, exact for representable values), on both the run-length
  path (width >= 8) and the flat path (width < 8), with long runs.
- Missing and malformed files give the documented status codes.
"""
from pathlib import Path

import numpy as np
import pytest

from althea_amd import hdr
from althea_amd._lib import AltheaError

GOLDEN = Path(__file__).parent / "golden" / "neoclassical_prefiltered5.hdr"


def py_decode(data: bytes) -> np.ndarray:
    """Independent decoder: header, then per scanline either the 2,2,hi,lo
    run-length form (channels in turn, count > 128 = run) or flat RGBE."""
    lines = data.split(b"\n")
    assert lines[0].startswith(b"#?")
    k = 1
    while lines[k] != b"":
        k += 1
    res = lines[k + 1].split()
    assert res[0] == b"-Y" and res[2] == b"+X"
    h, w = int(res[1]), int(res[3])
    pos = sum(len(l) + 1 for l in lines[:k + 2])
    buf = memoryview(data)[pos:]
    img = np.zeros((h, w, 4), np.uint8)
    i = 0
    for y in range(h):
        if 8 <= w < 32768 and buf[i] == 2 and buf[i + 1] == 2 and not (buf[i + 2] & 0x80):
            assert (buf[i + 2] << 8 | buf[i + 3]) == w
            i += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = buf[i]
                    i += 1
                    if n > 128:
                        img[y, x:x + n - 128, c] = buf[i]
                        i += 1
                        x += n - 128
                    else:
                        img[y, x:x + n, c] = np.frombuffer(buf[i:i + n], np.uint8)
                        i += n
                        x += n
        else:
            img[y] = np.frombuffer(buf[i:i + 4 * w], np.uint8).reshape(w, 4)
            i += 4 * w
    assert i == len(buf)
    e = img[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = img[..., :3] * scale[..., None]
    return out


def test_reference_cache_file_decodes_like_independent_decoder():
    assert hdr.info(GOLDEN) == (128, 64)
    got = hdr.read(GOLDEN)
    ref = py_decode(GOLDEN.read_bytes())
    np.testing.assert_array_equal(got, ref)
    assert np.isfinite(got).all() and (got[..., :3] >= 0).all() and got[..., :3].max() > 0


@pytest.mark.parametrize("w,h", [(128, 33), (7, 5), (8, 1), (300, 4)])
def test_write_read_round_trip(tmp_path, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    img = np.exp(rng.normal(0, 3, (h, w, 3))).astype(np.float32)
    img[:, : w // 2] = img[:, :1]  # long runs
    img[0, 0] = 0.0
    path = tmp_path / "x.hdr"
    hdr.write(path, img)
    back = hdr.read(path)
    assert back.shape == (h, w, 4) and np.all(back[..., 3] == 1.0)
    mx = img.max(-1, keepdims=True)
    assert np.all(np.abs(back[..., :3] - img) <= mx * 2.0 ** -7)
    assert np.all(back[0, 0, :3] == 0.0)
    # Values already on the RGBE grid come back exactly, and re-encoding is stable.
    hdr.write(path, back)
    np.testing.assert_array_equal(hdr.read(path), back)
    np.testing.assert_array_equal(py_decode(path.read_bytes()), back)


def test_error_codes(tmp_path):
    with pytest.raises(AltheaError) as e:
        hdr.info(tmp_path / "missing.hdr")
    assert e.value.status == 4  # ALTHEA_ERROR_NOT_FOUND
    bad = tmp_path / "bad.hdr"
    bad.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(AltheaError) as e:
        hdr.read(bad)
    assert e.value.status == 5  # ALTHEA_ERROR_IO
    trunc = tmp_path / "trunc.hdr"
    trunc.write_bytes(GOLDEN.read_bytes()[:3000])
    with pytest.raises(AltheaError):
        hdr.read(trunc)


def test_create_resources_missing_file_raises_not_found(tmp_path):
    import althea_amd as aa

    with pytest.raises(AltheaError) as e:
        aa.ImageBasedLighting.createResources(str(tmp_path / "NoSuchSkybox.hdr"))
    assert e.value.status == 4


@pytest.mark.gpu
def test_create_resources_from_file_equals_device_path(tmp_path):
    """The .hdr route precomputes exactly what the tensor route does on the
    decoded image."""
    import torch

    import althea_amd as aa

    env = aa.synth.env(seed=77, width=256, height=128)
    path = tmp_path / "env.hdr"
    hdr.write(path, env)
    decoded = hdr.read(path)
    desc = aa.IBLDesc(prefiltered_size=32, prefiltered_mips=3, prefilter_samples=64, irradiance_size=8,
                      brdf_lut_size=16, brdf_lut_samples=64)
    a = aa.ImageBasedLighting.createResources(str(path), desc)
    b = aa.ImageBasedLighting.createResources(torch.from_numpy(decoded).cuda(), desc)
    torch.cuda.synchronize()
    for name in ("prefiltered_tensor", "irradiance_tensor", "lut_tensor", "environment_tensor"):
        torch.testing.assert_close(getattr(a, name)(), getattr(b, name)(), rtol=0, atol=0)
