"""CPU pins of the oracle's screen-space reflection, glossy convolution and
reflection-buffer blend (no GPU).

- Convolution (Shaders/SSRGlossyConvolve.comp): constant in -> constant out;
  a linear ramp comes out as the ramp at the centre tap (uv = texel / extent,
  no half-texel offset); a unit impulse on a same-size level comes out as the
  7 linear taps built here from the trimmed binomial row C(16, 2..14) / 65502
  and the tent weights of the bilinear filter, including the vertical pass
  whose offsets are divided by the width (SSRGlossyConvolve.comp:42-52).
- Reflection-buffer read (DeferredPass.frag:29-31): level-constant buffers
  give the LINEAR-mipmap blend of the two levels; LOD 0 at a pixel centre is
  that texel.
- Blend (DeferredPass.frag:59-68): alpha 0 leaves the plain resolve; alpha 1
  replaces the prefiltered term by rgb; alpha 0.5 with rgb = c / 2 gives
  (term + c) / 2.
- March (Shaders/SSR.frag:80-149): the plane y = 0 reflects every ray upward,
  so nothing is hit; on a floor facing a wall, a floor pixel's hit lies on the
  wall at the analytic mirror intersection, within the march's step and cone,
  and the position rebuilt from the filtered depth is on the wall plane.
"""
import math

import numpy as np
import pytest

from althea_amd import synth
from tests import oracle_lib as ol

BINOM = np.array([math.comb(16, 8 + d) for d in range(-6, 7)], np.float64) / 65502.0


def _linear_taps():
    """(offset, weight) of the 7 linear taps folded from BINOM."""
    taps = [(0.0, BINOM[6])]
    for d1 in (1, 3, 5):
        b1, b2 = BINOM[6 + d1], BINOM[6 + d1 + 1]
        o = (d1 * b1 + (d1 + 1) * b2) / (b1 + b2)
        taps += [(o, b1 + b2), (-o, b1 + b2)]
    return taps


def _tent(t):
    return np.maximum(0.0, 1.0 - np.abs(t))


# ------------------------------------------------------------- convolution
def test_binomial_folding_is_a_normalised_kernel():
    taps = _linear_taps()
    assert abs(sum(w for _, w in taps) - 1.0) < 1e-15
    # The shader's constants (SSRGlossyConvolve.comp:42-52) are this folding.
    assert abs(taps[0][1] - 0.1964825501511404) < 1e-15
    assert abs(taps[1][0] - 1.411764705882353) < 1e-14 and abs(taps[1][1] - 0.2969069646728344) < 1e-15
    assert abs(taps[3][0] - 3.2941176470588234) < 1e-14 and abs(taps[3][1] - 0.09447039785044732) < 1e-15
    assert abs(taps[5][0] - 5.176470588235294) < 1e-14 and abs(taps[5][1] - 0.010381362401148057) < 1e-15


@pytest.mark.parametrize("vertical", [False, True])
def test_convolve_constant(vertical):
    src = np.full((24, 40, 4), 0.75, np.float32)
    out = ol.convolve_level(src, 20, 12, vertical)
    np.testing.assert_allclose(out, 0.75, rtol=0, atol=1e-15)


@pytest.mark.parametrize("vertical", [False, True])
def test_convolve_linear_ramp(vertical):
    sh, sw, dh, dw = 48, 64, 24, 32
    j, i = np.mgrid[0:sh, 0:sw].astype(np.float64)
    a, b, c = 0.37, -0.21, 5.0
    src = np.zeros((sh, sw, 4), np.float32)
    src[..., 0] = a * i + b * j + c
    src[..., 3] = 1.0
    out = ol.convolve_level(src, dw, dh, vertical)
    y, x = np.mgrid[0:dh, 0:dw].astype(np.float64)
    want = a * (x / dw * sw - 0.5) + b * (y / dh * sh - 0.5) + c
    # Interior: every tap inside [0, extent - 1] after the -0.5 of the filter.
    reach = 5.2 * (sh / dw if vertical else sw / dw)
    rx, ry = (0.0, reach) if vertical else (reach, 0.0)
    cx, cy = x / dw * sw - 0.5, y / dh * sh - 0.5
    ok = (cx - rx >= 0) & (cx + rx <= sw - 1) & (cy - ry >= 0) & (cy + ry <= sh - 1)
    assert ok.sum() > 100
    # The src values are FP32-rounded; allow their rounding.
    np.testing.assert_allclose(out[..., 0][ok], want[ok], rtol=0, atol=2e-5)
    np.testing.assert_allclose(out[..., 3], 1.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("vertical", [False, True])
def test_convolve_impulse_same_size(vertical):
    # Same-size level (sw = dw, sh = dh): the tap at uv + o / dw lands at
    # texel coordinate x + o s - 0.5 with s = 1 (x) or sh / dw (y).
    W, H = 32, 64
    x0, y0 = 15, 30
    src = np.zeros((H, W, 4), np.float32)
    src[y0, x0] = 1.0
    out = ol.convolve_level(src, W, H, vertical)
    taps = _linear_taps()
    s = H / W if vertical else 1.0
    want = np.zeros((H, W))
    for yy in range(H):
        for xx in range(W):
            if vertical:
                wx = _tent((xx - 0.5) - x0)  # every tap reads column coordinate xx - 0.5
                want[yy, xx] = wx * sum(w * _tent(yy + o * s - 0.5 - y0) for o, w in taps)
            else:
                wy = _tent((yy - 0.5) - y0)
                want[yy, xx] = wy * sum(w * _tent(xx + o * s - 0.5 - x0) for o, w in taps)
    np.testing.assert_allclose(out[..., 0], want, rtol=0, atol=1e-12)
    assert abs(out[..., 0].sum() - 1.0) < 1e-12


# ------------------------------------------------------ reflection buffer read
def _chain_offsets(W, H, mips):
    offs, o = [], 0
    for m in range(mips):
        offs.append(o)
        o += max(1, W >> m) * max(1, H >> m)
    return offs, o


def test_reflection_sample_level_constants():
    W, H, mips = 40, 24, 5
    offs, n = _chain_offsets(W, H, mips)
    chain = np.zeros((n, 4), np.float32)
    k = np.array([1.0, 2.0, 4.0, 8.0, 16.0])
    for m in range(mips):
        end = offs[m + 1] if m + 1 < mips else n
        chain[offs[m]:end] = k[m]
    rng = np.random.default_rng(3)
    pix = np.stack([rng.integers(0, W, 50), rng.integers(0, H, 50)], 1).astype(np.int32)
    lods = rng.uniform(-0.5, mips + 0.5, 50)
    out = ol.reflection_sample(chain, W, H, mips, pix, lods)
    d = np.clip(lods, 0, mips - 1)
    l0 = np.floor(d).astype(int)
    l1 = np.minimum(l0 + 1, mips - 1)
    want = k[l0] + (k[l1] - k[l0]) * (d - l0)
    np.testing.assert_allclose(out[:, 0], want, rtol=1e-14)


def test_reflection_sample_lod0_is_the_texel():
    W, H, mips = 40, 24, 3
    _, n = _chain_offsets(W, H, mips)
    chain = np.random.default_rng(4).uniform(0, 5, (n, 4)).astype(np.float32)
    pix = np.array([[0, 0], [39, 23], [17, 5], [3, 22]], np.int32)
    out = ol.reflection_sample(chain, W, H, mips, pix, np.zeros(4))
    np.testing.assert_array_equal(out, chain[pix[:, 1] * W + pix[:, 0]].astype(np.float64))


# ---------------------------------------------------------------- the blend
def _const_views(L=2.0, size=4, mips=3, lut_size=8):
    n = 6 * sum(max(1, size >> m) ** 2 for m in range(mips))
    pref = np.full((n, 4), L, np.float32)
    irr = np.full((6 * 4 * 4, 4), 0.5 * L, np.float32)
    lut = np.zeros((lut_size * lut_size, 2), np.float32)
    lut[:, 0], lut[:, 1] = 0.7, 0.1
    return {"prefiltered": pref, "pref_size": size, "pref_mips": mips, "irradiance": irr, "irr_size": 4,
            "lut": lut, "lut_size": lut_size}


@pytest.fixture(scope="module")
def small_scene():
    return synth.scene(seed=11, width=96, height=54, lights=3)


@pytest.fixture(scope="module")
def small_env():
    return ol.Env(synth.env(seed=12, width=64, height=32))


def test_blend_alpha_zero_alpha_one_alpha_half(small_scene, small_env):
    sc, mips = small_scene, 5
    views = _const_views()
    rng = np.random.default_rng(5)
    pix = np.stack([rng.integers(0, sc.width, 400), rng.integers(0, sc.height, 400)], 1).astype(np.int32)
    plain, _, _, q = ol.resolve(sc, views, small_env, pix, query=True)
    _, n = _chain_offsets(sc.width, sc.height, mips)
    # alpha 0 everywhere: untouched.
    zero = np.zeros((n, 4), np.float32)
    got, _, _ = ol.resolve(sc, views, small_env, pix, reflection=(zero, mips))
    np.testing.assert_array_equal(got, plain)
    # alpha 1, rgb c: the prefiltered term becomes c.
    c = np.array([3.0, 0.25, 1.5], np.float32)
    one = np.zeros((n, 4), np.float32)
    one[:, :3], one[:, 3] = c, 1.0
    got, _, _ = ol.resolve(sc, views, small_env, pix, reflection=(one, mips))
    want, _, _ = ol.resolve(sc, views, small_env, pix, refl=np.tile(c.astype(np.float64), (len(pix), 1)))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # alpha 1/2, rgb c/2: the term becomes (term + c) / 2.
    half = np.zeros((n, 4), np.float32)
    half[:, :3], half[:, 3] = c / 2, 0.5
    got, _, _ = ol.resolve(sc, views, small_env, pix, reflection=(half, mips))
    bg = np.isnan(q[:, 0])
    term = ol.cube_sample(views["prefiltered"], views["pref_size"], views["pref_mips"],
                          np.where(bg[:, None], 1.0, q[:, :3]), np.where(bg, 0.0, q[:, 3] * (views["pref_mips"] - 1)))
    mixed = 0.5 * term + 0.5 * c
    mixed[bg] = np.nan
    want, _, _ = ol.resolve(sc, views, small_env, pix, refl=mixed)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------------- the march
def test_ssr_plane_hits_nothing():
    sc = synth.scene(seed=13, width=120, height=68, lights=0, amplitude=0.0)
    views = _const_views()
    y, x = np.mgrid[0:sc.height, 0:sc.width]
    pix = np.stack([x.ravel(), y.ravel()], 1).astype(np.int32)
    rgba, _, step, _, _ = ol.ssr(sc, views, pix)
    covered = sc.position[..., 3].ravel() != 0
    assert covered.sum() > 1000
    assert np.all(step[covered] == -1) and np.all(step[~covered] == -2)
    assert np.all(rgba == 0.0)


def _mat(sc, k):
    return sc.mat(k).astype(np.float64).reshape(4, 4).T  # column-major -> [row, col]


def wall_scene(W=160, H=90, zw=-8.0, hw=3.0):
    """Floor y = 0 in front of a wall z = zw (normal +z, height hw), seen from
    (0, 1.5, 0); analytic G-buffer planes over a synth.scene camera."""
    sc = synth.scene(seed=14, width=W, height=H, lights=0, eye=(0.0, 1.5, 0.0), center=(0.0, 1.0, -8.0),
                     amplitude=0.0)
    ip, iv = _mat(sc, 1), _mat(sc, 4)
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    ndc = np.stack([(x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1, np.zeros_like(x), np.ones_like(x)], -1)
    e = ndc @ ip.T
    d = e[..., :3] @ iv[:3, :3].T
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    eye = iv[:3, 3]
    with np.errstate(divide="ignore", invalid="ignore"):
        tf = np.where(d[..., 1] < 0, -eye[1] / d[..., 1], np.inf)
        tw = np.where(d[..., 2] < 0, (zw - eye[2]) / d[..., 2], np.inf)
    pf = eye + tf[..., None] * d
    pw = eye + tw[..., None] * d
    floor_ok = np.isfinite(tf) & (pf[..., 2] > zw)
    wall_ok = np.isfinite(tw) & (pw[..., 1] >= 0) & (pw[..., 1] <= hw)
    use_w = wall_ok & (~floor_ok | (tw < tf))
    use_f = floor_ok & ~use_w
    pos = np.zeros((H, W, 4), np.float32)
    nrm = np.zeros((H, W, 4), np.float32)
    pos[use_f, :3], pos[use_f, 3], nrm[use_f, :3] = pf[use_f], 1.0, (0.0, 1.0, 0.0)
    pos[use_w, :3], pos[use_w, 3], nrm[use_w, :3] = pw[use_w], 1.0, (0.0, 0.0, 1.0)
    alb = np.zeros((H, W, 4), np.float32)
    mro = np.zeros((H, W, 4), np.float32)
    cov = use_f | use_w
    alb[cov] = (0.8, 0.6, 0.4, 1.0)
    mro[cov] = (0.0, 0.3, 1.0, 0.0)
    sc.position, sc.normal, sc.albedo, sc.material = pos, nrm, alb, mro
    return sc, use_f, use_w, d


def test_ssr_floor_pixels_hit_the_wall_at_their_mirror_point():
    zw, hw = -8.0, 3.0
    sc, use_f, use_w, d = wall_scene(zw=zw, hw=hw)
    views = _const_views()
    H, W = use_f.shape
    y, x = np.mgrid[0:H, 0:W]
    sel = use_f.ravel()
    pix = np.stack([x.ravel()[sel], y.ravel()[sel]], 1).astype(np.int32)
    rgba, _, step, uv, hpos = ol.ssr(sc, views, pix)
    P = sc.position.reshape(-1, 4)[sel, :3].astype(np.float64)
    dd = d.reshape(-1, 3)[sel]
    R = dd * np.array([1.0, -1.0, 1.0])  # mirror about the floor normal
    t = (zw - P[:, 2]) / R[:, 2]
    Wst = P + t[:, None] * R
    vp = _mat(sc, 0) @ _mat(sc, 2)
    clip = np.concatenate([Wst, np.ones((len(Wst), 1))], 1) @ vp.T
    uvw = 0.5 * clip[:, :2] / clip[:, 3:4] + 0.5
    on_wall = (Wst[:, 1] > 0.05) & (Wst[:, 1] < hw - 0.3) & np.all((uvw > 0.02) & (uvw < 0.98), 1)
    hit = step >= 0
    assert on_wall.sum() > 500
    # Most floor pixels whose mirror ray meets the visible wall find it ...
    assert hit[on_wall].mean() > 0.8, hit[on_wall].mean()
    # ... within about one march step of the analytic point, on the wall ...
    du = np.linalg.norm(uv[hit & on_wall] - uvw[hit & on_wall], axis=1)
    assert np.quantile(du, 0.99) < 0.0075 and du.max() < 0.012, (np.quantile(du, 0.99), du.max())
    # The hit point, rebuilt from the filtered depth, lies on the wall plane.
    assert np.all(np.abs(hpos[hit, 2] - zw) < 1e-6)
    assert np.all((hpos[hit, 1] > -1e-6) & (hpos[hit, 1] < hw + 1e-6))
    # ... and a hit is shaded (alpha 1, lit by the constant environment).
    assert np.all(rgba[hit, 3] == 1.0) and np.all(rgba[hit, :3] > 0.0)
    assert np.all(rgba[~hit] == 0.0)
    # Mirror rays that pass over the wall meet only sky: no hit.
    over = Wst[:, 1] > hw + 0.5
    assert over.sum() > 50 and not hit[over].any()

