"""CPU pins of the oracle's point-light shadow cubes and shadow test (no GPU).

Producer (Src/PointLight.cpp:235-282 with Shaders/ShadowMapBindless.*):
- no triangles: every texel keeps the clear value 1;
- a square below the light, facing it, covers its -Y face with the analytic
  distance h sqrt(1 + pc^2 + pt^2) / 1000 of the texel-centre ray;
- the same square wound the other way is culled; closer than the near plane
  0.01 or farther than the far plane 1000 (view depth) it is not written;
- random triangles against a brute-force Moller-Trumbore intersection of
  every texel-centre ray (front-facing = the normal (b - a) x (c - a) points
  at the light), texel by texel.

Consumer (PBRMaterial.glsl:119-140), through the oracle resolve over a plane:
- cubes of 1 (closest 1000) leave the plain resolve, cubes of 0 remove every
  light farther than 0.5;
- pixels whose ray to the light crosses an occluder on a side face are
  shadowed exactly there (the resolve without the light), the rest lit;
- on the -Y face the lookup (L.x, -L.y, -L.z) reads the reference's image
  turned by 180 degrees about the vertical: the shadow of an occluder below
  the light falls behind its mirror image (x, z) -> (-x, -z) about the light,
  and not behind the occluder.
"""
import numpy as np
import pytest

from althea_amd import synth
from tests import oracle_lib as ol
from tests.test_oracle_ssr import _const_views

FAR = 1000.0
NEAR = 0.01


def _square(center, u, v, facing):
    """Two triangles of the square center +- u +- v, wound so that the normal
    (b - a) x (c - a) has the sign of `facing` along u x v."""
    c = np.asarray(center, np.float64)
    u, v = np.asarray(u, np.float64), np.asarray(v, np.float64)
    pos = np.array([c - u - v, c + u - v, c + u + v, c - u + v], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    if facing < 0:
        idx = idx[:, ::-1].copy()
    return pos, idx


def _light(p, e=20.0):
    return np.array([[p[0], p[1], p[2], 0.0, e, e, e, 0.0]], np.float32)


def _turn(face, d):
    d = d.copy()
    if face in (2, 3):
        d[..., 2] *= -1.0
    else:
        d[..., 0] *= -1.0
    return d


def _face_rays(S):
    """Texel-centre ray directions R_f v of every texel (6, S, S, 3), major
    component +-1."""
    f, y, x = np.meshgrid(np.arange(6), np.arange(S), np.arange(S), indexing="ij")
    ids = np.stack([f.ravel(), x.ravel(), y.ravel()], 1)
    v = ol.cube_texel_dirs(S, ids).reshape(6, S, S, 3)
    v = v / np.abs(v).max(-1, keepdims=True)
    return np.stack([_turn(k, v[k]) for k in range(6)])


def test_no_triangles_clear_to_one():
    out = ol.shadow_cubes(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), _light((0, 1, 0)), 8)
    assert out.shape == (1, 6, 8, 8) and np.all(out == 1.0)


@pytest.mark.parametrize("h", [0.7, 3.0])
def test_square_below_the_light_gives_the_analytic_distance(h):
    S = 16
    # Square y = -h, normal +y (towards the light at the origin), large
    # enough to cover the whole -Y face (|x|, |z| <= h there).
    pos, idx = _square((0, -h, 0), (3 * h, 0, 0), (0, 0, -3 * h), +1)
    out = ol.shadow_cubes(pos, idx, _light((0, 0, 0)), S)[0]
    rays = _face_rays(S)
    d = rays[3]  # -Y face: d.y = -1
    want = float(np.float32(h)) * np.linalg.norm(d, axis=-1) / FAR  # positions are float32
    np.testing.assert_allclose(out[3], want, rtol=1e-12)
    assert np.all(out[2] == 1.0)  # nothing above


def test_reversed_winding_is_culled():
    pos, idx = _square((0, -1, 0), (3, 0, 0), (0, 0, -3), -1)
    out = ol.shadow_cubes(pos, idx, _light((0, 0, 0)), 16)
    assert np.all(out == 1.0)


def test_near_and_far_planes():
    S = 16
    rays = _face_rays(S)
    # h = 0.005: the -Y face's view depth 0.005 < 0.01, nothing written
    # there; a side texel looking down with slope s sees view depth h / s.
    h = float(np.float32(0.005))
    pos, idx = _square((0, -h, 0), (50, 0, 0), (0, 0, -50), +1)
    out = ol.shadow_cubes(pos, idx, _light((0, 0, 0)), S)[0]
    assert np.all(out[3] == 1.0)
    slope = -rays[0][..., 1]  # +X face, y of the ray at x = 1
    tau = np.where(slope > 0, h / np.where(slope > 0, slope, 1.0), np.inf)
    hit = (tau >= NEAR) & (tau <= 50.0 / 1.0) & np.isfinite(tau)
    # Inside the square (|x| <= 50 along the ray: tau <= 50).
    want = np.where(hit, np.minimum(tau * np.linalg.norm(rays[0], axis=-1) / FAR, 1.0), 1.0)
    np.testing.assert_allclose(out[0], want, rtol=1e-12)
    assert np.any(hit) and np.any(~hit & (slope > 0))
    # Beyond the far plane: not written.
    pos, idx = _square((0, -1500.0, 0), (5000, 0, 0), (0, 0, -5000), +1)
    assert np.all(ol.shadow_cubes(pos, idx, _light((0, 0, 0)), S)[0, 3] == 1.0)


def _brute_force(pos, idx, lp, S):
    rays = _face_rays(S).reshape(-1, 3)
    best = np.ones(rays.shape[0])
    for t in idx:
        a, b, c = (pos[t].astype(np.float64) - lp)
        e1, e2 = b - a, c - a
        n = np.cross(e1, e2)
        if not np.dot(n, -a) > 0.0:  # normal must point at the light (the origin)
            continue
        p = np.cross(rays, e2)
        det = p @ e1
        ok = np.abs(det) > 1e-300
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = -a
        u = (p @ tv) * inv
        q = np.cross(tv, e1)
        v = (rays @ q) * inv
        tau = (q @ e2) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (tau >= NEAR) & (tau <= FAR)
        depth = np.minimum(tau * np.linalg.norm(rays, axis=1) / FAR, 1.0)
        best = np.where(hit, np.minimum(best, depth), best)
    return best.reshape(6, S, S)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_triangles_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    S, T = 24, 40
    lp = np.array([0.3, -0.2, 0.1], np.float32).astype(np.float64)  # lights are float32
    centres = rng.normal(0.0, 3.0, (T, 1, 3))
    pos = (centres + rng.normal(0.0, 1.5, (T, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    got = ol.shadow_cubes(pos, idx, _light(lp), S)[0]
    want = _brute_force(pos, idx, lp, S)
    assert np.mean(want < 1.0) > 0.2
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=0)


# ------------------------------------------------- glTF (alpha-tested) pass
def _gltf_material(factor_alpha=1.0, cutoff=0.5):
    import ctypes

    from althea_amd import _lib as L
    return L.GltfMaterialC((ctypes.c_float * 4)(1.0, 1.0, 1.0, factor_alpha), 1.0, 1.0, 1.0, cutoff, 0, 0, 0, 0, 0)


def _alpha_texture(alpha, nearest=True):
    """RGBA8 texture (one level) whose alpha channel is `alpha` (h, w)."""
    h, w = alpha.shape
    lvl = np.full((h, w, 4), 200, np.uint8)
    lvl[..., 3] = alpha
    f = 0 if nearest else 1
    return ol.OracleTex(lvl.reshape(-1, 4), w, h, 1, srgb=0, mag=f, min_=f, mipmode=0)


def _gltf_prim(pos, idx, uv):
    return {"pos": pos, "tan": None, "bit": None, "nrm": None, "uvs": [uv], "idx": idx, "material": 0}


def _random_tris(seed, T=40):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 3.0, (T, 1, 3))
    pos = (centres + rng.normal(0.0, 1.5, (T, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    uv = rng.uniform(0.0, 1.0, (3 * T, 2)).astype(np.float32)
    return pos, idx, uv


@pytest.mark.parametrize("alpha,factor,kept", [(255, 1.0, True), (0, 1.0, False), (200, 0.5, False),
                                               (200, 0.7, True)])
def test_gltf_uniform_alpha_keeps_or_drops_every_fragment(alpha, factor, kept):
    """A texture of one alpha value: alpha / 255 x the factor's alpha against
    the 0.5 cut-off keeps every fragment (the cubes of the plain mesh pass) or
    none (every texel cleared to 1): 200/255 x 0.5 = 0.392 is cut, x 0.7 =
    0.549 is kept."""
    S = 16
    lp = np.array([0.3, -0.2, 0.1])
    pos, idx, uv = _random_tris(5)
    tex = _alpha_texture(np.full((4, 4), alpha, np.uint8))
    got, margin = ol.shadow_cubes_gltf([_gltf_prim(pos, idx, uv)], [_gltf_material(factor)], [tex], _light(lp), S)
    plain = ol.shadow_cubes(pos, idx, _light(lp), S)
    assert np.mean(plain < 1.0) > 0.1
    if kept:
        np.testing.assert_array_equal(got, plain)
    else:
        assert np.all(got == 1.0)
    # Every texel with a candidate fragment had its alpha decision taken.
    np.testing.assert_array_equal(margin[0] < 1e30, plain[0] < 1.0)


def test_gltf_checkerboard_cuts_the_analytic_squares():
    """A square below the light covering its -Y face, textured with a 2 x 2
    NEAREST checkerboard of alpha 0 and 255 (magnified: one texture texel
    spans many face texels).  Each -Y texel's ray meets the plane y = -h at
    P = d h / |d.y|; uv is P's affine coordinate across the square, and the
    texel is written (with the analytic distance) exactly where floor(2 uv)
    names an opaque texture texel."""
    S, h = 32, 1.0
    c = np.array([0.123, -h, -0.217])
    u, v = np.array([3.0, 0, 0]), np.array([0, 0, -3.0])
    pos, idx = _square(c, u, v, +1)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    board = np.array([[0, 255], [255, 0]], np.uint8)  # row j, column i
    tex = _alpha_texture(board)
    got, margin = ol.shadow_cubes_gltf([_gltf_prim(pos, idx, uv)], [_gltf_material()], [tex], _light((0, 0, 0)), S)
    d = _face_rays(S)[3]
    p = d * (np.float32(h) / np.abs(d[..., 1:2]))
    corner = pos[0].astype(np.float64)
    uu = (p - corner) @ (2 * u) / np.dot(2 * u, 2 * u)
    vv = (p - corner) @ (2 * v) / np.dot(2 * v, 2 * v)
    ti, tj = np.floor(2 * uu).astype(int), np.floor(2 * vv).astype(int)
    opaque = board[tj % 2, ti % 2] == 255
    clear = (np.abs(2 * uu - np.round(2 * uu)) > 1e-6) & (np.abs(2 * vv - np.round(2 * vv)) > 1e-6)
    want = np.where(opaque, float(np.float32(h)) * np.linalg.norm(d, axis=-1) / FAR, 1.0)
    assert 0.3 < opaque.mean() < 0.7 and clear.mean() > 0.9
    np.testing.assert_allclose(got[0, 3][clear], want[clear], rtol=1e-12)
    assert np.all(got[0, 2] == 1.0)


def test_gltf_derivatives_pick_the_mip_level():
    """The mip level comes from the texture coordinate's differences to the
    next texel right and below.  A two-level texture whose level 0 is opaque
    and level 1 transparent (LINEAR, nearest mip): a quad whose uv spans few
    texture texels per face texel reads level 0 (kept); with the uv scaled up
    so that one face texel spans more than 1.5 texture texels, level 1
    (cut)."""
    S, h = 16, 1.0
    pos, idx = _square((0.05, -h, 0.07), (3.0, 0, 0), (0, 0, -3.0), +1)
    lvl = np.full((8 * 8 + 4 * 4, 4), 200, np.uint8)
    lvl[:64, 3] = 255
    lvl[64:, 3] = 0
    tex = ol.OracleTex(lvl, 8, 8, 2, srgb=0, mag=1, min_=1, mipmode=0)
    out = {}
    for scale in (1.0, 40.0):
        uv = (np.array([[0, 0], [1, 0], [1, 1], [0, 1]]) * scale).astype(np.float32)
        got, _ = ol.shadow_cubes_gltf([_gltf_prim(pos, idx, uv)], [_gltf_material()], [tex], _light((0, 0, 0)), S)
        out[scale] = got[0, 3]
    # One face texel at the -Y face centre spans 2h/S = 0.125 of the 6-wide
    # square: 0.125 / 6 x 8 = 0.17 texture texels (magnified) at scale 1 and
    # 6.7 (level 1 of the 2) at scale 40.
    assert np.all(out[1.0] < 1.0)
    assert np.all(out[40.0] == 1.0)


# ---------------------------------------------------------------- consumer
@pytest.fixture(scope="module")
def plane_scene():
    sc = synth.scene(seed=31, width=240, height=135, lights=1, amplitude=0.0, eye=(0.0, 9.0, 4.0))
    sc.ssao_enabled = False
    return sc


@pytest.fixture(scope="module")
def env():
    return ol.Env(synth.env(seed=32, width=64, height=32))


LP = np.array([0.0, 3.0, 0.0])


def _resolve(sc, env, lights, shadows=None):
    sc.lights = lights
    pix = np.stack(np.meshgrid(np.arange(sc.width), np.arange(sc.height)), -1).reshape(-1, 2).astype(np.int32)
    rgb, _, margin = ol.resolve(sc, _const_views(), env, pix, shadows=shadows)
    return rgb, margin


def test_constant_cubes(plane_scene, env):
    sc = plane_scene
    lit, _ = _resolve(sc, env, _light(LP))
    dark, _ = _resolve(sc, env, np.zeros((0, 8), np.float32))
    got1, _ = _resolve(sc, env, _light(LP), np.ones((1, 6, 4, 4), np.float32))
    got0, _ = _resolve(sc, env, _light(LP), np.zeros((1, 6, 4, 4), np.float32))
    np.testing.assert_array_equal(got1, lit)
    np.testing.assert_array_equal(got0, dark)
    assert np.abs(lit - dark).max() > 1e-3


def _crossing(P, lp, plane_axis, plane_value):
    """Point where the segment P -> lp crosses plane (axis = value)."""
    den = lp[plane_axis] - P[:, plane_axis]
    s = (plane_value - P[:, plane_axis]) / np.where(den == 0.0, 1e-30, den)
    return P + (lp - P) * s[:, None], s


def _check_shadow(sc, env, pos, idx, region, S=256):
    """Resolve over the plane with the oracle's cubes of (pos, idx) for one
    light at LP: pixels whose segment to the light enters `region` (a mask
    function of the plane pixels, True / False / None for the edge band) must
    equal the unlit resolve, the others the lit one."""
    cubes = ol.shadow_cubes(pos, idx, _light(LP), S).astype(np.float32)
    lit, _ = _resolve(sc, env, _light(LP))
    dark, _ = _resolve(sc, env, np.zeros((0, 8), np.float32))
    got, _ = _resolve(sc, env, _light(LP), cubes)
    P = sc.position.reshape(-1, 4)
    covered = P[:, 3] != 0
    inside, outside = region(P[:, :3].astype(np.float64))
    inside &= covered
    outside &= covered
    assert inside.sum() > 100 and outside.sum() > 1000, (inside.sum(), outside.sum())
    np.testing.assert_array_equal(got[inside], dark[inside])
    np.testing.assert_array_equal(got[outside], lit[outside])
    return inside


def test_side_face_occluder_shadows_behind_itself(plane_scene, env):
    # Vertical square x = 2, y in [1, 2], z in [-0.6, 0.6], facing the light
    # (normal -x): every ray that meets it is on the +X face.
    pos, idx = _square((2.0, 1.5, 0.0), (0, 0.5, 0), (0, 0, 0.6), -1)
    a, b, c = pos[idx[0]]
    assert np.cross(b - a, c - a)[0] < 0.0

    def region(P):
        X, s = _crossing(P, LP, 0, 2.0)
        ok = (P[:, 0] > 2.0) & (s > 0) & (s < 1)
        dy, dz = np.abs(X[:, 1] - 1.5), np.abs(X[:, 2])
        inside = ok & (dy < 0.5 - 0.05) & (dz < 0.6 - 0.05)
        outside = ~ok | (dy > 0.5 + 0.05) | (dz > 0.6 + 0.05)
        return inside, outside

    _check_shadow(plane_scene, env, pos, idx, region)


def test_minus_y_face_reads_the_mirrored_image(plane_scene, env):
    # Horizontal square at y = 1 centred (1.2, 1, 0.8) about the light's
    # vertical, half-size 0.5, facing up: on the -Y face (|dy| = 2 > |dx|).
    c = np.array([1.2, 1.0, 0.8])
    pos, idx = _square(c, (0.5, 0, 0), (0, 0, -0.5), +1)

    def region_at(cx, cz):
        def region(P):
            X, s = _crossing(P, LP, 1, 1.0)
            ok = (s > 0) & (s < 1)
            dx, dz = np.abs(X[:, 0] - cx), np.abs(X[:, 2] - cz)
            inside = ok & (dx < 0.5 - 0.03) & (dz < 0.5 - 0.03)
            outside = ~ok | (dx > 0.5 + 0.03) | (dz > 0.5 + 0.03)
            return inside, outside
        return region

    # The shadow falls behind the mirror image (-1.4, 1, -0.9) ...
    _check_shadow(plane_scene, env, pos, idx, region_at(-c[0], -c[2]))
    # ... and the pixels behind the occluder itself stay lit.
    P = plane_scene.position.reshape(-1, 4)[:, :3].astype(np.float64)
    inside_true, _ = region_at(c[0], c[2])(P)
    cubes = ol.shadow_cubes(pos, idx, _light(LP), 256).astype(np.float32)
    lit, _ = _resolve(plane_scene, env, _light(LP))
    got, _ = _resolve(plane_scene, env, _light(LP), cubes)
    sel = inside_true & (plane_scene.position.reshape(-1, 4)[:, 3] != 0)
    assert sel.sum() > 100
    np.testing.assert_array_equal(got[sel], lit[sel])
