"""CPU tests pinning the ray-tracing oracle (a restatement of
rt_final.cpp's image) against closed forms computed here in float64 from the
geometry alone: the sky gradient of an empty scene, a perfect mirror floor
(sky of the reflected direction), a partial mirror with albedo != 1 at bounce
limits 0, 1 and 2 (the split of rt_final.cpp:581-609), two facing partial
mirrors (the attenuation compounded as (albedo * refl)^k), an analytic sphere
hit, the affine dependence of a Lambertian image on albedo, and the ray count
bound.  A copy of the oracle built with a planted mistake in the reflection
attenuation (albedo dropped) must fail the reflective checks."""
import numpy as np
import pytest

LIGHT = np.array([-0.5, 1.0, 0.2])
AMB = 0.05
SKY_HI = np.array([0.5, 0.7, 1.0])


def _camera_dirs(W, H, cam_f=(0, -0.2, -1.0)):
    f = np.asarray(cam_f, float)
    f = f / np.linalg.norm(f)
    r, u = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    a = W / H
    uu = ((px + 0.5) / W - 0.5) * 2 * a
    vv = -((py + 0.5) / H - 0.5) * 2
    d = f[None, None] + uu[..., None] * r + vv[..., None] * u
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _sky(dy):
    t = 0.5 * (dy + 1)
    return np.clip((1 - t)[..., None] + t[..., None] * SKY_HI, 0, 1)


def _render(orc, spheres, planes, W=160, H=90, cam_o=(0, 0.3, 0.5), cam_f=(0, -0.2, -1), max_depth=4, lib=None):
    cam = np.array(list(cam_o) + list(cam_f) + [1, 0, 0, 0, 1, 0], np.float32)
    sph = np.asarray(spheres, np.float32).reshape(-1, 8)
    pl = np.asarray(planes, np.float32).reshape(-1, 8)
    return orc.rt_render(sph, pl, cam, LIGHT.astype(np.float32), np.full(3, AMB, np.float32), W, H,
                         max_depth=max_depth, threads=2, lib=lib)


def test_empty_scene_is_sky(orc):
    img, rays = _render(orc, [], [])
    d = _camera_dirs(160, 90)
    assert rays == 160 * 90
    np.testing.assert_allclose(img, _sky(d[..., 1]), atol=2e-6)


def test_perfect_mirror_floor(orc):
    # Floor y = -1 with albedo 1, refl 1: a hit adds clamp(contrib * 0) = 0
    # and the reflected ray (d.x, -d.y, d.z) sees the sky.
    floor = [0, 1, 0, 1.0, 1, 1, 1, 1.0]
    img, rays = _render(orc, [], [floor])
    d = _camera_dirs(160, 90)
    down = d[..., 1] < -1e-3
    expect = np.where(down[..., None], _sky(-d[..., 1]), _sky(d[..., 1]))
    np.testing.assert_allclose(img, expect, atol=2e-6)
    assert rays == 160 * 90 + int(down.sum())
    # Without bounces the floor is terminal: ambient + albedo * (n . L).
    img0, rays0 = _render(orc, [], [floor], max_depth=0)
    lam = AMB + LIGHT[1] / np.linalg.norm(LIGHT)
    np.testing.assert_allclose(img0[down], np.full((int(down.sum()), 3), min(lam, 1.0)), atol=2e-6)
    assert rays0 == 160 * 90


@pytest.mark.parametrize("cx", [0.0, 0.6])
def test_single_lambertian_sphere(orc, cx):
    c, rad, alb = np.array([cx, 0.3, -3.0]), 0.8, np.array([0.7, 0.4, 0.2])
    img, _ = _render(orc, [[*c, rad, *alb, 0.0]], [], cam_f=(0, 0, -1))
    d = _camera_dirs(160, 90, cam_f=(0, 0, -1))
    o = np.array([0, 0.3, 0.5])
    oc = o - c
    b = d @ oc
    disc = b * b - (oc @ oc - rad * rad)
    inside = disc > 0.02  # away from the silhouette
    t = -b - np.sqrt(np.maximum(disc, 0))
    p = o + d * t[..., None]
    n = (p - c) / rad
    L = LIGHT / np.linalg.norm(LIGHT)
    shade = np.clip(AMB + alb[None, None] * np.maximum(0, n @ L)[..., None], 0, 1)
    np.testing.assert_allclose(img[inside], shade[inside], atol=2e-5)
    miss = disc < -0.02
    np.testing.assert_allclose(img[miss], _sky(d[..., 1])[miss], atol=2e-6)
    # the sphere is where the geometry puts it (catches a flipped u or v)
    assert inside.sum() > 100
    assert np.all(np.abs(img[inside] - _sky(d[..., 1])[inside]).max(axis=-1) > 1e-3)


def test_lambertian_image_is_affine_in_albedo(orc):
    # refl = 0 everywhere: pixel = ambient + albedo * max(0, n.L), so
    # doubling every albedo doubles (pixel - ambient) on hit pixels.
    rng = np.random.default_rng(1)
    sph = [[rng.uniform(-1, 1), rng.uniform(-0.5, 0.5), rng.uniform(-5, -2), rng.uniform(0.2, 0.6),
            *rng.uniform(0.05, 0.45, 3), 0.0] for _ in range(6)]
    floor = [0, 1, 0, 1.0, 0.3, 0.3, 0.3, 0.0]
    a, _ = _render(orc, sph, [floor])
    sph2 = [s[:4] + [2 * x for x in s[4:7]] + [0.0] for s in sph]
    floor2 = floor[:4] + [0.6, 0.6, 0.6, 0.0]
    b, _ = _render(orc, sph2, [floor2])
    d = _camera_dirs(160, 90)
    hit = np.abs(a - _sky(d[..., 1])).max(axis=-1) > 1e-4
    assert hit.sum() > 1000
    np.testing.assert_allclose(b[hit] - AMB, 2 * (a[hit] - AMB), atol=3e-6)


def test_ray_count_bounded_by_depth(orc, gcq):
    from gcq_amd import rt
    with rt.Scene.preset(rt.CORNELL) as s:
        arrs = s.arrays()
    for md in (0, 1, 4):
        img, rays = orc.rt_render(*arrs, 320, 180, max_depth=md, threads=2)
        assert 320 * 180 <= rays <= (md + 1) * 320 * 180
        assert np.all(img >= 0) and np.all(img <= md + 1)


def test_presets_match_reference_scene_sizes(gcq):
    from gcq_amd import rt
    with rt.Scene.preset(rt.COMPLEX) as s:
        sph, pl, cam, li, am = s.arrays()
        assert sph.shape == (103, 8) and pl.shape == (1, 8)  # rt_final.cpp:650-671
        assert np.allclose(cam[:6], [0, 0.3, 0.5, 0, -0.2, -1])
    with rt.Scene.preset(rt.CORNELL) as s:
        sph, pl, cam, li, am = s.arrays()
        assert sph.shape == (2, 8) and pl.shape == (5, 8)  # :673-684
    with rt.Scene.preset(rt.RANDOM, 64, 3) as s:
        sph, pl, *_ = s.arrays()
        assert sph.shape == (64, 8) and pl.shape == (1, 8)
        assert np.all((sph[:, 3] >= 0.2) & (sph[:, 3] <= 1.5))
        assert np.allclose(sph[:, 1], -1 + sph[:, 3])  # resting on the plane y = -1
        mirrors = np.sum(sph[:, 7] > 0)
        assert 8 <= mirrors <= 32 and set(np.unique(sph[:, 7]).tolist()) <= {0.0, float(np.float32(0.8))}


# ---------------------------------------------------------------- partial mirrors
# rt_final.cpp:581-614: a hit adds clamp01(contrib * (1 - refl)) and sends the
# reflected ray on with atten * albedo * refl while depth < max_depth;
# otherwise it adds clamp01(contrib) and ends.  contrib = (ambient + albedo *
# max(0, n.L)) * atten.
ALB = np.array([0.6, 0.5, 0.4])
REFL = 0.35
NL_UP = LIGHT[1] / np.linalg.norm(LIGHT)  # n.L for the upward normal (0, 1, 0)


def _partial_floor_expect(d, max_depth, alb=ALB, refl=REFL):
    dy = d[..., 1]
    lam = AMB + alb * NL_UP
    if max_depth == 0:
        hit = np.broadcast_to(np.clip(lam, 0, 1), d.shape)
    else:  # the reflected ray (d.x, -d.y, d.z) leaves the floor and sees the sky
        hit = np.clip(lam * (1 - refl), 0, 1) + np.clip(_sky(-dy) * alb * refl, 0, 1)
    return np.where((dy < 0)[..., None], hit, _sky(dy))


def _partial_floor_err(render, max_depth):
    floor = [0, 1, 0, 1.0, *ALB, REFL]
    img, rays = render([], [floor], max_depth=max_depth)
    d = _camera_dirs(160, 90)
    ok = np.abs(d[..., 1]) > 1e-3
    err = np.abs(img - _partial_floor_expect(d, max_depth))[ok].max()
    down = int((d[..., 1] < 0).sum())
    return err, rays, down


@pytest.mark.parametrize("max_depth", [0, 1, 2])
def test_partial_mirror_floor(orc, max_depth):
    err, rays, down = _partial_floor_err(lambda *a, **k: _render(orc, *a, **k), max_depth)
    assert err < 3e-6, err
    # one bounce at most: the reflected ray escapes, so depth 2 equals depth 1
    assert rays == 160 * 90 + (down if max_depth >= 1 else 0)


# Two facing partial mirrors, floor y = -1 (normal up) and ceiling y = +1
# (normal down, n.L < 0 so only ambient), camera between them: every ray that
# is not near-horizontal bounces between them until max_depth.  The k-th hit
# adds clamp01((ambient + alb * nl_k) * (alb * refl)^k * (1 - refl)), the
# last one clamp01((ambient + alb * nl_m) * (alb * refl)^m).
ALB2 = np.array([0.9, 0.7, 0.5])
REFL2 = 0.6


def _two_mirrors_expect(d, max_depth):
    dy = d[..., 1]
    out = np.zeros(d.shape)
    for k in range(max_depth + 1):
        on_floor = (dy < 0) if k % 2 == 0 else (dy > 0)  # the first hit is the plane the ray points at
        nl = np.where(on_floor, NL_UP, 0.0)[..., None]
        contrib = (AMB + ALB2 * nl) * (ALB2 * REFL2) ** k
        out += np.clip(contrib * ((1 - REFL2) if k < max_depth else 1.0), 0, 1)
    return out


def _two_mirrors_err(render, max_depth=4):
    floor = [0, 1, 0, 1.0, *ALB2, REFL2]
    ceiling = [0, -1, 0, 1.0, *ALB2, REFL2]
    img, rays = render([], [floor, ceiling], cam_o=(0, 0, 0.5), cam_f=(0, -0.2, -1), max_depth=max_depth)
    d = _camera_dirs(160, 90)
    ok = np.abs(d[..., 1]) > 0.05  # hit points close enough that the 1e-4 offset clears the plane
    err = np.abs(img - _two_mirrors_expect(d, max_depth))[ok].max()
    return err, rays, int(ok.sum())


def test_two_facing_mirrors(orc):
    err, rays, n_ok = _two_mirrors_err(lambda *a, **k: _render(orc, *a, **k))
    assert n_ok > 5000
    assert err < 5e-6, err
    # every non-horizontal pixel traces max_depth + 1 rays
    assert rays >= n_ok * 5


def test_planted_attenuation_mistake_is_caught(orc, tmp_path):
    """The same checks against a copy of the oracle whose reflection ray
    carries atten * refl (albedo dropped) must fail: the partial floor at one
    bounce, and the compounded attenuation of the facing mirrors."""
    bad = orc.build_rt_variant(tmp_path / "librt_bad.so", "PLANT_RT_ATTEN_NO_ALBEDO")

    def render(*a, **k):
        return _render(orc, *a, lib=bad, **k)

    err1, _, _ = _partial_floor_err(render, 1)
    err2, _, _ = _two_mirrors_err(render)
    assert err1 > 1e-3, err1
    assert err2 > 1e-3, err2
    # depth 0 has no reflection ray, so it cannot tell the two apart
    err0, _, _ = _partial_floor_err(render, 0)
    assert err0 < 3e-6
