"""Textures on the CPU checker: a white or uniform texture changes no bit against the equivalent
untextured material, bilinear sampling reproduces a checkerboard's cells, an sRGB-encoded texture is
decoded by IEC 61966-2-1, a normal map tilts the shading normal by its encoded angle (closed-form
irradiance), an emissive texture scales the emitted radiance; on the GPU, textured renders match the
checker for all three integrators."""
import math

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

NEE, BPT, BDPT = _lib.YPT_INTEGRATOR_NEE, _lib.YPT_INTEGRATOR_BPT, _lib.YPT_INTEGRATOR_BDPT


def _plane_scene(albedo=0.5, tex=None, slot="albedo_tex", light=None, emissive=None, uv_scale=1.0):
    b = scenes.MeshBuilder()
    m = scenes.preset_material("PLASTIC")
    m.albedo[:] = (albedo,) * 3
    m.specular_amount = 0.0
    m.roughness = 1.0
    if emissive is not None:
        m.emissive[:] = (emissive,) * 3
    if tex is not None:
        setattr(m, slot, 0)
        getattr(m, slot + "_scale")[:] = (uv_scale, uv_scale)
    b.quad((-1, 0, 1), (2, 0, 0), (0, 0, -2), b.material(m))      # +y, uv (0,0) at (-1, 0, 1)
    uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    lights = [light] if light is not None else []
    return b.arrays(lights, uvs=uvs, textures=[tex] if tex is not None else [])


def _sun(theta=0.0, intensity=2.0):
    l = _lib.Light(_lib.Light.DIRECTIONAL)
    l.lookAt((0, 5, 0), (math.sin(theta), 5 - math.cos(theta), 0))
    l.setIntensity(intensity)
    return l


def _film(oracle, scene, integ=NEE, w=32, h=32, spp=2, depth=1, eye=(0, 3, 0), target=(0, 0, 0), fovy=0.55):
    cam = camera.Camera(eye=eye, target=target, up=(0, 0, -1), fovy=fovy)
    fr = camera.make_frame(cam, w, h, max_depth=depth, firefly_clamp_scale=0.0, flags=_lib.YPT_FLAG_DISABLE_JITTER)
    return oracle.render(scene, fr, integ, 0, spp, use_bvh=False)["film"]


def _tex(value, w=4, h=4, srgb=True):
    t = np.zeros((h, w, 4), np.uint8)
    t[...] = value
    return (t, srgb)


@pytest.mark.parametrize("srgb", [True, False])
def test_uniform_texture_equals_the_untextured_material(oracle, srgb):
    lut = (128 / 255.0 + 0.055) / 1.055
    val = np.float32(lut ** 2.4) if srgb else np.float32(128) / np.float32(255)
    a = _film(oracle, _plane_scene(0.8, _tex(128, srgb=srgb), light=_sun()))
    b = _film(oracle, _plane_scene(float(np.float32(0.8) * val), light=_sun()))
    assert a[..., 1].max() > 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    c = _film(oracle, _plane_scene(0.8, _tex(255, srgb=srgb), light=_sun()))
    d = _film(oracle, _plane_scene(0.8, light=_sun()))
    assert np.array_equal(c.view(np.uint32), d.view(np.uint32))


def test_checkerboard_texture_is_sampled_per_cell(oracle):
    # 8x8 texels in 2x2 blocks of 4x4: bilinear filtering is uniform inside each block, so each block's
    # pixels match an untextured plane of albedo 0.5 x texel on the same pixels and wavelengths
    t = np.full((8, 8, 4), 255, np.uint8)
    t[:4, 4:, :3] = 64
    t[4:, :4, :3] = 64
    lo = np.float32(64) / np.float32(255)
    fovy = 2 * math.atan(1 / 3)                                     # the view covers the plane exactly
    f = _film(oracle, _plane_scene(0.5, (t, False), light=_sun()), w=64, h=64, fovy=fovy)[..., :3]
    white = _film(oracle, _plane_scene(0.5, light=_sun()), w=64, h=64, fovy=fovy)[..., :3]
    dark = _film(oracle, _plane_scene(float(np.float32(0.5) * lo), light=_sun()), w=64, h=64, fovy=fovy)[..., :3]
    # image right is +x (u), image up is -z (v): the top-left block is u < 1/2, v > 1/2, which is
    # texel rows 4..7 (row index grows with v, as glTF lays images out), columns 0..3
    blocks = {(0, 0): dark, (1, 1): dark, (0, 1): white, (1, 0): white}
    for (r, c), want in blocks.items():
        sl = (slice(32 * r + 10, 32 * r + 22), slice(32 * c + 10, 32 * c + 22))
        assert np.allclose(f[sl], want[sl], rtol=1e-6, atol=0), (r, c)
    # wrap: uv scale 2 repeats the pattern twice across the plane
    g = _film(oracle, _plane_scene(0.5, (t, False), light=_sun(), uv_scale=2.0), w=64, h=64, fovy=fovy)[..., :3]
    for r in range(4):
        for c in range(4):
            sl = (slice(16 * r + 5, 16 * r + 11), slice(16 * c + 5, 16 * c + 11))
            want = dark if (r + c) % 2 == 0 else white
            assert np.allclose(g[sl], want[sl], rtol=1e-6, atol=0), (r, c)


def test_normal_map_tilts_the_shading_normal(oracle):
    theta = 0.6
    n = np.array([math.sin(theta), 0.0, math.cos(theta)])            # tangent space: x = tangent, z = normal
    texel = np.round((n * 0.5 + 0.5) * 255).astype(np.uint8)
    dec = texel / 255.0 * 2 - 1
    dec = dec / np.linalg.norm(dec)
    flat = _film(oracle, _plane_scene(0.5, light=_sun()))[..., 1]
    t = np.zeros((4, 4, 4), np.uint8)
    t[..., :3] = texel
    t[..., 3] = 255
    tilted = _film(oracle, _plane_scene(0.5, (t, False), slot="normal_tex", light=_sun()))[..., 1]
    # the sun is straight above: a diffuse surface receives cos(angle to the mapped normal)
    ratio = tilted.mean() / flat.mean()
    assert abs(ratio - dec[2]) < 2e-3, (ratio, dec[2])


def test_emissive_texture_scales_the_emission(oracle):
    plain = _film(oracle, _plane_scene(0.0, emissive=2.0))[..., 1]
    tex = _film(oracle, _plane_scene(0.0, _tex(200, srgb=True), slot="emissive_tex", emissive=2.0))[..., 1]
    want = ((200 / 255.0 + 0.055) / 1.055) ** 2.4
    assert abs(tex.mean() / plain.mean() - want) < 1e-4


def _textured_scene():
    b = scenes.MeshBuilder()
    rng = np.random.default_rng(0)
    albedo = rng.integers(0, 256, (16, 16, 4)).astype(np.uint8)
    normal = np.zeros((8, 8, 4), np.uint8)
    normal[..., 0] = rng.integers(96, 160, (8, 8))
    normal[..., 1] = rng.integers(96, 160, (8, 8))
    normal[..., 2] = 240
    orm = rng.integers(64, 256, (8, 8, 4)).astype(np.uint8)
    emis = rng.integers(0, 256, (4, 4, 4)).astype(np.uint8)
    m = scenes.preset_material("PLASTIC")
    m.albedo[:] = (0.9, 0.7, 0.5)
    m.roughness = 0.6
    m.albedo_tex, m.normal_tex, m.orm_tex = 0, 1, 2
    m.albedo_tex_scale[:] = (2.0, 3.0)
    g = scenes.preset_material("METAL_COPPER")
    g.roughness = 0.5
    g.orm_tex = 2
    e = scenes.blank_material()
    e.emissive[:] = (8.0, 6.0, 4.0)
    e.emissive_tex = 3
    fm, gm, em = b.material(m), b.material(g), b.material(e)
    b.quad((-3, 0, 3), (6, 0, 0), (0, 0, -6), fm)
    b.quad((-3, 0, -3), (6, 0, 0), (0, 4, 0), gm)
    b.quad((-0.8, 3.5, -0.8), (0, 0, 1.6), (1.6, 0, 0), em)    # emitter facing down
    uv = np.tile(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), (3, 1))
    scene = b.arrays(uvs=uv, textures=[(albedo, True), (normal, False), (orm, False), (emis, True)])
    cam = camera.Camera(eye=(0, 3, 6), target=(0, 1, 0), fovy=0.9)
    return scene, cam


def test_textured_scene_renders_on_the_checker(oracle):
    scene, cam = _textured_scene()
    fr = camera.make_frame(cam, 32, 24, max_depth=4)
    for integ in (NEE, BDPT):
        f = oracle.render(scene, fr, integ, 0, 2)["film"]
        assert np.isfinite(f).all() and f[..., 1].max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [BPT, NEE, BDPT])
def test_textured_render_matches_the_checker(ylib, oracle, integrator):
    scene, cam = _textured_scene()
    frame = camera.make_frame(cam, 64, 48, max_depth=5, light_path_vertices=5)
    it = ylib.Integrator(integrator, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 3)
    g_last, g = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(scene, frame, integrator, 0, 3)
    assert np.array_equal(g_last[..., 3].view(np.uint32), ref["last"][..., 3].view(np.uint32))
    assert np.allclose(g[..., :3], ref["film"][..., :3], rtol=1e-4, atol=1e-6)


# ---------------------------------------------------------------- alpha testing
CUT = 0.5


def _alpha_texture():
    """8 x 8 texels in 2 x 2 blocks: alpha 255 in the diagonal blocks, 0 (rows 0-3, columns 4-7) and
    100 (rows 4-7, columns 0-3, below the cutoff) in the others"""
    t = np.full((8, 8, 4), 255, np.uint8)
    t[:4, 4:, 3] = 0
    t[4:, :4, 3] = 100
    return t


def _alpha_scene(cutoff=CUT, occluder=True, light=None):
    """white ground at y = 0 (2 triangles), an alpha-tested occluder quad at y = 1 over x, z in [-1, 1]
    with uv (0, 0) at (-1, 1, 1), u along +x, v along -z"""
    b = scenes.MeshBuilder()
    g = scenes.preset_material("PLASTIC")
    g.albedo[:] = (0.7, 0.7, 0.7)
    g.specular_amount = 0.0
    g.roughness = 1.0
    b.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), b.material(g))
    uv = [[0, 0], [1, 0], [1, 1], [0, 1]]
    if occluder:
        o = scenes.preset_material("PLASTIC")
        o.albedo[:] = (0.5, 0.5, 0.5)
        o.specular_amount = 0.0
        o.albedo_tex = 0
        o.alpha_cutoff = cutoff
        o.material_mask |= 1                                          # two sided
        b.quad((-1, 1, 1), (2, 0, 0), (0, 0, -2), b.material(o))
        uv += [[0, 0], [1, 0], [1, 1], [0, 1]]
    lights = [light] if light is not None else []
    return b.arrays(lights, uvs=np.array(uv, np.float32), textures=[(_alpha_texture(), False)] if occluder else [])


def _alpha_expected(x, z):
    """the occluder is present at (x, z) where its texel block is opaque"""
    u, v = (x + 1) / 2, (1 - z) / 2
    col, row = int(u * 8) // 4, int(v * 8) // 4
    return (row, col) in ((0, 0), (1, 1))


def _down_rays(n=9):
    xs = (np.arange(n) + 0.5) / n * 2 - 1                             # avoids the block edges at 0
    rays = []
    for x in xs:
        for z in xs:
            if abs(x) < 0.15 or abs(z) < 0.15:
                continue
            rays.append([x, 3, z, 1e-4, 0, -1, 0, 1e30])
    return np.array(rays, np.float32)


@pytest.mark.parametrize("use_bvh", [False, True])
def test_alpha_test_cuts_closest_hits(oracle, use_bvh):
    scene = _alpha_scene()
    rays = _down_rays()
    t, tri, _ = oracle.trace_closest(scene, rays, use_bvh=use_bvh)
    for r, tt, k in zip(rays, t, tri):
        if _alpha_expected(r[0], r[2]):
            assert k >= 2 and abs(tt - 2) < 1e-5, r                   # occluder triangles follow the ground's
        else:
            assert k < 2 and abs(tt - 3) < 1e-5, r
    # without the cutoff the occluder is opaque everywhere
    t2, tri2, _ = oracle.trace_closest(_alpha_scene(cutoff=0.0), rays, use_bvh=use_bvh)
    assert (tri2 >= 2).all()
    # a cutoff above every alpha but 255 keeps only the opaque blocks; below 100/255 the 100 block joins
    _, tri3, _ = oracle.trace_closest(_alpha_scene(cutoff=0.3), rays, use_bvh=use_bvh)
    for r, k in zip(rays, tri3):
        u, v = (r[0] + 1) / 2, (1 - r[2]) / 2
        cut_away = v < 0.5 and u > 0.5                                # alpha 0 block only
        assert (k < 2) == cut_away, r


def test_alpha_tested_occluder_casts_a_patterned_shadow(oracle):
    sun = _sun(0.0, 2.0)
    cam = camera.Camera(eye=(0, 0.8, 3.0), target=(0, 0, 0), fovy=0.6)

    def film(scene):
        fr = camera.make_frame(cam, 64, 48, max_depth=1, firefly_clamp_scale=0.0, flags=_lib.YPT_FLAG_DISABLE_JITTER)
        return oracle.render(scene, fr, NEE, 0, 2, use_bvh=False)["film"][..., :3]
    cut = film(_alpha_scene(light=sun))
    lit = film(_alpha_scene(occluder=False, light=sun))
    dark = film(_alpha_scene(cutoff=0.0, light=sun))
    # every pixel sees the ground (the camera stays below the occluder): it is either lit or shadowed
    same_lit = np.all(cut == lit, axis=-1)
    same_dark = np.all(cut == dark, axis=-1)
    assert np.all(same_lit | same_dark)
    under = ~np.all(lit == dark, axis=-1)                             # pixels under the occluder
    frac = same_lit[under].mean()
    assert under.sum() > 100 and 0.25 < frac < 0.75, frac


@pytest.mark.gpu
def test_alpha_test_gpu_trace_matches_the_checker(ylib, oracle):
    scene = _alpha_scene()
    rays = _down_rays(33)
    it = ylib.Integrator(NEE, 0)
    it.set_scene(scene)
    t, tri, b = it.trace_closest(rays)
    it.close()
    rt, rtri, rb = oracle.trace_closest(scene, rays)
    assert np.array_equal(tri, rtri)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))


def _alpha_render_scene():
    b = scenes.MeshBuilder()
    g = scenes.preset_material("PLASTIC")
    g.albedo[:] = (0.8, 0.75, 0.7)
    b.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), b.material(g))
    rng = np.random.default_rng(5)
    leaf = rng.integers(0, 256, (16, 16, 4)).astype(np.uint8)
    o = scenes.preset_material("PLASTIC")
    o.albedo[:] = (0.3, 0.7, 0.2)
    o.albedo_tex = 0
    o.alpha_cutoff = 0.5
    o.material_mask |= 1
    b.quad((-1.5, 1.2, 1.5), (3, 0, 0), (0, 0, -3), b.material(o))
    b.quad((-1.5, 0, -1.2), (3, 0, 0), (0, 2.5, 0), b.material(o))  # a cut-out wall
    e = scenes.blank_material()
    e.emissive[:] = (10.0, 9.0, 8.0)
    e.albedo_tex = 0                                                 # an alpha-tested emitter
    e.alpha_cutoff = 0.5
    b.quad((-0.6, 3.0, -0.6), (0, 0, 1.2), (1.2, 0, 0), b.material(e))
    uv = np.tile(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), (4, 1))
    sun = _sun(0.3, 1.5)
    scene = b.arrays([sun], uvs=uv, textures=[(leaf, True)])
    cam = camera.Camera(eye=(0.5, 2.5, 5), target=(0, 0.8, 0), fovy=0.9)
    return scene, cam


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [BPT, NEE, BDPT])
def test_alpha_tested_render_matches_the_checker(ylib, oracle, integrator):
    scene, cam = _alpha_render_scene()
    frame = camera.make_frame(cam, 64, 48, max_depth=5, light_path_vertices=5)
    it = ylib.Integrator(integrator, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 3)
    g_last, g = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(scene, frame, integrator, 0, 3)
    assert np.array_equal(g_last[..., 3].view(np.uint32), ref["last"][..., 3].view(np.uint32))
    assert np.allclose(g[..., :3], ref["film"][..., :3], rtol=1e-4, atol=1e-6)


def test_alpha_tested_scene_renders_on_the_checker(oracle):
    scene, cam = _alpha_render_scene()
    fr = camera.make_frame(cam, 32, 24, max_depth=4, light_path_vertices=4)
    for integ in (NEE, BDPT):
        f = oracle.render(scene, fr, integ, 0, 2)["film"]
        assert np.isfinite(f).all() and f[..., 1].max() > 0
