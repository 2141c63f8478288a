"""Environment textures (the World.Skycube renderer variable): the inverse trigonometric functions of the
shared core, the long-lat and cube-map lookups of getSkyBoxColor (materialsCommonResources.hlsl:181-223)
against texel centres, an independent NumPy lookup and the face table of D3D / Vulkan, validation, and
renders under an environment: a white sphere in a constant sky, and bidirectional light subpaths from a
textured sky against next-event estimation."""
import math
import struct

import numpy as np
import pytest

from yapt_amd import _lib, scenes

LONGLAT, CUBE = _lib.YPT_ENVIRONMENT_LONGLAT, _lib.YPT_ENVIRONMENT_CUBE


def test_inverse_trigonometric_functions(oracle):
    x = np.linspace(-1.0, 1.0, 40001).astype(np.float32)
    for name, ref in (("acos", np.arccos), ("asin", np.arcsin)):
        got = oracle.evaluate(name, x[:, None])[:, 0]
        assert np.abs(got - ref(x.astype(np.float64))).max() < 4e-7, name
    rng = np.random.default_rng(1)
    yx = (rng.normal(size=(40000, 2)) * np.exp(rng.uniform(-12, 12, size=(40000, 1)))).astype(np.float32)
    got = oracle.evaluate("atan2", yx)[:, 0]
    ref = np.arctan2(yx[:, 0].astype(np.float64), yx[:, 1].astype(np.float64))
    assert np.abs(got - ref).max() < 4e-7
    special = np.array([[0, 0], [0, 1], [0, -1], [1, 0], [-1, 0], [1, 1], [-1, -1], [1, -1]], np.float32)
    got = oracle.evaluate("atan2", special)[:, 0]
    want = [0, 0, math.pi, math.pi / 2, -math.pi / 2, math.pi / 4, -3 * math.pi / 4, 3 * math.pi / 4]
    assert np.allclose(got, want, atol=3e-7)
    ends = oracle.evaluate("acos", np.array([[1.0], [-1.0], [0.0], [0.5]], np.float32))[:, 0]
    assert np.allclose(ends, [0.0, math.pi, math.pi / 2, math.pi / 3], atol=3e-7) and ends[0] == 0.0


def test_texture_less_sky_is_the_absolute_direction(oracle):
    d = np.random.default_rng(2).normal(size=(100, 3))
    assert np.array_equal(oracle.env_color(None, d), np.abs(d.astype(np.float32)))


@pytest.mark.parametrize("w,h", [(16, 8), (7, 5)])
def test_longlat_texel_centres_read_their_texel(oracle, w, h):
    img = np.random.default_rng(3).uniform(0, 4, size=(h, w, 4)).astype(np.float32)
    d = scenes.longlat_directions(w, h).reshape(-1, 3)
    got = oracle.env_color((LONGLAT, img), d)
    assert np.abs(got - img[..., :3].reshape(-1, 3)).max() < 1e-5
    # the axes: row 0 looks up, +x reads the middle column boundary, +z three quarters across
    img = np.zeros((4, 8, 4), np.float32)
    img[0, :, 0] = 1.0                      # top row red
    img[:, 5, 1] = 1.0                      # column 5 (u in [0.625, 0.75]) green
    up, z = oracle.env_color((LONGLAT, img), [[0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert up[0] == 0.5                     # the repeat sampler averages the top and bottom rows at the pole
    assert z[1] == 0.5                      # u = 0.75 sits between columns 5 and 6


def _bilinear_numpy(img, u, v, wrap):
    """an independent restatement of a bilinear lookup between texel centres (float64)"""
    h, w = img.shape[:2]
    if wrap:
        u, v = u - np.floor(u), v - np.floor(v)
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x).astype(int), np.floor(y).astype(int)
    ax, ay = (x - x0)[:, None], (y - y0)[:, None]
    if wrap:
        xa, xb, ya, yb = x0 % w, (x0 + 1) % w, y0 % h, (y0 + 1) % h
    else:
        xa, xb, ya, yb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1), np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    im = img[..., :3].astype(np.float64)
    top = im[ya, xa] * (1 - ax) + im[ya, xb] * ax
    bot = im[yb, xa] * (1 - ax) + im[yb, xb] * ax
    return top * (1 - ay) + bot * ay


def test_longlat_lookup_matches_an_independent_bilinear(oracle):
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 2, size=(24, 48, 4)).astype(np.float32)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u = np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi) + 0.5
    v = np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi
    want = _bilinear_numpy(img, u, v, wrap=True)
    got = oracle.env_color((LONGLAT, img), d)
    assert np.abs(got - want).max() < 2e-4


def test_cube_texel_centres_and_face_axes(oracle):
    cube = np.random.default_rng(5).uniform(0, 4, size=(6, 8, 8, 4)).astype(np.float32)
    d = scenes.cube_face_directions(8).reshape(-1, 3)
    assert np.abs(oracle.env_color((CUBE, cube), d) - cube[..., :3].reshape(-1, 3)).max() < 1e-5
    # one colour per face: each axis reads its own face, in the order +X, -X, +Y, -Y, +Z, -Z
    faces = np.zeros((6, 4, 4, 4), np.float32)
    for f in range(6):
        faces[f, ..., 0] = f + 1
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32) * 3.0
    assert np.array_equal(oracle.env_color((CUBE, faces), axes)[:, 0], np.arange(1, 7, dtype=np.float32))
    # the corner texels of +X: looking along +x a little up (+y) and towards -z reads row 0 (t = 0 is up)
    # and the right-hand end of the row (s = 1 is -z)
    px = np.zeros((6, 2, 2, 4), np.float32)
    px[0, 0, 1, 0] = 1.0
    assert oracle.env_color((CUBE, px), [[1.0, 0.99, -0.99]])[0, 0] == 1.0


@pytest.mark.parametrize("kind", [LONGLAT, CUBE])
def test_smooth_sky_through_either_layout(oracle, kind):
    # a smooth sky sampled into either layout reads back as the sky between the texels, across face edges and
    # the long-lat seam; a flipped or rotated face would differ by the sky's own contrast (0.4 and more)
    fn = lambda d: scenes.sky_radiance(d, sun=0.0)
    env = scenes.longlat_environment(256, 128, fn) if kind == LONGLAT else scenes.cube_environment(96, fn)
    rng = np.random.default_rng(6)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d[np.abs(d[:, 1]) > 0.03]          # away from the horizon, where the sky has a step
    got = oracle.env_color(env, d)
    assert np.abs(got - fn(d)).max() < 2e-2


def test_environment_validation(oracle):
    with pytest.raises(ValueError):
        _lib.environment_texels(CUBE, np.zeros((6, 4, 5, 4)))
    with pytest.raises(ValueError):
        _lib.environment_texels(LONGLAT, np.zeros((6, 4, 4, 4)))
    for bad in (-1.0, np.nan, np.inf):
        img = np.ones((2, 4, 4), np.float32)
        img[1, 2, 0] = bad
        with pytest.raises(RuntimeError):
            oracle.env_color((LONGLAT, img), [[0, 1, 0]])


def _sphere_in_sky(albedo, environment, centre=(0.0, 0.0, 0.0)):
    from yapt_amd import camera
    b = scenes.MeshBuilder()
    m = scenes.preset_material("PLASTIC")
    m.albedo[:] = (albedo, albedo, albedo)
    m.roughness = 1.0
    m.specular_amount = 0.0
    P, N, I = scenes.sphere_template(24, 24)
    b.add(P + np.asarray(centre, np.float32), N, I, b.material(m))
    cam = camera.Camera(eye=(0.0, 0.0, 4.0), target=(0.0, 0.0, 0.0), fovy=math.pi * 0.3)
    return b.arrays(environment=environment), cam


@pytest.mark.parametrize("integrator", [0, 1])
def test_white_sphere_in_a_constant_sky_disappears(oracle, integrator):
    # a convex Lambertian sphere of albedo 1 under a constant sky reflects exactly the sky's radiance: each
    # sample of each pixel carries what the same sample (same wavelengths) sees with the sphere moved out
    # of view
    from yapt_amd import camera
    const = (LONGLAT, np.full((4, 8, 4), 0.75, np.float32))
    scene, cam = _sphere_in_sky(1.0, const)
    empty, _ = _sphere_in_sky(1.0, const, centre=(0.0, 0.0, 40.0))      # behind the camera
    frame = camera.make_frame(cam, 24, 24, max_depth=6, env_intensity_scale=2.0)
    a = oracle.render(scene, frame, integrator, 0, 2, use_bvh=True)["last"]
    b = oracle.render(empty, frame, integrator, 0, 2, use_bvh=True)["last"]
    assert np.array_equal(a[..., 3], b[..., 3])
    assert np.allclose(a[..., :3], b[..., :3], rtol=1e-4, atol=0)
    # grey: the sphere's pixels average albedo x sky
    scene, cam = _sphere_in_sky(0.5, const)
    film = oracle.render(scene, frame, integrator, 0, 64, use_bvh=True)["film"][..., 1].astype(np.float64)
    sky = oracle.render(empty, frame, integrator, 0, 64, use_bvh=True)["film"][..., 1].astype(np.float64)
    assert abs(film[8:16, 8:16].mean() / sky[8:16, 8:16].mean() - 0.5) < 0.03


@pytest.mark.parametrize("kind", [LONGLAT, CUBE])
def test_bdpt_light_subpaths_from_a_textured_sky_agree_with_nee(oracle, kind):
    # with no lights, BDPT starts light subpaths at the sky (lightSampling.hlsl:78-120) and weights them
    # against camera paths that escape; both estimate what NEE's escaping paths see
    from yapt_amd import camera
    env = scenes.longlat_environment(64, 32) if kind == LONGLAT else scenes.cube_environment(24)
    scene, cam, scale = scenes.punctual_lights("env", sphere_res=(12, 12), environment=env)
    frame = camera.make_frame(cam, 48, 32, max_depth=6, env_intensity_scale=scale, light_path_vertices=6)
    nee = oracle.render(scene, frame, _lib.YPT_INTEGRATOR_NEE, 0, 256, use_bvh=True)["film"][..., 1].mean()
    bdpt = oracle.render(scene, frame, _lib.YPT_INTEGRATOR_BDPT, 0, 128, use_bvh=True)["film"][..., 1].mean()
    assert nee > 0.02
    assert abs(bdpt / nee - 1.0) < 0.02, (nee, bdpt)


# ---------------------------------------------------------------- image files and the py_yapt surface
def test_radiance_hdr_texels_and_runs():
    from yapt_amd import hdr_images
    # flat texels, decoded by the format's definition: value = mantissa * 2^(E - 136)
    flat = bytes([128, 64, 0, 129, 255, 255, 255, 0, 1, 2, 3, 140])
    data = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 3\n" + flat
    got = hdr_images.decode_hdr(data)
    assert got.shape == (1, 3, 4)
    assert np.array_equal(got[0, 0, :3], [1.0, 0.5, 0.0]) and np.array_equal(got[0, 1, :3], [0, 0, 0])
    assert np.array_equal(got[0, 2, :3], [16.0, 32.0, 48.0]) and np.all(got[..., 3] == 1)
    # an adaptive run-length scanline of 8 texels: a run of 8 per channel, and literals
    rle = bytes([2, 2, 0, 8]) + bytes([128 + 8, 100]) + bytes([8]) + bytes(range(8)) + bytes([128 + 8, 7]) + bytes([128 + 8, 130])
    got = hdr_images.decode_hdr(b"#?RADIANCE\n\n-Y 1 +X 8\n" + rle)
    assert np.allclose(got[0, :, 0], 100 * 2.0 ** -6) and np.allclose(got[0, :, 1], np.arange(8) * 2.0 ** -6)
    # a written image reads back within RGBE's precision (8 mantissa bits of the largest channel)
    img = np.exp(np.random.default_rng(7).uniform(-6, 6, size=(9, 40, 3)))
    for rle in (True, False):
        back = hdr_images.decode_hdr(hdr_images.encode_hdr(img, rle=rle))[..., :3]
        assert np.all(np.abs(back - img) <= img.max(axis=2, keepdims=True) / 128.0)
    with pytest.raises(ValueError):
        hdr_images.decode_hdr(b"#?RADIANCE\n\n-Y 1 +X 8\n" + bytes([2, 2, 0, 8, 128 + 9, 1]))


def _exr_bytes(planes, compression, half=True):
    """an OpenEXR scanline file written from the file layout document: header attributes, the chunk
    offset table, chunks of (y, size, data) with channels in alphabetical order"""
    import zlib
    h, w = next(iter(planes.values())).shape
    names = sorted(planes)

    def attr(name, kind, value):
        return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value
    ptype = 1 if half else 2
    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for n in names) + b"\0"
    header = struct.pack("<iI", 20000630, 2)
    header += attr("channels", "chlist", chl) + attr("compression", "compression", bytes([compression]))
    header += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    header += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    header += attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0)) + attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    lines = 16 if compression == 3 else 1
    dt = np.float16 if half else np.float32
    chunks = []
    for y in range(0, h, lines):
        raw = b"".join(planes[n][yy].astype(dt).tobytes() for yy in range(y, min(h, y + lines)) for n in names)
        if compression in (2, 3):
            b = np.frombuffer(raw, np.uint8)
            inter = np.concatenate([b[0::2], b[1::2]]).astype(np.int64)
            pred = np.concatenate([inter[:1], (inter[1:] - inter[:-1] + 128) & 0xFF]).astype(np.uint8)
            packed = zlib.compress(pred.tobytes())
            raw = packed if len(packed) < len(raw) else raw
        chunks.append(struct.pack("<ii", y, len(raw)) + raw)
    table_start = len(header)
    offs, cur = [], table_start + 8 * len(chunks)
    for c in chunks:
        offs.append(cur)
        cur += len(c)
    return header + struct.pack("<%dQ" % len(offs), *offs) + b"".join(chunks)


@pytest.mark.parametrize("compression,half", [(0, True), (0, False), (2, True), (3, True), (3, False)])
def test_openexr_scanline_files(compression, half):
    from yapt_amd import hdr_images
    rng = np.random.default_rng(8)
    planes = {c: rng.uniform(0, 50, size=(21, 13)).astype(np.float32) for c in "RGB"}
    planes["R"][3, 4] = 0.0
    got = hdr_images.decode_exr(_exr_bytes(planes, compression, half))
    for k, c in enumerate("RGB"):
        want = planes[c].astype(np.float16).astype(np.float32) if half else planes[c]
        assert np.array_equal(got[..., k], want), c
    assert np.all(got[..., 3] == 1)
    with pytest.raises(ValueError, match="PIZ"):
        hdr_images.decode_exr(_exr_bytes(planes, 0, half)[:0] + _exr_bytes(planes, 0, half).replace(
            b"compression\0compression\0\x01\0\0\0\0", b"compression\0compression\0\x01\0\0\0\x04"))


def test_ktx2_float_cube_map():
    from yapt_amd import ktx
    faces = np.random.default_rng(9).uniform(0, 9, size=(6, 4, 4, 4)).astype(np.float16)
    header = ktx.KTX2_IDENTIFIER + struct.pack("<9I", 97, 2, 4, 4, 0, 0, 6, 1, 0)   # R16G16B16A16_SFLOAT
    index = struct.pack("<IIII QQ", 0, 0, 0, 0, 0, 0)
    level = struct.pack("<3Q", 80 + 24, faces.nbytes, faces.nbytes)
    texels, cube = ktx.decode_float(header + index + level + faces.tobytes())
    assert cube and np.array_equal(texels, faces.astype(np.float32))


def test_skycube_variable_commits_the_sky_with_the_scene(oracle, tmp_path):
    from yapt_amd import hdr_images, py_yapt as Y
    app = Y.Application(24, 16)
    r, s = app.get_renderer(), app.get_scene()
    P, N, I = scenes.sphere_template(12, 12)
    mesh = r.createMesh("sphere", P, I, normals=N)
    obj = s.createRenderObject("ball")
    obj.setMesh(mesh)
    mat = r.createMaterial("m")
    obj.setMaterial(mat)
    s.getMainCamera().getTransform().setTranslation(Y.vec3([0, 0, 5]))
    assert r.getRendererVariable("World.Skycube").getTexture() is None
    plain = s.arrays()
    assert plain.env is None
    path = tmp_path / "sky.hdr"
    sky = scenes.longlat_environment(32, 16)[1]
    path.write_bytes(hdr_images.encode_hdr(sky))
    tex = app.get_resources().load_texture_2D(str(path), "R16G16B16A16_SFLOAT")
    assert tex.is_float and tex.getDimension() == Y.ResourceDimension.TEXTURE_2D
    r.getRendererVariable("World.Skycube").setTexture(tex)
    arr = s.arrays()
    assert arr.env.type == LONGLAT and arr.env.width == 32 and arr.env.height == 16
    assert np.all(np.abs(arr.env_texels[..., :3] - sky[..., :3]) <= sky[..., :3].max(axis=2, keepdims=True) / 128)
    a = oracle.render(arr, r.frame_desc(), _lib.YPT_INTEGRATOR_NEE, 0, 2)["film"]
    b = oracle.render(plain, r.frame_desc(), _lib.YPT_INTEGRATOR_NEE, 0, 2)["film"]
    assert not np.array_equal(a, b)
    cube = Y.Texture("cube", scenes.cube_environment(8)[1], cube=True)
    r.getRendererVariable("World.Skycube").setTexture(cube)
    assert s.arrays().env.type == CUBE
    mat.setAlbedoTexture(tex)                # float textures are for the sky only
    with pytest.raises(ValueError):
        s.arrays()
