"""The py_yapt scene surface (yapt_amd/py_yapt.py): transforms against closed forms, the world-space
commit of render objects (instancing by transform, non-uniform scale, mirroring, submeshes and their
materials), materials and presets, renderer variables with the reference's defaults and limits,
glTF meshes through Resources, a scene script run the way the app runs one, and on the GPU a
render through the surface against the checker with the FinalColor readback."""
import base64
import json
import math
import struct

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes
from yapt_amd.gltf import load as gltf_load
from yapt_amd import py_yapt as Y
from yapt_amd.py_yapt import Light, Material, vec3, vec4


def _quad_mesh(renderer, uvs=True, submeshes=None):
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float64)   # facing +y
    N = np.tile([0.0, 1.0, 0.0], (4, 1))
    T = np.tile([1.0, 0.0, 0.0, 1.0], (4, 1))
    I = np.array([0, 2, 1, 0, 3, 2])
    UV = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64) if uvs else None
    return renderer.createMesh("quad", P, I, normals=N, uvs=UV, tangents=T, submeshes=submeshes)


# ---------------------------------------------------------------- transforms
def test_axis_angle_orientation_rotates_the_axes():
    t = Y.Transform()
    t.setOrientation(Y.axis_angle_quat((1, 0, 0), 90))      # +y -> +z, -z -> +y
    assert np.allclose(t.up(), [0, 0, 1], atol=1e-6)
    assert np.allclose(t.forward(), [0, 1, 0], atol=1e-6)
    assert np.allclose(t.right(), [1, 0, 0], atol=1e-6)
    t.addOrientation(Y.axis_angle_quat((0, 1, 0), 90))      # applied in the local frame first
    # local +x rotated about y by 90 -> -z, then about x by 90 -> +y
    assert np.allclose(t.right(), [0, 1, 0], atol=1e-6)


def test_look_at_points_forward_at_the_target():
    t = Y.Transform()
    t.setScale(vec3(3.0))
    t.lookAt(vec3([1, 2, 3]), vec3([4, 6, 3]), vec3([0, 1, 0]))
    assert np.allclose(t.forward(), [0.6, 0.8, 0], atol=1e-6)
    assert abs(float(np.dot(t.up(), t.forward()))) < 1e-6 and t.up()[1] > 0
    assert np.allclose(t.getScale(), 1.0) and np.allclose(t.getTranslation(), [1, 2, 3])


def test_parent_chain_and_add_scale():
    parent, child = Y.Transform(), Y.Transform()
    parent.setTranslation(vec3([0, 10, 0]))
    parent.setScale(vec3(2.0))
    child.setParent(parent)
    child.setTranslation(vec3([1, 0, 0]))
    assert np.allclose(child.getMatrixWS()[:3, 3], [2, 10, 0])
    assert child.getParentId() == parent.getId()
    child.addScale(vec3(5.0))                                # assigns, as Transform::addScale does
    assert np.allclose(child.getScale(), 5.0)


# ---------------------------------------------------------------- the commit
def test_render_objects_are_instanced_by_transform():
    app = Y.Application(64, 48)
    r, s = app.get_renderer(), app.get_scene()
    mesh = _quad_mesh(r)
    red, blue = r.createMaterial("red"), r.createMaterial("blue")
    red.setAlbedo(vec3([1, 0, 0]))
    blue.setAlbedo(vec3([0, 0, 1]))
    a = s.create_render_object("a")
    a.setMesh(mesh)
    a.setMaterial(red, 0)
    a.getTransform().setTranslation(vec3([0, 5, 0]))
    b = s.create_render_object("b")
    b.setMesh(mesh)
    b.setMaterial(blue, 0)
    b.getTransform().setScale(vec3([2, 1, 4]))
    b.getTransform().setOrientation(Y.axis_angle_quat((0, 0, 1), 90))    # +y -> -x
    arr = s.arrays()
    assert arr.triangle_count == 4 and list(arr.inst_count) == [2, 2] and list(arr.inst_first) == [0, 2]
    assert np.allclose(arr.positions[:4], mesh.positions + [0, 5, 0])
    # b: scale then rotate 90 degrees about z: (x, y, z) -> (2x, y, 4z) -> (-y, 2x, 4z)
    Pb = np.stack([-mesh.positions[:, 1], 2 * mesh.positions[:, 0], 4 * mesh.positions[:, 2]], 1)
    assert np.allclose(arr.positions[4:], Pb, atol=1e-6)
    assert np.allclose(arr.normals[4:], [-1, 0, 0], atol=1e-6)
    assert np.allclose(arr.tangents[4:, :3], [0, 1, 0], atol=1e-6)
    assert [m.albedo[0] for m in arr.materials] == [1.0, 0.0]
    assert arr.uvs is not None and np.allclose(arr.uvs[:4], mesh.uvs)


def test_non_uniform_scale_uses_the_inverse_transpose():
    app = Y.Application(64, 48)
    r, s = app.get_renderer(), app.get_scene()
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    n = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    mesh = r.createMesh("tri", P, [0, 1, 2], normals=np.tile(n, (3, 1)))
    o = s.create_render_object("o")
    o.setMesh(mesh)
    o.getTransform().setScale(vec3([4, 1, 1]))
    arr = s.arrays()
    want = np.array([0.25, 1.0, 0.0]) / math.hypot(0.25, 1.0)          # n / scale, normalized
    assert np.allclose(arr.normals, want, atol=1e-6)


def test_mirroring_reverses_winding_and_handedness():
    app = Y.Application(64, 48)
    r, s = app.get_renderer(), app.get_scene()
    mesh = _quad_mesh(r)
    o = s.create_render_object("o")
    o.setMesh(mesh)
    o.getTransform().setScale(vec3([-1, 1, 1]))
    arr = s.arrays()
    P, I = arr.positions.astype(np.float64), arr.indices.reshape(-1, 3)
    face = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    assert (face[:, 1] > 0).all()                        # still facing +y, as the normals say
    assert np.allclose(arr.normals, [0, 1, 0]) and np.allclose(arr.tangents[:, 3], -1.0)
    assert np.allclose(arr.tangents[:, :3], [-1, 0, 0])


def test_submeshes_become_instances_with_their_materials():
    app = Y.Application(64, 48)
    r, s = app.get_renderer(), app.get_scene()
    mesh = _quad_mesh(r, submeshes=[(0, 3, 0), (3, 3, 0)])
    m0 = r.createMaterial("m0")
    o = s.create_render_object("o")
    o.setMesh(mesh)
    o.setMaterial(m0, 0)                                  # submesh 1 keeps the default material
    arr = s.arrays()
    assert list(arr.inst_count) == [1, 1] and len(arr.materials) == 2
    assert o.getMaterials() == [m0, None]
    tex = r.createTexture("t", np.full((2, 2, 4), 255, np.uint8), srgb=True)
    m0.setAlbedoTexture(tex, Y.vec2(2.0))
    m1 = r.createMaterial("m1")
    m1.setORMTexture(tex)
    o.setMaterials([m0, m1])
    arr = s.arrays()
    assert len(arr.texel_arrays) == 1                     # one texture, bound twice
    assert arr.materials[0].albedo_tex == 0 and list(arr.materials[0].albedo_tex_scale) == [2.0, 2.0]
    assert arr.materials[1].orm_tex == 0 and arr.materials[1].albedo_tex == -1


def test_lights_follow_their_transforms():
    app = Y.Application(64, 48)
    r, s = app.get_renderer(), app.get_scene()
    o = s.create_render_object("floor")
    o.setMesh(_quad_mesh(r))
    spot = s.create_light("spot", Light.LightType.SPOT)
    spot.getTransform().lookAt(vec3([0, 4, 0]), vec3([0, 0, 0]), vec3([0, 0, 1]))
    spot.setColor(vec3([1, 0.5, 0.25]))
    spot.setIntensity(8.0)
    spot.setSpotConeAngles(0.2, 3.0)                      # outer clamps to pi / 2
    assert abs(spot.getSpotOuterAngle() - math.pi / 2) < 1e-7
    p = s.arrays().lights[0]
    assert p.type == _lib.YPT_LIGHT_SPOT
    assert np.allclose(list(p.position), [0, 4, 0]) and np.allclose(list(p.direction), [0, -1, 0], atol=1e-6)
    assert np.allclose(list(p.intensity), [8, 4, 2])


# ---------------------------------------------------------------- materials and variables
def test_presets_and_setters_match_the_reference_parameters():
    r = Y.Renderer()
    m = r.createMaterial("m")
    m.setAlbedo(vec3([0.1, 0.2, 0.3]))
    m.setFromMaterialPreset(Material.MaterialPreset.METAL_COPPER)   # replaces every parameter
    want = scenes.preset_material("METAL_COPPER")
    assert bytes(m.record(lambda t: 0)) == bytes(want)
    m.setTwoSided(True)
    m.setEnableDispersion(True)
    m.setAlphaBlend(True)
    assert m.record(lambda t: 0).material_mask == 7
    m.setTwoSided(False)
    assert m.getTwoSided() is False and m.getAlphaBlend() is True
    m.setCauchysCoefficients(Y.vec2([1.6, 0.01]))
    assert np.allclose(m.getCauchysCoefficients(), [1.6, 0.01])


def test_renderer_variables_have_the_reference_defaults_and_limits():
    r = Y.Renderer()
    assert r.getRendererVariable("RenderPipeline").selected() == "BPT_WITH_NEE"
    assert r.getRendererVariable("Generic.RenderResolution").getAsIntArray() == [1920, 1080]
    v = r.getRendererVariable("Firefly.FireflyClampRelaxPower")
    v.setFromFloatArray([9.0])
    assert v.getAsFloatArray() == [4.0]                   # clamped to the limit
    with pytest.raises(ValueError):
        r.getRendererVariable("Denoise.Mode").setSelectedOption(9)
    assert r.getRendererVariable("no such variable") is None
    assert "Tonemap.ManualExposure" in r.getAllRendererVariableNames()


def test_camera_frame_matches_a_look_at_camera():
    app = Y.Application(96, 64)
    r, s = app.get_renderer(), app.get_scene()
    o = s.create_render_object("floor")
    o.setMesh(_quad_mesh(r))
    cam = s.get_main_camera()
    cam.getTransform().lookAt(vec3([0, 3, 6]), vec3([0, 0, 0]), vec3([0, 1, 0]))
    cam.setFOVY(0.8)
    f = r.frame_desc()
    g = camera.make_frame(camera.Camera(eye=(0, 3, 6), target=(0, 0, 0), fovy=0.8), 96, 64)
    assert np.allclose(list(f.view_to_world), list(g.view_to_world), atol=1e-6)
    assert np.allclose(list(f.uv_to_view), list(g.uv_to_view), atol=1e-6)
    assert (f.width, f.height, f.max_depth) == (96, 64, 32)   # maxRayDepth of RaytraceCommonResources.cpp:80


# ---------------------------------------------------------------- resources and scripts
def _glb(tmp_path):
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    N = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    I = np.array([0, 2, 1, 0, 3, 2], np.uint16)
    blob = P.tobytes() + N.tobytes() + I.tobytes() + b"\0\0"
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0, "translation": [0, 9, 0]}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2},
                                      {"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2}]}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 48}, {"buffer": 0, "byteOffset": 48, "byteLength": 48},
                           {"buffer": 0, "byteOffset": 96, "byteLength": 12}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                         {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
                         {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR"}],
           "buffers": [{"byteLength": len(blob)}]}
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
    p = tmp_path / "plane.glb"
    p.write_bytes(b"glTF" + struct.pack("<II", 2, 12 + len(body)) + body)
    return str(p)


def test_resources_load_object_space_meshes(tmp_path):
    app = Y.Application(32, 32)
    meshes = app.get_resources().load_meshes_from_path(_glb(tmp_path))
    mesh = meshes[0][0][0]                                 # the indexing scene scripts use
    assert mesh.getSubmeshCount() == 2 and meshes[0][1] == {0: (0, 0), 1: (0, 1)}
    assert np.allclose(mesh.positions[:, 1], 0.0)          # the node's translation is not applied


SCRIPT = '''
from yapt_amd.py_yapt import Material, Light, vec3, vec4, axis_angle_quat
renderer = yapt_instance.get_renderer()
scene = yapt_instance.get_scene()
plane = yapt_instance.get_resources().load_meshes_from_path(PLANE)[0][0][0]
mat = renderer.createMaterial("floor_mat")
mat.setFromMaterialPreset(Material.MaterialPreset.PLASTIC)
mat.setRoughness(1.0)
mat.setSpecularAmount(0.0)
for k, x in enumerate((-3.0, 3.0)):
    obj = scene.create_render_object("floor_%d" % k)
    obj.setMesh(plane)
    obj.setMaterial(mat, 0)
    obj.setMaterial(mat, 1)
    obj.getTransform().setTranslation(vec3([x, 0, 0]))
    obj.getTransform().setScale(vec3(2.5))
sun = scene.create_light("sun", Light.LightType.DIRECTIONAL)
sun.getTransform().lookAt(vec3([-6, 12, 8]), vec3([0, 0, 0]), vec3([0, 1, 0]))
cam = scene.get_main_camera().getTransform()
cam.setTranslation(vec3([0, 8, 12]))
cam.setOrientation(axis_angle_quat((1, 0, 0), -30))
renderer.getRendererVariable("World.EnvIntensityScale").setFromFloatArray([0.0])
renderer.getRendererVariable("Tonemap.UseAutoExposure").setSelectedOption(0)
'''


def test_a_scene_script_runs_against_the_surface(tmp_path, oracle):
    script = tmp_path / "scene.py"
    script.write_text(SCRIPT)
    app = Y.Application(48, 32)
    app.run_script(str(script), {"PLANE": repr(_glb(tmp_path))})
    r, s = app.get_renderer(), app.get_scene()
    arr = s.arrays()
    assert arr.triangle_count == 8 and len(arr.lights) == 1
    assert np.allclose(arr.positions[:, 0].min(), -5.5) and np.allclose(arr.positions[:, 0].max(), 5.5)
    film = oracle.render(arr, r.frame_desc(), _lib.YPT_INTEGRATOR_NEE, 0, 2)["film"]
    assert np.isfinite(film).all() and film[..., 1].max() > 0


@pytest.mark.gpu
def test_rendering_through_the_surface_matches_the_checker(tmp_path, oracle):
    script = tmp_path / "scene.py"
    script.write_text(SCRIPT)
    app = Y.Application(96, 64)
    app.run_script(str(script), {"PLANE": repr(_glb(tmp_path))})
    r = app.get_renderer()
    for pipeline in (0, 1, 2):
        r.getRendererVariable("RenderPipeline").setSelectedOption(pipeline)
        r.resetRenderOutput()
        assert r.render(4) == 4
        g = r.film()
        ref = oracle.render(r._arrays, r.frame_desc(), r.pipeline(), 0, 4)["film"]
        assert np.allclose(g[..., :3], ref[..., :3], rtol=1e-4, atol=1e-6), pipeline
        rb = r.readback(Y.ReadbackTarget.FinalColor)
        assert rb.getState() == Y.ReadbackState.Ready and rb.getData().shape == (64, 96, 4)
        assert rb.getData()[..., :3].max() > 0
    app.shutdown()


def test_scene_loader_builds_the_scene_gltf_load_bakes(tmp_path):
    """the node tree, materials, textures and lights that SceneLoader creates commit to the arrays that
    gltf.load bakes directly (matrix nodes included)"""
    import test_gltf as G
    path = str(G._write_gltf(tmp_path))
    want, _ = gltf_load(path)
    app = Y.Application(32, 32)
    loader = Y.SceneLoader(app.get_resources(), app.get_scene(), app.get_renderer())
    roots = loader.load_scene_gltf(path)
    assert len(roots) == 2 and loader.get_material("m0") is not None
    got = app.get_scene().arrays()
    assert got.triangle_count == want.triangle_count
    assert np.allclose(np.sort(got.positions, axis=0), np.sort(want.positions, axis=0), atol=1e-5)
    assert np.allclose(np.sort(got.normals, axis=0), np.sort(want.normals, axis=0), atol=1e-5)
    gl = sorted(got.lights, key=lambda l: l.type)
    wl = sorted(want.lights, key=lambda l: l.type)
    for a, b in zip(gl, wl):
        assert a.type == b.type
        assert np.allclose(list(a.position), list(b.position), atol=1e-5)
        assert np.allclose(list(a.direction), list(b.direction), atol=1e-5)
        assert np.allclose(list(a.intensity), list(b.intensity))
    assert bytes(loader.get_material("m0").record(lambda t: -1)) == bytes(want.materials[0])
    doc, blob, _ = G._textured_doc("MASK", 0, in_buffer=True)
    tpath = G._save(tmp_path, doc, blob, True)
    want, _ = gltf_load(tpath)
    app = Y.Application(32, 32)
    Y.SceneLoader(app.get_resources(), app.get_scene(), app.get_renderer()).load_scene_gltf(tpath)
    got = app.get_scene().arrays()
    assert np.allclose(got.uvs, want.uvs) and len(got.texel_arrays) == len(want.texel_arrays) == 3
    m = got.materials[0]
    assert abs(m.alpha_cutoff - 0.3) < 1e-7
    for slot in ("albedo", "orm", "normal"):
        k = getattr(m, slot + "_tex")
        assert np.array_equal(got.texel_arrays[k], want.texel_arrays[getattr(want.materials[0], slot + "_tex")])


def test_matrix_nodes_decompose_into_trs():
    q = Y.axis_angle_quat((1, 2, 3), 40)
    t = Y.Transform()
    t.setOrientation(q)
    t.setScale(vec3([2, 3, -4]))
    t.setTranslation(vec3([1, -2, 5]))
    m = t.getMatrixWS()
    tt, qq, ss = Y._decompose(m)
    u = Y.Transform()
    u.setTranslation(tt)
    u.setOrientation(qq)
    u.setScale(ss)
    assert np.allclose(u.getMatrixWS(), m, atol=1e-9)
