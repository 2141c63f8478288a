"""glTF 2.0 loading: container formats (.gltf with a data-URI buffer, .glb), node transforms baked to
world space, the reference scene loader's material and KHR_lights_punctual mapping
(pythonScripts/yapt/scene_loader.py), and a render of a loaded scene on the GPU against the checker."""
import base64
import json
import math
import struct

import numpy as np
import pytest

from yapt_amd import _lib, gltf


def _doc():
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)   # quad in xz, facing +y
    N = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    I = np.array([0, 2, 1, 0, 3, 2], np.uint16)
    blob = P.tobytes() + N.tobytes() + I.tobytes() + b"\0\0"
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 48}, {"buffer": 0, "byteOffset": 48, "byteLength": 48},
             {"buffer": 0, "byteOffset": 96, "byteLength": 12}]
    acc = [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
           {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
           {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR"}]
    s = math.sqrt(0.5)
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 3]}],
        "nodes": [
            {"mesh": 0, "translation": [1, 2, 3], "rotation": [0, s, 0, s], "scale": [2, 2, 2], "children": [1, 2]},
            {"mesh": 1, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 5, 0, 1]},
            {"rotation": [-s, 0, 0, s], "translation": [0, 4, 0], "extensions": {"KHR_lights_punctual": {"light": 0}}},
            {"translation": [0, 10, 0], "extensions": {"KHR_lights_punctual": {"light": 1}}},
        ],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2, "material": 0}]},
                   {"primitives": [{"attributes": {"POSITION": 0}, "indices": 2}]}],
        "materials": [{"name": "m0", "doubleSided": True, "emissiveFactor": [0.5, 0.25, 0.0],
                       "pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.6, 0.4, 0.5], "metallicFactor": 0.0,
                                                "roughnessFactor": 0.3},
                       "alphaMode": "BLEND",
                       "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                                      "KHR_materials_ior": {"ior": 1.33},
                                      "KHR_materials_sheen": {"sheenColorFactor": [0.5, 0.25, 0.25], "sheenRoughnessFactor": 0.6},
                                      "KHR_materials_volume": {"thicknessFactor": 1.0, "attenuationDistance": 2.0,
                                                               "attenuationColor": [0.5, 1.0, 0.25]}}}],
        "extensionsUsed": ["KHR_lights_punctual"],
        "extensions": {"KHR_lights_punctual": {"lights": [
            {"type": "spot", "color": [1, 0.5, 0.25], "intensity": 683.0, "spot": {"innerConeAngle": 0.2, "outerConeAngle": 0.4}},
            {"type": "point", "intensity": 1366.0}]}},
        "bufferViews": views, "accessors": acc,
    }
    return doc, blob


def _write_gltf(tmp_path):
    doc, blob = _doc()
    doc["buffers"] = [{"byteLength": len(blob), "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}]
    p = tmp_path / "scene.gltf"
    p.write_text(json.dumps(doc))
    return p


def _write_glb(tmp_path):
    doc, blob = _doc()
    doc["buffers"] = [{"byteLength": len(blob)}]
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
    p = tmp_path / "scene.glb"
    p.write_bytes(b"glTF" + struct.pack("<II", 2, 12 + len(body)) + body)
    return p


def test_transforms_are_baked_to_world_space(tmp_path):
    scene, info = gltf.load(str(_write_gltf(tmp_path)))
    assert scene.triangle_count == 4
    P = scene.positions
    # node 0: scale 2, rotate 90 degrees about +y (x -> -z), translate (1, 2, 3): (-1, 0, -1) -> (-1, 2, 5)
    assert np.allclose(P[0], [1 - 2, 2, 3 + 2], atol=1e-5)
    assert np.allclose(P[2], [1 + 2, 2, 3 - 2], atol=1e-5)
    # node 1 (child, matrix = translate (0, 5, 0) in the parent's frame): scaled by the parent, 10 units up
    assert np.allclose(P[4], P[0] + [0, 10, 0], atol=1e-5)
    assert np.allclose(scene.normals[:8], [0, 1, 0], atol=1e-6)   # given normals and area normals agree here


def test_glb_and_gltf_load_the_same_scene(tmp_path):
    a, _ = gltf.load(str(_write_gltf(tmp_path)))
    b, _ = gltf.load(str(_write_glb(tmp_path)))
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.indices, b.indices)


def test_material_mapping_follows_the_reference_loader(tmp_path):
    scene, _ = gltf.load(str(_write_gltf(tmp_path)))
    m, d = scene.materials[0], scene.materials[1]
    assert np.allclose(list(m.albedo), [0.8, 0.6, 0.4]) and abs(m.metalness) < 1e-7 and abs(m.roughness - 0.3) < 1e-6
    assert np.allclose(list(m.emissive), [2.0, 1.0, 0.0])                    # x emissiveStrength
    assert abs(m.dielectric_ior - 1.33) < 1e-6
    assert abs(m.sheen_amount - 0.5) < 1e-7 and np.allclose(list(m.sheen_tint), [1.0, 0.5, 0.5])
    assert abs(m.sheen_roughness - 0.36) < 1e-6
    assert np.allclose(list(m.absorption), [math.log(2) / 2, 0.0, math.log(4) / 2], atol=1e-6)
    assert abs(m.transparency - 0.5) < 1e-7                                 # BLEND coverage 0.5, no transmission
    assert m.material_mask & gltf.MASK_TWO_SIDED
    assert abs(d.specular_amount) < 1e-7 and abs(d.metalness) < 1e-7      # default: PLASTIC, specular amount 0


def test_lights_follow_khr_lights_punctual(tmp_path):
    _, info = gltf.load(str(_write_gltf(tmp_path)))
    spot, point = [l.params() for l in info["lights"]]
    assert spot.type == _lib.YPT_LIGHT_SPOT and point.type == _lib.YPT_LIGHT_POINT
    # child of node 0: position = node0 * (0, 4, 0) = (1, 2 + 8, 3); rotated -90 degrees about x, then the
    # parent's 90 degrees about y: -z -> -y -> -y
    assert np.allclose(list(spot.position), [1, 10, 3], atol=1e-5)
    assert np.allclose(list(spot.direction), [0, -1, 0], atol=1e-5)
    assert np.allclose(list(spot.intensity), [1.0, 0.5, 0.25])             # 683 cd / 683 lm/W
    assert abs(spot.cos_inner - math.cos(0.2)) < 1e-6 and abs(spot.cos_outer - math.cos(0.4)) < 1e-6
    assert np.allclose(list(point.position), [0, 10, 0]) and np.allclose(list(point.intensity), [2, 2, 2])


@pytest.mark.gpu
def test_loaded_scene_renders_like_the_checker(tmp_path, oracle):
    from yapt_amd import camera
    scene, info = gltf.load(str(_write_gltf(tmp_path)))
    cam = camera.Camera(eye=(0, 14, 12), target=(1, 4, 2), fovy=0.9)
    frame = camera.make_frame(cam, 64, 48, max_depth=5)
    it = _lib.Integrator(_lib.YPT_INTEGRATOR_NEE, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 4)
    g = it.film()
    it.close()
    c = oracle.render(scene, frame, _lib.YPT_INTEGRATOR_NEE, 0, 4)["film"]
    assert g[..., 1].max() > 0
    assert np.allclose(g[..., :3], c[..., :3], rtol=1e-4, atol=1e-6)


def _png(rgba):
    import io
    from PIL import Image
    f = io.BytesIO()
    Image.fromarray(rgba, "RGBA").save(f, format="PNG")
    return f.getvalue()


def _textured_doc(alpha_mode, albedo_alpha, in_buffer):
    """one quad with TEXCOORD_0 and albedo, ORM (occlusion + metallic-roughness) and normal textures"""
    rng = np.random.default_rng(3)
    albedo = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    albedo[..., 3] = albedo_alpha
    orm = rng.integers(0, 256, (2, 2, 4)).astype(np.uint8)
    normal = np.full((2, 2, 4), 128, np.uint8)
    normal[..., 2] = 255
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    UV = np.array([[0, 0], [2, 0], [2, 1], [0, 1]], np.float32)
    I = np.array([0, 2, 1, 0, 3, 2], np.uint16)
    blob = bytearray(P.tobytes() + UV.tobytes() + I.tobytes() + b"\0\0")
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 48}, {"buffer": 0, "byteOffset": 48, "byteLength": 32},
             {"buffer": 0, "byteOffset": 80, "byteLength": 12}]
    images = []
    for name, img in (("base", albedo), ("orm", orm), ("normalmap", normal)):
        data = _png(img)
        if in_buffer:
            views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data)})
            images.append({"name": name, "bufferView": len(views) - 1, "mimeType": "image/png"})
            blob += data + b"\0" * ((4 - len(data) % 4) % 4)
        else:
            images.append({"name": name, "uri": "data:image/png;base64," + base64.b64encode(data).decode()})
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 0}]}],
        "materials": [{"name": "tex", "alphaMode": alpha_mode, "alphaCutoff": 0.3,
                       "normalTexture": {"index": 2}, "occlusionTexture": {"index": 1},
                       "pbrMetallicRoughness": {"baseColorFactor": [1, 1, 1, 0.5], "baseColorTexture": {"index": 0},
                                                "metallicRoughnessTexture": {"index": 1}}}],
        "textures": [{"source": 0}, {"source": 1}, {"source": 2}], "images": images,
        "bufferViews": views,
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC2"},
                      {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR"}],
    }
    return doc, bytes(blob), (albedo, orm, normal)


def _save(tmp_path, doc, blob, glb):
    if glb:
        doc["buffers"] = [{"byteLength": len(blob)}]
        js = json.dumps(doc).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        p = tmp_path / "tex.glb"
        p.write_bytes(b"glTF" + struct.pack("<II", 2, 12 + len(body)) + body)
    else:
        doc["buffers"] = [{"byteLength": len(blob), "uri": "data:application/octet-stream;base64,"
                           + base64.b64encode(blob).decode()}]
        p = tmp_path / "tex.gltf"
        p.write_text(json.dumps(doc))
    return str(p)


@pytest.mark.parametrize("glb", [False, True])
def test_textures_are_decoded_and_bound(tmp_path, glb):
    doc, blob, (albedo, orm, normal) = _textured_doc("OPAQUE", 255, in_buffer=glb)
    scene, info = gltf.load(_save(tmp_path, doc, blob, glb))
    m = scene.materials[0]
    assert not info["images_failed"]
    slots = {"albedo": m.albedo_tex, "orm": m.orm_tex, "normal": m.normal_tex}
    assert sorted(slots.values()) == [0, 1, 2] and m.emissive_tex == -1
    assert info["textures"]["tex"] == slots
    assert list(m.albedo_tex_scale) == [1.0, 1.0] and list(m.orm_tex_scale) == [1.0, 1.0]
    texels, desc = scene.texel_arrays, scene.texture_array
    for slot, img, srgb in (("albedo", albedo, 1), ("orm", orm, 0), ("normal", normal, 0)):
        # colour is sRGB; metallic-roughness, occlusion and normal maps are data
        assert np.array_equal(texels[slots[slot]], img) and desc[slots[slot]].srgb == srgb, slot
    assert np.allclose(scene.uvs, [[0, 0], [2, 0], [2, 1], [0, 1]])
    assert abs(m.transparency) < 1e-7 and m.alpha_cutoff == 0.0


def test_alpha_modes_follow_the_reference_loader(tmp_path):
    def load(mode, alpha):
        doc, blob, _ = _textured_doc(mode, alpha, in_buffer=False)
        return gltf.load(_save(tmp_path, doc, blob, False))[0].materials[0]
    m = load("MASK", 0)                                   # texels below the cutoff: alpha testing at the cutoff
    assert abs(m.alpha_cutoff - 0.3) < 1e-7 and not (m.material_mask & gltf.MASK_ALPHA_BLEND)
    m = load("MASK", 200)                                 # every texel passes: stays opaque
    assert m.alpha_cutoff == 0.0
    m = load("BLEND", 2)                                  # cutouts: alpha testing at 1/2
    assert m.alpha_cutoff == 0.5 and abs(m.transparency) < 1e-7
    m = load("BLEND", 128)                                # tinted coverage: per-texel alpha, factor 0.5 folded in
    assert m.alpha_cutoff == 0.0 and m.material_mask & gltf.MASK_ALPHA_BLEND
    assert abs(m.transparency - 0.5) < 1e-7 and m.material_mask & gltf.MASK_TWO_SIDED
