"""glTF 2.0 scenes (.gltf with external or data-URI buffers, .glb) as a resident-ready scene.

The reference loads glTF through pygltflib in its Python scene loader (pythonScripts/yapt/scene_loader.py)
and its C++ mesh loader; pygltflib is not part of this build, so the container, accessors and node
hierarchy are parsed here. What follows the reference:
  * materials (scene_loader.py:105-243): base colour factor -> albedo, transparency 0, metallic and
    roughness factors (default 1), emissive factor x KHR_materials_emissive_strength, doubleSided ->
    two sided, KHR_materials_transmission / specular / clearcoat (IOR 1.5) / sheen / volume
    (absorption = -ln(attenuationColor) / attenuationDistance) / iridescence / ior, and BLEND coverage
    folded into the transparency (1 - coverage * (1 - transmission)); primitives without a material get
    the PLASTIC preset with specular amount 0;
  * KHR_lights_punctual (scene_loader.py:362-390): colour, intensity / 683 (lm/W at 555 nm), spot cone
    angles; the light looks down its node's -Z axis;
  * node transforms (translation, rotation quaternion, scale, or a matrix) composed down the hierarchy
    and baked into world space, normals by the inverse transpose, tangents by the matrix (their
    handedness flipped, and the winding reversed, under a mirroring transform).
  * textures (scene_loader.py:39-59, 88-99, 286-309): images decoded to RGBA8, sRGB unless an image
    holds data (a normal, occlusion or metallic-roughness map, or an image named "normal"), a uv scale
    of (1, 1) and TEXCOORD_0; occlusion and metallic-roughness share the one ORM slot, the
    metallic-roughness texture winning; MASK and BLEND alpha become alpha testing or per-texel
    coverage by the smallest texel alpha (scene_loader.py:150-173).
The reference loads external image files only; images in a bufferView or a data URI are decoded
here as well, and KTX images (KHR_texture_basisu sources included) go through ktx.py, keeping the
encoding that their format states. Triangle lists (mode 4) only.

    scene, info = load("model.glb")
"""
import base64
import io
import json
import math
import os
import struct

import numpy as np

from . import _lib, ktx, scenes

GLTF_LUMENS_PER_WATT = 683.0
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_LIGHT_TYPES = {"point": _lib.Light.POINT, "spot": _lib.Light.SPOT, "directional": _lib.Light.DIRECTIONAL}
MASK_TWO_SIDED, MASK_ALPHA_BLEND = _lib.YPT_MATERIAL_TWO_SIDED, _lib.YPT_MATERIAL_ALPHA_BLEND


def _read_container(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"glTF":
        version, length = struct.unpack_from("<II", data, 4)
        if version != 2:
            raise ValueError("glTF version %d is not supported" % version)
        off, doc, binchunk = 12, None, None
        while off < length:
            clen, ctype = struct.unpack_from("<II", data, off)
            chunk = data[off + 8: off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                binchunk = chunk
            off += 8 + clen
        return doc, binchunk
    return json.loads(data.decode("utf-8")), None


def _buffers(doc, base, binchunk):
    out = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if binchunk is None:
                raise ValueError("buffer %d has no uri and the file has no BIN chunk" % i)
            out.append(binchunk)
        elif uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                out.append(f.read())
    return out


def _accessor(doc, bufs, index):
    a = doc["accessors"][index]
    dtype = np.dtype(_COMPONENT[a["componentType"]])
    n, count = _NCOMP[a["type"]], a["count"]
    if "bufferView" not in a:
        return np.zeros((count, n), np.float32)
    bv = doc["bufferViews"][a["bufferView"]]
    raw = bufs[bv["buffer"]]
    start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    elem = dtype.itemsize * n
    stride = bv.get("byteStride", 0) or elem
    if stride == elem:
        arr = np.frombuffer(raw, dtype, count * n, start).reshape(count, n)
    else:
        rows = np.frombuffer(raw, np.uint8, (count - 1) * stride + elem, start)
        arr = np.stack([np.frombuffer(rows[i * stride: i * stride + elem].tobytes(), dtype) for i in range(count)])
    arr = arr.astype(np.float64)
    if a.get("normalized") and dtype != np.float32:
        arr = np.maximum(arr / np.iinfo(dtype).max, -1.0)
    return arr


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_matrix(node):
    """local transform of a node: its matrix (column-major in the file) or T * R * S"""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T
    m = np.eye(4)
    m[:3, :3] = _quat_matrix(node.get("rotation", [0, 0, 0, 1])) * np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    m[:3, 3] = node.get("translation", [0, 0, 0])
    return m


def _linear_images(doc):
    """scene_loader.py:286-300: images used as normal, occlusion or metallic-roughness maps hold data"""
    out, texs = set(), doc.get("textures", [])
    for mat in doc.get("materials", []):
        defs = [mat.get("normalTexture"), mat.get("occlusionTexture"),
                (mat.get("pbrMetallicRoughness") or {}).get("metallicRoughnessTexture")]
        for d in defs:
            if d is not None and d.get("index") is not None and d["index"] < len(texs):
                src = texs[d["index"]].get("source")
                if src is not None:
                    out.add(src)
    return out


class _Images:
    """glTF images decoded on first use into RGBA8 texture slots (scene_loader.py:39-59, 302-309)"""

    def __init__(self, doc, bufs, base):
        self.doc, self.bufs, self.base = doc, bufs, base
        self.linear = _linear_images(doc)
        self.slot, self.textures, self.min_alpha, self.failed = {}, [], [], []

    def _bytes(self, img):
        uri = img.get("uri")
        if uri is not None:
            if uri.startswith("data:"):
                return base64.b64decode(uri.split(",", 1)[1])
            with open(os.path.join(self.base, uri), "rb") as f:
                return f.read()
        bv = self.doc["bufferViews"][img["bufferView"]]
        off = bv.get("byteOffset", 0)
        return bytes(self.bufs[bv["buffer"]][off: off + bv["byteLength"]])

    def texture(self, tex_index):
        """texture slot of a glTF texture, -1 if its image cannot be loaded (the reference then keeps the
        material untextured, scene_loader.py:88-99)"""
        texs, imgs = self.doc.get("textures", []), self.doc.get("images", [])
        if tex_index is None or tex_index >= len(texs):
            return -1
        tex = texs[tex_index]
        src = tex.get("source")
        if src is None:                               # KHR_texture_basisu names its KTX2 image here
            src = (tex.get("extensions") or {}).get("KHR_texture_basisu", {}).get("source")
        if src is None or src >= len(imgs):
            return -1
        if src not in self.slot:
            img = imgs[src]
            name = img.get("name") or ""
            srgb = not (src in self.linear or "normal" in name.casefold())
            try:
                raw = self._bytes(img)
                if ktx.is_ktx(raw):                   # a KTX file states its own encoding
                    rgba, srgb = ktx.decode(raw)
                else:
                    from PIL import Image
                    rgba = np.asarray(Image.open(io.BytesIO(raw)).convert("RGBA"), np.uint8)
            except Exception as e:                    # an unreadable image leaves the slot unbound
                self.failed.append((img.get("name") or img.get("uri") or src, str(e)))
                self.slot[src] = -1
                return -1
            self.slot[src] = len(self.textures)
            self.textures.append((np.ascontiguousarray(rgba), srgb))
            self.min_alpha.append(float(rgba[..., 3].min()) / 255.0)
        return self.slot[src]

    def texture_min_alpha(self, slot):
        """smallest texel alpha, 0 when the image is unknown so that alpha testing stays on
        (scene_loader.py:280-284)"""
        return self.min_alpha[slot] if slot >= 0 else 0.0


def _bind(m, slot_name, slot):
    if slot >= 0:
        setattr(m, slot_name + "_tex", slot)
        getattr(m, slot_name + "_tex_scale")[:] = (1.0, 1.0)


def _material(doc, mat, images):
    """scene_loader.py _load_material_gltf (105-243)"""
    m = scenes.blank_material()
    ext = mat.get("extensions", {}) or {}
    strength = ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0)
    if mat.get("emissiveFactor"):
        m.emissive[:] = [c * strength for c in mat["emissiveFactor"]]
    if mat.get("doubleSided"):
        m.material_mask |= MASK_TWO_SIDED
    coverage, blended, transmission = 1.0, False, 0.0
    pbr = mat.get("pbrMetallicRoughness")
    if pbr is not None:
        color = pbr.get("baseColorFactor", [1, 1, 1, 1])
        m.albedo[:] = color[:3]
        m.transparency = 0.0
        has_tex = "baseColorTexture" in pbr
        albedo = images.texture(pbr["baseColorTexture"].get("index")) if has_tex else -1
        _bind(m, "albedo", albedo)
        mode = mat.get("alphaMode", "OPAQUE")
        factor_alpha = color[3] if len(color) > 3 else 1.0
        if mode == "MASK" and has_tex:
            cutoff = mat.get("alphaCutoff", 0.5)
            if images.texture_min_alpha(albedo) < cutoff:
                m.alpha_cutoff = cutoff
        elif mode == "BLEND":
            tex_min_alpha = images.texture_min_alpha(albedo) if has_tex else 1.0
            if tex_min_alpha < 0.02:                  # cutouts: alpha testing at 1/2
                m.alpha_cutoff = 0.5
            elif tex_min_alpha < 1.0 or factor_alpha < 1.0:
                coverage, blended = factor_alpha, True
                if tex_min_alpha < 1.0:
                    m.material_mask |= MASK_ALPHA_BLEND
        m.metalness = pbr.get("metallicFactor", 1.0)
        m.roughness = pbr.get("roughnessFactor", 1.0)
    if "KHR_materials_transmission" in ext:
        transmission = ext["KHR_materials_transmission"].get("transmissionFactor", 0.0)
        m.transparency = transmission
        vol = ext.get("KHR_materials_volume")
        thin = vol is None or vol.get("thicknessFactor", 0.0) == 0
        m.material_mask = (m.material_mask | MASK_TWO_SIDED) if thin else (m.material_mask & ~MASK_TWO_SIDED)
    if "KHR_materials_specular" in ext:
        m.specular_amount = ext["KHR_materials_specular"].get("specularFactor", 1.0)
    if "KHR_materials_clearcoat" in ext:
        cc = ext["KHR_materials_clearcoat"]
        m.clear_coat_amount = cc.get("clearcoatFactor", 0.0)
        m.clear_coat_roughness = cc.get("clearcoatRoughnessFactor", 0.0)
        m.clear_coat_ior = 1.5
    if "KHR_materials_sheen" in ext:
        sh = ext["KHR_materials_sheen"]
        col = sh.get("sheenColorFactor", [0.0, 0.0, 0.0])
        amount = max(col)
        if amount > 0:
            m.sheen_amount = amount
            m.sheen_tint[:] = [c / amount for c in col]
            r = sh.get("sheenRoughnessFactor", 0.0)
            m.sheen_roughness = r * r
    if "KHR_materials_volume" in ext:
        vol = ext["KHR_materials_volume"]
        if vol.get("thicknessFactor", None) == 0:
            m.material_mask |= MASK_TWO_SIDED
        dist = vol.get("attenuationDistance", -1)
        col = vol.get("attenuationColor", [1, 1, 1])
        if dist > 0 and any(c != 1 for c in col):
            m.absorption[:] = [-math.log(c) / dist for c in col]
    if "KHR_materials_iridescence" in ext:
        ir = ext["KHR_materials_iridescence"]
        m.clear_coat_ior = ir.get("iridescenceIor", 1.3)
        m.clear_coat_amount = ir.get("iridescenceFactor", 0.0)
        m.thin_film_thickness_nm = ir.get("iridescenceThicknessMaximum", 400)
    if "KHR_materials_ior" in ext:
        m.dielectric_ior = ext["KHR_materials_ior"].get("ior", 1.5)
    if blended:
        m.transparency = 1.0 - coverage * (1.0 - transmission)
        if transmission == 0.0:
            m.material_mask |= MASK_TWO_SIDED
    slots = {}
    for key, slot in (("emissiveTexture", "emissive"), ("normalTexture", "normal"), ("occlusionTexture", "orm")):
        if key in mat:
            slots[slot] = images.texture(mat[key].get("index"))
    if pbr is not None and "metallicRoughnessTexture" in pbr:   # one ORM slot: metallic-roughness wins
        slots["orm"] = images.texture(pbr["metallicRoughnessTexture"].get("index"))
    for slot, t in slots.items():
        _bind(m, slot, t)
    if pbr is not None and "baseColorTexture" in pbr:
        slots["albedo"] = m.albedo_tex
    return m, slots


def _area_normals(P, I):
    n = np.zeros_like(P)
    fn = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    for k in range(3):
        np.add.at(n, I[:, k], fn)
    l = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(l > 0, n / np.where(l > 0, l, 1), np.array([0.0, 1.0, 0.0]))


def load(path, scene_index=None):
    """Returns (SceneArrays, info) with info = {"lights": [...], "materials": [names], "textures": {material: {slot:
    texture index, -1 where the image could not be loaded}}, "images_failed": [(image, error)],
    "camera": (eye, target, yfov) of the first camera node or None, "triangles": n}"""
    doc, binchunk = _read_container(path)
    bufs = _buffers(doc, os.path.dirname(os.path.abspath(path)), binchunk)
    b = scenes.MeshBuilder()
    images = _Images(doc, bufs, os.path.dirname(os.path.abspath(path)))
    mats, tex_info, uvs = [], {}, []
    for k, mat in enumerate(doc.get("materials", [])):
        m, t = _material(doc, mat, images)
        mats.append(b.material(m))
        if t:
            tex_info[mat.get("name", "material_%d" % k)] = t
    default = None
    lights, camera = [], None
    light_defs = doc.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])
    tangents = []

    def visit(ni, parent):
        nonlocal default, camera
        node = doc["nodes"][ni]
        world = parent @ node_matrix(node)
        if "mesh" in node:
            nmat = np.linalg.inv(world[:3, :3]).T
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                if prim.get("mode", 4) != 4:
                    continue
                attr = prim["attributes"]
                P = _accessor(doc, bufs, attr["POSITION"])[:, :3]
                I = (_accessor(doc, bufs, prim["indices"]).astype(np.int64).reshape(-1, 3) if "indices" in prim
                     else np.arange(len(P)).reshape(-1, 3))
                Pw = P @ world[:3, :3].T + world[:3, 3]
                if np.linalg.det(world[:3, :3]) < 0:        # mirrored: keep the front faces facing out
                    I = I[:, [0, 2, 1]]
                if "NORMAL" in attr:
                    N = _accessor(doc, bufs, attr["NORMAL"])[:, :3] @ nmat.T
                    N = N / np.maximum(np.linalg.norm(N, axis=1, keepdims=True), 1e-30)
                else:
                    N = _area_normals(Pw, I)
                if "TANGENT" in attr:
                    T4 = _accessor(doc, bufs, attr["TANGENT"])
                    T = T4[:, :3] @ world[:3, :3].T
                    T = T / np.maximum(np.linalg.norm(T, axis=1, keepdims=True), 1e-30)
                    flip = -1.0 if np.linalg.det(world[:3, :3]) < 0 else 1.0   # mirrored: bitangent flips
                    tangents.append(np.concatenate([T, T4[:, 3:4] * flip], 1))
                else:
                    tangents.append(scenes.arbitrary_tangents(N.astype(np.float32)))
                uvs.append(_accessor(doc, bufs, attr["TEXCOORD_0"])[:, :2] if "TEXCOORD_0" in attr
                           else np.zeros((len(P), 2)))
                if "material" in prim:
                    mi = mats[prim["material"]]
                else:
                    if default is None:
                        dm = scenes.preset_material("PLASTIC")
                        dm.specular_amount = 0.0
                        default = b.material(dm)
                    mi = default
                b.add(Pw, N, I, mi)
        ext = node.get("extensions", {}).get("KHR_lights_punctual")
        if ext is not None and ext.get("light") is not None and ext["light"] < len(light_defs):
            ld = light_defs[ext["light"]]
            t = _LIGHT_TYPES.get(ld.get("type"))
            if t is not None:
                l = _lib.Light(t)
                pos = world[:3, 3]
                fwd = -world[:3, 2]
                l.lookAt(pos, pos + fwd)
                l.setColor(ld.get("color", [1.0, 1.0, 1.0]))
                l.setIntensity(ld.get("intensity", 1.0) / GLTF_LUMENS_PER_WATT)
                if t == _lib.Light.SPOT:
                    spot = ld.get("spot", {})
                    l.setSpotConeAngles(spot.get("innerConeAngle", 0.0), spot.get("outerConeAngle", math.pi / 4.0))
                lights.append(l)
        if "camera" in node and camera is None:
            cam = doc["cameras"][node["camera"]]
            if cam.get("type") == "perspective":
                eye = world[:3, 3]
                camera = (tuple(eye), tuple(eye - world[:3, 2]), cam["perspective"].get("yfov", math.pi * 0.3))
        for c in node.get("children", []):
            visit(c, world)

    si = scene_index if scene_index is not None else doc.get("scene", 0)
    roots = doc["scenes"][si]["nodes"] if doc.get("scenes") else list(range(len(doc.get("nodes", []))))
    for r in roots:
        visit(r, np.eye(4))
    if not b.pos:
        raise ValueError("no triangle geometry in %s" % path)
    nrm = np.concatenate(b.nrm)
    arrays = _lib.SceneArrays(np.concatenate(b.pos), nrm, np.concatenate(b.idx), b.inst_first, b.inst_count, b.inst_mat,
                              b.materials, tangents=np.concatenate(tangents).astype(np.float32), lights=lights,
                              uvs=np.concatenate(uvs).astype(np.float32) if images.textures else None,
                              textures=images.textures)
    info = {"lights": lights, "materials": [m.get("name") for m in doc.get("materials", [])], "textures": tex_info,
            "images_failed": images.failed, "camera": camera, "triangles": arrays.triangle_count}
    return arrays, info


def mesh_primitives(path):
    """Object-space triangle primitives of every mesh of a glTF file, in file order: a list per mesh of
    dicts with positions, normals, tangents (x, y, z, handedness), uvs (TEXCOORD_0) or None where the
    file has no such attribute, indices and the glTF material index (or None). Node transforms are
    not applied (the role of MeshLoader.create_and_upload_from_gltf, mesh_utils.py:87-97)."""
    doc, binchunk = _read_container(path)
    bufs = _buffers(doc, os.path.dirname(os.path.abspath(path)), binchunk)
    out = []
    for mesh in doc.get("meshes", []):
        prims = []
        for prim in mesh.get("primitives", []):
            if prim.get("mode", 4) != 4:
                continue
            attr = prim["attributes"]
            P = _accessor(doc, bufs, attr["POSITION"])[:, :3].astype(np.float32)
            I = (_accessor(doc, bufs, prim["indices"]).astype(np.uint32).reshape(-1) if "indices" in prim
                 else np.arange(len(P), dtype=np.uint32))
            get = lambda k, n: _accessor(doc, bufs, attr[k])[:, :n].astype(np.float32) if k in attr else None
            prims.append({"positions": P, "normals": get("NORMAL", 3), "tangents": get("TANGENT", 4),
                          "uvs": get("TEXCOORD_0", 2), "indices": I, "material": prim.get("material")})
        out.append(prims)
    return out
