"""The reference's Python scene surface over this library: the py_yapt classes a scene script uses
(Source/Python/srcBindings: PyRenderer.cpp, PyScene.cpp, PyMaterial.cpp, PyMesh.cpp,
PyRenderObject.cpp, PySceneObject.cpp, PyTransform.cpp, PyCamera.cpp, PyLight.cpp,
PyRendererVar.cpp, PyReadback.cpp) and the pythonScripts/yapt wrappers around them
(scene_wrapper.py, resources.py, scene_loader.py), so that a script written against them builds the
same scene here.

    app = Application(width=960, height=704)
    renderer, scene, resources = app.get_renderer(), app.get_scene(), app.get_resources()
    mat = renderer.createMaterial("m"); mat.setFromMaterialPreset(Material.MaterialPreset.PLASTIC)
    obj = scene.create_render_object("o"); obj.setMesh(mesh); obj.setMaterial(mat, 0)
    obj.getTransform().setTranslation(vec3([0, 1, 0]))
    renderer.render(64)                                    # 64 sample indices merged into the film
    rgb8 = renderer.readback(ReadbackTarget.FinalColor).getData()

What differs from the reference, and why:
  * Meshes are object-space arrays (createMesh(name, positions, indices, normals, uvs, tangents)),
    not GPU vertex buffers with layouts: the data is uploaded once per scene commit, where render
    objects' world transforms are applied. Textures likewise are RGBA8 arrays (createTexture).
  * The render loop is explicit: render(n) accumulates n sample indices; nothing runs in between.
  * Auto exposure and the debug pipelines (DBG_Normals, DBG_Tangents, DBG_UV) are not built;
    with Tonemap.UseAutoExposure ON the manual exposure is used.
  * Orthographic cameras are not built (the integrators trace from a pinhole or thin lens).

Instancing: each render object references a mesh and has its own transform, and a commit writes
one world-space copy of the mesh's triangles per render object (positions by the object's world
matrix, normals by its inverse transpose, tangents by the matrix with the handedness flipped and
the winding reversed under a mirroring transform, as the object-space front-face rule keeps them).
Traversal then runs over one BVH with no per-instance ray transforms. When only transforms change,
the new world-space copy goes to the library as a refit (ypt_refit_scene): the same tree with new
boxes, no rebuild.
"""
import enum
import math
import os
import sys

import numpy as np

from . import _lib
from . import camera as _camera
from . import scenes as _scenes


# ---------------------------------------------------------------- glm vector helpers
def _vec(n, v, rest):
    if rest:
        return np.array([v] + list(rest), np.float32)
    a = np.asarray(v, np.float32)
    if a.ndim == 0:
        return np.full(n, float(a), np.float32)
    if a.size != n:
        raise ValueError("vec%d needs %d components, got %d" % (n, n, a.size))
    return a.reshape(n).copy()


def vec2(v=0.0, *rest):
    return _vec(2, v, rest)


def vec3(v=0.0, *rest):
    """vec3(s), vec3([x, y, z]) or vec3(x, y, z)"""
    return _vec(3, v, rest)


def vec4(v=0.0, *rest):
    return _vec(4, v, rest)


v2, v3, v4 = vec2, vec3, vec4


# ---------------------------------------------------------------- quaternions (x, y, z, w)
def quat_mul(a, b):
    ax, ay, az, aw = [float(c) for c in a]
    bx, by, bz, bw = [float(c) for c in b]
    return np.array([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz], np.float64)


def quat_to_matrix(q):
    """glm::mat4_cast of a unit quaternion, as a 3x3 rotation (column vectors)"""
    x, y, z, w = [float(c) for c in q]
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def matrix_to_quat(m):
    """glm::quat_cast: the quaternion of a rotation matrix (the largest-component branch)"""
    m = np.asarray(m, np.float64)
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        return np.array([(m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s, 0.25 * s])
    if m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        return np.array([0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s, (m[2, 1] - m[1, 2]) / s])
    if m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        return np.array([(m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s, (m[0, 2] - m[2, 0]) / s])
    s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
    return np.array([(m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s, (m[1, 0] - m[0, 1]) / s])


def axis_angle_quat(axis, degrees):
    """the rotation QQuaternion.fromAxisAndAngle gives, as (x, y, z, w)"""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    h = math.radians(degrees) / 2
    return vec4([*(a * math.sin(h)), math.cos(h)])


# ---------------------------------------------------------------- transforms and scene objects
WORLD_RIGHT, WORLD_UP, WORLD_FORWARD = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, -1.0])


class Transform:
    """Scene/Transform.inl: translation, orientation (x, y, z, w), scale and a parent; the world
    matrix is parent * T * R * S; right/up/forward are the world matrix applied to +x, +y and -z."""

    def __init__(self, on_change=None):
        self._t = np.zeros(3)
        self._s = np.ones(3)
        self._q = np.array([0.0, 0.0, 0.0, 1.0])
        self._parent = None
        self._children = []
        self._on_change = on_change

    def _changed(self):
        if self._on_change is not None:
            self._on_change()
        for c in self._children:
            c._changed()

    def setTranslation(self, v):
        self._t = np.asarray(v, np.float64).reshape(3).copy()
        self._changed()

    def addTranslation(self, v):
        self._t = self._t + np.asarray(v, np.float64).reshape(3)
        self._changed()

    def getTranslation(self):
        return vec3(self._t)

    def setScale(self, v):
        self._s = np.asarray(vec3(v), np.float64)
        self._changed()

    def addScale(self, v):
        """Transform::addScale assigns the scale (Transform.inl), kept as is"""
        self.setScale(v)

    def getScale(self):
        return vec3(self._s)

    def setOrientation(self, q):
        q = np.asarray(q, np.float64).reshape(4)
        self._q = q / np.linalg.norm(q)
        self._changed()

    def addOrientation(self, q):
        self._q = quat_mul(self._q, np.asarray(q, np.float64).reshape(4))
        self._changed()

    def getOrientation(self):
        return vec4(self._q)

    def lookAt(self, position, at, up):
        """inverse(glm::lookAt): the -z axis towards `at`, scale 1"""
        view = _camera.look_at(position, at, up).astype(np.float64)
        inv = np.linalg.inv(view)
        self._t = np.asarray(position, np.float64).reshape(3).copy()
        self._q = matrix_to_quat(inv[:3, :3])
        self._s = np.ones(3)
        self._changed()

    def local_matrix(self):
        m = np.eye(4)
        m[:3, :3] = quat_to_matrix(self._q) * self._s
        m[:3, 3] = self._t
        return m

    def getMatrixWS(self):
        m = self.local_matrix()
        return m if self._parent is None else self._parent.getMatrixWS() @ m

    def right(self):
        return vec3(self.getMatrixWS()[:3, :3] @ WORLD_RIGHT)

    def up(self):
        return vec3(self.getMatrixWS()[:3, :3] @ WORLD_UP)

    def forward(self):
        return vec3(self.getMatrixWS()[:3, :3] @ WORLD_FORWARD)

    def world_position(self):
        return self.getMatrixWS()[:3, 3].copy()

    def setParent(self, parent):
        if parent is self:
            return
        if self._parent is not None:
            self._parent._children.remove(self)
        self._parent = parent
        if parent is not None:
            parent._children.append(self)
        self._changed()

    def getId(self):
        return id(self)

    def getParentId(self):
        return 0 if self._parent is None else id(self._parent)


class SceneObject:
    def __init__(self, scene, name, on_change=None):
        self._scene = scene
        self._name = name
        self._transform = Transform(on_change or scene._touch_transform)

    def getName(self):
        return self._name

    def getTransform(self):
        return self._transform


class RenderObject(SceneObject):
    """PyRenderObject.cpp: a mesh and one material per submesh"""

    def __init__(self, scene, name):
        super().__init__(scene, name)
        self._mesh = None
        self._materials = {}

    def setMesh(self, mesh):
        self._mesh = mesh
        self._scene._touch()

    def getMesh(self):
        return self._mesh

    def setMaterial(self, material, submesh_index=0):
        self._materials[int(submesh_index)] = material
        self._scene._touch()

    def setMaterials(self, materials):
        self._materials = {k: m for k, m in enumerate(materials)}
        self._scene._touch()

    def getMaterials(self):
        n = self._mesh.getSubmeshCount() if self._mesh is not None else (max(self._materials) + 1 if self._materials else 0)
        return [self._materials.get(k) for k in range(n)]


class LightType(enum.IntEnum):
    POINT = _lib.YPT_LIGHT_POINT
    SPOT = _lib.YPT_LIGHT_SPOT
    DIRECTIONAL = _lib.YPT_LIGHT_DIRECTIONAL


class Light(SceneObject):
    """PyLight.cpp / LightProxy: type, colour, intensity and spot cone angles; the light sits at its
    transform's world position and shines along its forward (-z) axis"""
    LightType = LightType

    def __init__(self, scene, name, light_type=LightType.POINT):
        super().__init__(scene, name)
        self._light = _lib.Light(int(light_type))

    def setType(self, t):
        self._light.setType(int(t))
        self._scene._touch()

    def getType(self):
        return LightType(self._light.type)

    def setColor(self, rgb):
        self._light.setColor(vec3(rgb))
        self._scene._touch()

    def getColor(self):
        return vec3(self._light.color)

    def setIntensity(self, v):
        self._light.setIntensity(v)
        self._scene._touch()

    def getIntensity(self):
        return self._light.intensity

    def setSpotConeAngles(self, inner, outer):
        self._light.setSpotConeAngles(inner, outer)
        self._scene._touch()

    def getSpotInnerAngle(self):
        return self._light.spot_inner

    def getSpotOuterAngle(self):
        return self._light.spot_outer

    def _world_light(self):
        t = self._transform
        pos = t.world_position()
        self._light.lookAt(pos, pos + np.asarray(t.forward(), np.float64))
        return self._light


class ProjectionType(enum.IntEnum):
    PERSPECTIVE = 0
    ORTHOGRAPHIC = 1


class Camera(SceneObject):
    """PyCamera.cpp / Camera.inl: a perspective camera at its transform, looking down -z; default
    field of view 0.3 pi, near 0.1, far 1000"""
    ProjectionType = ProjectionType

    def __init__(self, scene, name="main_camera"):
        super().__init__(scene, name, on_change=lambda: None)
        self._fovy, self._near, self._far, self._aspect = math.pi * 0.3, 0.1, 1000.0, 16.0 / 9.0
        self._ortho_size = 1.0
        self._projection = ProjectionType.PERSPECTIVE

    def _tf(self):
        self._scene._renderer._touch_frame()

    def setFOVY(self, v):
        self._fovy = float(v)
        self._tf()

    def getFOVY(self):
        return self._fovy

    def setNearPlane(self, v):
        self._near = float(v)
        self._tf()

    def getNearPlane(self):
        return self._near

    def setFarPlane(self, v):
        self._far = float(v)
        self._tf()

    def getFarPlane(self):
        return self._far

    getFarPlance = getFarPlane    # the binding's spelling (PyCamera.cpp)

    def setAspectRatio(self, v):
        self._aspect = float(v)
        self._tf()

    def getAspectRatio(self):
        return self._aspect

    def setOrthographicSize(self, v):
        self._ortho_size = float(v)

    def getOrthographicSize(self):
        return self._ortho_size

    def setProjectionType(self, t):
        self._projection = ProjectionType(int(t))
        self._tf()

    def getProjectionType(self):
        return self._projection

    def setProjectionPerspective(self, fovy=None, aspect=None, near=None, far=None):
        self._projection = ProjectionType.PERSPECTIVE
        for name, v in (("_fovy", fovy), ("_aspect", aspect), ("_near", near), ("_far", far)):
            if v is not None:
                setattr(self, name, float(v))
        self._tf()

    def setProjectionOrthographic(self, size=None, aspect=None, near=None, far=None):
        self._projection = ProjectionType.ORTHOGRAPHIC
        for name, v in (("_ortho_size", size), ("_aspect", aspect), ("_near", near), ("_far", far)):
            if v is not None:
                setattr(self, name, float(v))
        self._tf()

    def _camera(self, aperture_radius, focus_distance):
        if self._projection != ProjectionType.PERSPECTIVE:
            raise NotImplementedError("orthographic cameras are not supported by the integrators")
        t = self._transform
        eye = t.world_position()
        return _camera.Camera(eye=tuple(eye), target=tuple(eye + np.asarray(t.forward(), np.float64)),
                              up=tuple(np.asarray(t.up(), np.float64)), fovy=self._fovy, near=self._near,
                              far=self._far, aperture_radius=aperture_radius, focus_distance=focus_distance)


# ---------------------------------------------------------------- materials, meshes, textures
class MaterialPreset(enum.IntEnum):
    BLANK = 0
    METAL_GOLD = 1
    METAL_SILVER = 2
    METAL_COPPER = 3
    METAL_BRASS = 4
    METAL_ALUMINIUM = 5
    GLASS = 6
    PLASTIC = 7


_TEX_SLOTS = ("albedo", "normal", "orm", "emissive")


class Material:
    """PyMaterial.cpp / MaterialProxy.cpp: MaterialParameters with the reference's setters; the
    validation and packing clamps are applied by the library when the scene is committed"""
    MaterialPreset = MaterialPreset

    def __init__(self, renderer, name):
        self._renderer = renderer
        self._name = name
        self._p = _scenes.blank_material()
        self._tex = {}

    def _touch(self):
        self._renderer._touch()

    def getName(self):
        return self._name

    def setFromMaterialPreset(self, preset):
        """the preset replaces every parameter, texture bindings included"""
        self._p = _scenes.preset_material(MaterialPreset(int(preset)).name)
        self._tex = {}
        self._touch()

    def _set3(self, field, v):
        getattr(self._p, field)[:] = [float(c) for c in vec3(v)]
        self._touch()

    def _set(self, field, v):
        setattr(self._p, field, float(v))
        self._touch()

    def _mask(self, bit, on):
        self._p.material_mask = (self._p.material_mask | bit) if on else (self._p.material_mask & ~bit)
        self._touch()

    def setAlbedo(self, v): self._set3("albedo", v)
    def getAlbedo(self): return vec3(list(self._p.albedo))
    def setSpecularTint(self, v): self._set3("specular", v)
    def getSpecularTint(self): return vec3(list(self._p.specular))
    def setAbsorption(self, v): self._set3("absorption", v)
    def getAbsorption(self): return vec3(list(self._p.absorption))
    def setEmission(self, v): self._set3("emissive", v)
    def getEmission(self): return vec3(list(self._p.emissive))
    def setSheenTint(self, v): self._set3("sheen_tint", v)
    def getSheenTint(self): return vec3(list(self._p.sheen_tint))
    def setSubsurfaceRadius(self, v): self._set3("subsurface_radius", v)
    def getSubsurfaceRadius(self): return vec3(list(self._p.subsurface_radius))
    def setTransparency(self, v): self._set("transparency", v)
    def getTransparency(self): return self._p.transparency
    def setMetalness(self, v): self._set("metalness", v)
    def getMetalness(self): return self._p.metalness
    def setDielectricIOR(self, v): self._set("dielectric_ior", v)
    def getDielectricIOR(self): return self._p.dielectric_ior
    def setRoughness(self, v): self._set("roughness", v)
    def getRoughness(self): return self._p.roughness
    def setAnisotropy(self, v): self._set("anisotropy", v)
    def getAnisotropy(self): return self._p.anisotropy
    def setAnisotropyRotation(self, v): self._set("anisotropy_rotation", v)
    def getAnisotropyRotation(self): return self._p.anisotropy_rotation
    def setSpecularAmount(self, v): self._set("specular_amount", v)
    def getSpecularAmount(self): return self._p.specular_amount
    def setClearCoatAmount(self, v): self._set("clear_coat_amount", v)
    def getClearCoatAmount(self): return self._p.clear_coat_amount
    def setClearCoatIOR(self, v): self._set("clear_coat_ior", v)
    def getClearCoatIOR(self): return self._p.clear_coat_ior
    def setClearCoatRoughness(self, v): self._set("clear_coat_roughness", v)
    def getClearCoatRoughness(self): return self._p.clear_coat_roughness
    def setSheenRoughness(self, v): self._set("sheen_roughness", v)
    def getSheenRoughness(self): return self._p.sheen_roughness
    def setSheenAmount(self, v): self._set("sheen_amount", v)
    def getSheenAmount(self): return self._p.sheen_amount
    def setThinFilmThicknessNM(self, v): self._set("thin_film_thickness_nm", v)
    def getThinFilmThicknessNM(self): return self._p.thin_film_thickness_nm
    def setEmissionFocus(self, v): self._set("emission_focus", v)
    def getEmissionFocus(self): return self._p.emission_focus
    def setAlphaCutoff(self, v): self._set("alpha_cutoff", v)
    def getAlphaCutoff(self): return self._p.alpha_cutoff
    def setSubsurfaceAmount(self, v): self._set("subsurface_amount", v)
    def getSubsurfaceAmount(self): return self._p.subsurface_amount
    def setSubsurfaceRadiusScale(self, v): self._set("subsurface_radius_scale", v)
    def getSubsurfaceRadiusScale(self): return self._p.subsurface_radius_scale

    # extensions with no counterpart in the reference (0 turns them off)
    def setBlackbodyTemperature(self, kelvin): self._set("blackbody_kelvin", kelvin)
    def getBlackbodyTemperature(self): return self._p.blackbody_kelvin

    def setSellmeierCoefficients(self, b, c):
        self._p.sellmeier_b[:] = [float(x) for x in vec3(b)]
        self._p.sellmeier_c[:] = [float(x) for x in vec3(c)]
        self._touch()

    def getSellmeierCoefficients(self):
        return vec3(list(self._p.sellmeier_b)), vec3(list(self._p.sellmeier_c))

    def setCauchysCoefficients(self, v):
        self._p.cauchy_coeffs[:] = [float(c) for c in vec2(v)]
        self._touch()

    def getCauchysCoefficients(self):
        return vec2(list(self._p.cauchy_coeffs))

    def setTwoSided(self, on): self._mask(_lib.YPT_MATERIAL_TWO_SIDED, on)
    def getTwoSided(self): return bool(self._p.material_mask & _lib.YPT_MATERIAL_TWO_SIDED)
    def setEnableDispersion(self, on): self._mask(_lib.YPT_MATERIAL_DISPERSION, on)
    def getEnableDispersion(self): return bool(self._p.material_mask & _lib.YPT_MATERIAL_DISPERSION)
    def setAlphaBlend(self, on): self._mask(_lib.YPT_MATERIAL_ALPHA_BLEND, on)
    def getAlphaBlend(self): return bool(self._p.material_mask & _lib.YPT_MATERIAL_ALPHA_BLEND)

    def _bind(self, slot, tex, scale):
        if tex is None:
            self._tex.pop(slot, None)
        else:
            self._tex[slot] = (tex, vec2(1.0) if scale is None else vec2(scale))
        self._touch()

    def setAlbedoTexture(self, tex, scale=None): self._bind("albedo", tex, scale)
    def setNormalTexture(self, tex, scale=None): self._bind("normal", tex, scale)
    def setORMTexture(self, tex, scale=None): self._bind("orm", tex, scale)
    def setEmissiveTexture(self, tex, scale=None): self._bind("emissive", tex, scale)

    def record(self, texture_index):
        """the ypt_material of this material, with its textures as indices from texture_index(tex)"""
        m = _lib.Material.from_buffer_copy(self._p)
        for slot in _TEX_SLOTS:
            bound = self._tex.get(slot)
            setattr(m, slot + "_tex", -1 if bound is None else texture_index(bound[0]))
            if bound is not None:
                getattr(m, slot + "_tex_scale")[:] = [float(c) for c in bound[1]]
        return m


class ResourceDimension(enum.IntEnum):
    """the texture dimensions of PyRendererBasicTypes.cpp that textures here take"""
    TEXTURE_2D = 3
    TEXTURE_CUBEMAP = 5


class Texture:
    """An RGBA8 image (srgb: the texels are sRGB-encoded colour, an _SRGB format), or a linear float image
    (an _SFLOAT format: H x W x 4, or six square faces +X, -X, +Y, -Y, +Z, -Z for a cube map), the kind an
    HDR environment is. Materials bind RGBA8 textures; the World.Skycube variable takes either."""

    def __init__(self, name, texels, srgb=True, cube=False):
        a = np.asarray(texels)
        self._name = name
        self.is_float = a.dtype.kind == "f"
        self.dimension = ResourceDimension.TEXTURE_CUBEMAP if cube else ResourceDimension.TEXTURE_2D
        if self.is_float:
            kind = _lib.YPT_ENVIRONMENT_CUBE if cube else _lib.YPT_ENVIRONMENT_LONGLAT
            self.texels, self.srgb = _lib.environment_texels(kind, a), False
        else:
            if cube:
                raise ValueError("cube maps take float texels (6 x N x N x 4)")
            a = np.ascontiguousarray(a, dtype=np.uint8)
            if a.ndim != 3 or a.shape[2] != 4:
                raise ValueError("textures are H x W x 4 uint8")
            self.texels, self.srgb = a, bool(srgb)

    def getName(self):
        return self._name

    def getDimension(self):
        return self.dimension

    def min_alpha(self):
        if self.is_float:
            return None
        return float(self.texels[..., 3].min()) / 255.0

    def environment(self):
        """(YPT_ENVIRONMENT_*, linear float texels) for this texture as the sky: a 2D texture is a long-lat
        image (RaytraceCommonResources.cpp:290-303); RGBA8 texels are decoded (sRGB or unorm) first"""
        if self.dimension == ResourceDimension.TEXTURE_CUBEMAP:
            return _lib.YPT_ENVIRONMENT_CUBE, self.texels
        if self.is_float:
            return _lib.YPT_ENVIRONMENT_LONGLAT, self.texels
        c = self.texels.astype(np.float64) / 255.0
        if self.srgb:
            c[..., :3] = np.where(c[..., :3] <= 0.04045, c[..., :3] / 12.92, ((c[..., :3] + 0.055) / 1.055) ** 2.4)
        return _lib.YPT_ENVIRONMENT_LONGLAT, c.astype(np.float32)


class Mesh:
    """object-space vertex arrays and submeshes (index offset, index count, vertex offset), the parts
    of Mesh.setSubmesh the integrators read"""

    def __init__(self, name, positions, indices, normals=None, uvs=None, tangents=None, submeshes=None):
        self._name = name
        self.positions = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, np.int64).reshape(-1)
        if len(self.indices) % 3:
            raise ValueError("indices must be whole triangles")
        nv = len(self.positions)
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(nv, 3)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, np.float64).reshape(nv, 2)
        self.tangents = None if tangents is None else np.ascontiguousarray(tangents, np.float64).reshape(nv, 4)
        self._submeshes = list(submeshes) if submeshes else [(0, len(self.indices), 0)]

    def getName(self):
        return self._name

    def getSubmeshCount(self):
        return len(self._submeshes)

    def setSubmesh(self, index, index_offset, index_count, vertex_offset, bounds_min=None, bounds_max=None):
        while len(self._submeshes) <= index:
            self._submeshes.append((0, 0, 0))
        self._submeshes[index] = (int(index_offset), int(index_count), int(vertex_offset))

    def submesh_triangles(self, k):
        off, cnt, voff = self._submeshes[k]
        if cnt % 3:
            raise ValueError("submesh %d of %s is not whole triangles" % (k, self._name))
        return self.indices[off: off + cnt].reshape(-1, 3) + voff


# ---------------------------------------------------------------- renderer variables and readback
class RendererVariableType(enum.IntEnum):
    FLOAT = 0
    INT = 1
    TEXTURE = 2
    BUFFER = 3
    OPTIONS = 4


class RendererVariable:
    """PyRendererVar.cpp: a numeric variable with limits (values are clamped), or an options list"""
    Type = RendererVariableType

    def __init__(self, renderer, name, kind, default, lo=None, hi=None, options=None):
        self._renderer, self._name, self._kind = renderer, name, kind
        self._lo = None if lo is None else np.atleast_1d(np.asarray(lo, np.float64))
        self._hi = None if hi is None else np.atleast_1d(np.asarray(hi, np.float64))
        self._options = list(options or [])
        self._value = np.atleast_1d(np.asarray(default, np.float64)).copy()

    def getType(self):
        return self._kind

    def getComponentCount(self):
        return len(self._value)

    def _set(self, values):
        v = np.asarray(values, np.float64).reshape(-1)
        if len(v) != len(self._value):
            raise ValueError("%s has %d components" % (self._name, len(self._value)))
        if self._lo is not None:
            v = np.minimum(np.maximum(v, self._lo), self._hi)
        self._value = v
        self._renderer._touch_frame()

    def setTexture(self, texture):
        """a TEXTURE variable (World.Skycube): the texture, or None to clear it"""
        if self._kind != RendererVariableType.TEXTURE:
            raise TypeError("%s is not a texture variable" % self._name)
        if texture is not None and not isinstance(texture, Texture):
            raise TypeError("%s takes a Texture" % self._name)
        self._texture = texture
        self._renderer._touch()     # the sky travels with the committed scene

    def getTexture(self):
        return getattr(self, "_texture", None)

    def setFromFloatArray(self, values):
        if self._kind != RendererVariableType.FLOAT:
            raise TypeError("%s is not a float variable" % self._name)
        self._set(values)

    def setFromIntArray(self, values):
        if self._kind != RendererVariableType.INT:
            raise TypeError("%s is not an int variable" % self._name)
        self._set([int(x) for x in values])

    def getAsFloatArray(self):
        return [float(x) for x in self._value]

    def getAsIntArray(self):
        return [int(x) for x in self._value]

    def getLimitsFloat(self):
        return [float(x) for x in self._lo], [float(x) for x in self._hi]

    getLimitsInt = getLimitsFloat

    def getOptions(self):
        return list(self._options)

    def getSelectedOption(self):
        return int(self._value[0])

    def setSelectedOption(self, index):
        if self._kind != RendererVariableType.OPTIONS:
            raise TypeError("%s is not an options variable" % self._name)
        if not 0 <= int(index) < len(self._options):
            raise ValueError("%s has %d options" % (self._name, len(self._options)))
        self._value = np.array([int(index)], np.float64)
        self._renderer._touch_frame()

    def selected(self):
        return self._options[self.getSelectedOption()]


def _renderer_variables(r):
    F, I, O, T = RendererVariableType.FLOAT, RendererVariableType.INT, RendererVariableType.OPTIONS, RendererVariableType.TEXTURE
    on_off = ("OFF", "ON")
    defs = [   # RendererVarsList.h:21-54
        ("World.Skycube", T, 0, None, None, None),
        ("World.EnvIntensityScale", F, 1.0, 0.0, 100000.0, None),
        ("Generic.RenderResolution", I, (1920, 1080), (2, 2), (99999999, 99999999), None),
        ("RenderPipeline", O, 1, None, None, ("BPT", "BPT_WITH_NEE", "BDPT", "DBG_Normals", "DBG_Tangents", "DBG_UV")),
        ("Denoise.Mode", O, 0, None, None, ("OFF", "ON", "DEBUG_MAT_PARAMS", "DEBUG_VARIANCE", "DEBUG_SHOW_SAMPLE_SOURCES",
                                            "DEBUG_FEATURE_VARIANCE")),
        ("Denoise.DepthSigmaScale", F, 0.02, 0.00001, 10.0, None),
        ("Denoise.NormalSigma", F, 0.06, 0.00001, 1000.0, None),
        ("Denoise.MaterialDiffSigma", F, 0.05, 0.00001, 1000.0, None),
        ("Denoise.ColorDiffSigma", F, 3.0, 0.00001, 1000.0, None),
        ("Denoise.FeatureVarianceScale", F, 1.0, 0.0, 1000.0, None),
        ("Denoise.FeaturePrefilter", O, 1, None, None, on_off),
        ("Camera.ApertureRadius", F, 0.0, 0.0, 1000.0, None),
        ("Camera.FocusDistance", F, 10.0, 0.001, 1000000.0, None),
        ("Firefly.RegularizeAfterVertices", I, 8, 0, 64, None),
        ("Firefly.FireflyClampScale", F, 12.0, 0.0, 100000.0, None),
        ("Firefly.FireflyClampRelaxPower", F, 0.5, 0.0, 4.0, None),
        ("Firefly.KillNaNSamples", O, 1, None, None, on_off),
        ("Tonemap.Enable", O, 1, None, None, on_off),
        ("Tonemap.Toe", F, 1.0, 0.0, 100.0, None),
        ("Tonemap.Mid", F, 1.0, 0.0, 100.0, None),
        ("Tonemap.Shoulder", F, 1.0, 0.0, 100.0, None),
        ("Tonemap.UseAutoExposure", O, 1, None, None, on_off),
        ("Tonemap.ExposureCompensation", F, 0.0, -10000.0, 10000.0, None),
        ("Tonemap.EyeAdaptationSpeed", F, 1.0, 0.0, 1000.0, None),
        ("Tonemap.ManualExposure", F, 1.0, 0.0, 9999999999999999.0, None),
        # the bounce limit of the path loop (the reference fixes maxRayDepth = 32,
        # RaytraceCommonResources.cpp:80) and BDPT's light subpath length, exposed here as variables
        ("_DEBUG.BDPT.EffectiveLightPathNodes", I, (0, 128), (0, 0), (99999999, 99999999), None),
        ("_DEBUG.BDPT.EffectiveCameraPathNodes", I, (0, 128), (0, 0), (99999999, 99999999), None),
        ("Integrator.MaxDepth", I, 32, 1, 64, None),
        ("Integrator.LightPathVertices", I, 16, 1, 16, None),
    ]
    return {name: RendererVariable(r, name, kind, d, lo, hi, opts) for name, kind, d, lo, hi, opts in defs}


class ReadbackTarget(enum.IntEnum):
    FinalColor = 0


class ReadbackState(enum.IntEnum):
    Created = 0
    Pending = 1
    Ready = 2
    Failed = 3
    Freed = 4


class Readback:
    """PyReadback.cpp: the FinalColor target as RGBA8 (sRGB-encoded), ready at once"""

    def __init__(self, rgba8):
        self._data = rgba8
        self._state = ReadbackState.Ready

    def getState(self):
        return self._state

    def isReady(self):
        return self._state == ReadbackState.Ready

    def getData(self):
        return self._data

    def getBytes(self):
        return self._data.tobytes()

    def getWidth(self):
        return self._data.shape[1]

    def getHeight(self):
        return self._data.shape[0]

    def getDimensions(self):
        return (self.getWidth(), self.getHeight(), 1)

    def getFormat(self):
        return "RGBA8_SRGB"

    def getTexelSizeInBytes(self):
        return 4

    def getRowPitchInBytes(self):
        return 4 * self.getWidth()

    def getSizeInBytes(self):
        return self._data.nbytes

    def release(self):
        self._data = None
        self._state = ReadbackState.Freed


# ---------------------------------------------------------------- scene and renderer
class Scene:
    """py_yapt Scene with the scene_wrapper.py helpers (both spellings)"""

    def __init__(self, renderer):
        self._renderer = renderer
        self._objects = []
        self._camera = Camera(self)
        self._camera._transform._on_change = renderer._touch_frame

    def _touch(self):
        self._renderer._touch()

    def _touch_transform(self):
        self._renderer._touch_transform()

    def createRenderObject(self, name):
        o = RenderObject(self, name)
        self._objects.append(o)
        self._touch()
        return o

    def createSceneObject(self, name):
        o = SceneObject(self, name)
        self._objects.append(o)
        return o

    def createLight(self, name, light_type):
        o = Light(self, name, light_type)
        self._objects.append(o)
        self._touch()
        return o

    create_render_object = createRenderObject
    create_scene_object = createSceneObject
    create_light = createLight

    def destroy_object(self, obj):
        self._objects.remove(obj)
        self._touch()

    def get_all_objects_list(self):
        return list(self._objects)

    def getMainCamera(self):
        return self._camera

    get_main_camera = getMainCamera

    def update(self, dt=0.0):
        pass

    def arrays(self):
        """the committed scene: every render object's submeshes in world space, one instance each"""
        pos, nrm, tan, uv, idx = [], [], [], [], []
        inst_first, inst_count, inst_mat = [], [], []
        materials, mat_index, textures, tex_index = [], {}, [], {}
        any_uv = any(isinstance(o, RenderObject) and o._mesh is not None and o._mesh.uvs is not None for o in self._objects)
        default = None
        nverts = ntris = 0

        def texture_slot(t):
            if t.is_float:
                raise ValueError("material textures are RGBA8; %s holds float texels" % t.getName())
            if id(t) not in tex_index:
                tex_index[id(t)] = len(textures)
                textures.append((t.texels, t.srgb))
            return tex_index[id(t)]

        def material_slot(m):
            nonlocal default
            if m is None:
                if default is None:
                    default = len(materials)
                    materials.append(_scenes.blank_material())
                return default
            if id(m) not in mat_index:
                mat_index[id(m)] = len(materials)
                materials.append(m.record(texture_slot))
            return mat_index[id(m)]

        for o in self._objects:
            if not isinstance(o, RenderObject) or o._mesh is None:
                continue
            mesh = o._mesh
            M = o._transform.getMatrixWS()
            A = M[:3, :3]
            det = np.linalg.det(A)
            P = mesh.positions @ A.T + M[:3, 3]
            if mesh.normals is not None:
                N = mesh.normals @ np.linalg.inv(A)
            else:
                N = None
            for k in range(mesh.getSubmeshCount()):
                tri = mesh.submesh_triangles(k)
                if len(tri) == 0:
                    continue
                if det < 0:
                    tri = tri[:, [0, 2, 1]]
                used, local = np.unique(tri.reshape(-1), return_inverse=True)
                p = P[used]
                if N is not None:
                    n = N[used]
                    n = n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
                else:
                    n = _area_normals(p, local.reshape(-1, 3))
                if mesh.tangents is not None:
                    t3 = mesh.tangents[used, :3] @ A.T
                    t3 = t3 / np.maximum(np.linalg.norm(t3, axis=1, keepdims=True), 1e-30)
                    w = mesh.tangents[used, 3:4] * (1.0 if det >= 0 else -1.0)
                    t = np.concatenate([t3, w], 1)
                else:
                    t = _scenes.arbitrary_tangents(n.astype(np.float32)).astype(np.float64)
                pos.append(p)
                nrm.append(n)
                tan.append(t)
                if any_uv:
                    uv.append(mesh.uvs[used] if mesh.uvs is not None else np.zeros((len(used), 2)))
                idx.append(local.reshape(-1, 3) + nverts)
                inst_first.append(ntris)
                inst_count.append(len(tri))
                inst_mat.append(material_slot(o._materials.get(k)))
                nverts += len(used)
                ntris += len(tri)
        if not pos:
            raise ValueError("the scene has no render object with a mesh")
        lights = [o._world_light() for o in self._objects if isinstance(o, Light)]
        sky = self._renderer.getRendererVariable("World.Skycube").getTexture()
        return _lib.SceneArrays(np.concatenate(pos), np.concatenate(nrm), np.concatenate(idx), inst_first, inst_count,
                                inst_mat, materials, tangents=np.concatenate(tan).astype(np.float32), lights=lights,
                                uvs=np.concatenate(uv) if any_uv else None, textures=textures,
                                environment=None if sky is None else sky.environment())


def _area_normals(P, I):
    n = np.zeros_like(P)
    fn = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    for k in range(3):
        np.add.at(n, I[:, k], fn)
    l = np.linalg.norm(n, axis=1, keepdims=True)
    return np.where(l > 0, n / np.where(l > 0, l, 1), np.array([0.0, 1.0, 0.0]))


PIPELINES = {"BPT": _lib.YPT_INTEGRATOR_BPT, "BPT_WITH_NEE": _lib.YPT_INTEGRATOR_NEE, "BDPT": _lib.YPT_INTEGRATOR_BDPT}


class Renderer:
    """PyRenderer.cpp: materials, meshes, textures, renderer variables and the FinalColor readback,
    with the integrator of the selected RenderPipeline doing the rendering"""

    def __init__(self, device=0):
        self._device = device
        self._vars = _renderer_variables(self)
        self._scene = None
        self._scene_gen = 0
        self._frame_gen = 0
        self._xform_gen = 0
        self._committed = None     # (pipeline, scene, frame generations, denoise features, transform generation)
        self._integrators = {}
        self._arrays = None
        self._samples = 0
        self._warned_auto = False

    # generation counters: a scene change re-uploads the scene, a frame change re-sets the frame, a
    # transform change of scene objects refits the BVH to the moved triangles (no rebuild), and each
    # restarts the accumulation as the reference's resetRenderOutput does
    def _touch(self):
        self._scene_gen += 1

    def _touch_transform(self):
        self._xform_gen += 1

    def _touch_frame(self):
        self._frame_gen += 1

    def init(self, *args, **kwargs):
        return self

    def shutdown(self):
        for it in self._integrators.values():
            it.close()
        self._integrators = {}
        self._committed = None

    def createMaterial(self, name):
        return Material(self, name)

    def createMesh(self, name, positions, indices, normals=None, uvs=None, tangents=None, submeshes=None):
        return Mesh(name, positions, indices, normals, uvs, tangents, submeshes)

    def createTexture(self, name, texels, srgb=True, cube=False):
        return Texture(name, texels, srgb, cube)

    def getRendererVariable(self, name):
        return self._vars.get(name)

    def getAllRendererVariableNames(self):
        return list(self._vars)

    def resetRenderOutput(self):
        self._committed = None

    def _var(self, name):
        return self._vars[name]

    def frame_desc(self):
        v = self._var
        w, h = v("Generic.RenderResolution").getAsIntArray()
        cam = self._scene.getMainCamera()._camera(v("Camera.ApertureRadius").getAsFloatArray()[0],
                                                  v("Camera.FocusDistance").getAsFloatArray()[0])
        return _camera.make_frame(
            cam, w, h, max_depth=v("Integrator.MaxDepth").getAsIntArray()[0],
            env_intensity_scale=v("World.EnvIntensityScale").getAsFloatArray()[0],
            regularize_after=v("Firefly.RegularizeAfterVertices").getAsIntArray()[0],
            firefly_clamp_scale=v("Firefly.FireflyClampScale").getAsFloatArray()[0],
            firefly_clamp_relax_power=v("Firefly.FireflyClampRelaxPower").getAsFloatArray()[0],
            kill_nan_samples=v("Firefly.KillNaNSamples").getSelectedOption(),
            light_path_vertices=v("Integrator.LightPathVertices").getAsIntArray()[0],
            bdpt_light_nodes=[min(x, 0xFFFFFFFF) for x in v("_DEBUG.BDPT.EffectiveLightPathNodes").getAsIntArray()],
            bdpt_camera_nodes=[min(x, 0xFFFFFFFF) for x in v("_DEBUG.BDPT.EffectiveCameraPathNodes").getAsIntArray()])

    def post_params(self):
        v = self._var
        mode = v("Denoise.Mode").selected()
        if mode not in ("OFF", "ON"):
            raise NotImplementedError("Denoise.Mode %s is a debug view that is not built" % mode)
        if v("Tonemap.UseAutoExposure").selected() == "ON" and not self._warned_auto:
            print("py_yapt: auto exposure is not built, using Tonemap.ManualExposure", file=sys.stderr)
            self._warned_auto = True
        exposure = v("Tonemap.ManualExposure").getAsFloatArray()[0]
        return _lib.PostParams.defaults(
            denoise=1 if mode == "ON" else 0, prefilter_features=v("Denoise.FeaturePrefilter").getSelectedOption(),
            depth_sigma_scale=v("Denoise.DepthSigmaScale").getAsFloatArray()[0],
            normal_sigma=v("Denoise.NormalSigma").getAsFloatArray()[0],
            material_sigma=v("Denoise.MaterialDiffSigma").getAsFloatArray()[0],
            color_sigma=v("Denoise.ColorDiffSigma").getAsFloatArray()[0],
            feature_variance_scale=v("Denoise.FeatureVarianceScale").getAsFloatArray()[0],
            tonemap=v("Tonemap.Enable").getSelectedOption(), exposure=exposure)

    def pipeline(self):
        name = self._var("RenderPipeline").selected()
        if name not in PIPELINES:
            raise NotImplementedError("RenderPipeline %s is a debug view that is not built" % name)
        return PIPELINES[name]

    def _integrator(self):
        """the integrator of the selected pipeline with the current scene and frame"""
        if self._scene is None:
            raise RuntimeError("no scene: create one with Scene(renderer) or Application")
        kind = self.pipeline()
        features = self._var("Denoise.Mode").selected() == "ON"
        state = (kind, self._scene_gen, self._frame_gen, features, self._xform_gen)
        it = self._integrators.get(kind)
        if it is None:
            it = self._integrators[kind] = _lib.native().Integrator(kind, self._device)
            self._committed = None
        if self._committed != state:
            if self._committed is None or self._committed[1] != self._scene_gen or self._committed[0] != kind:
                self._arrays = self._scene.arrays()
                it.set_scene(self._arrays)
            elif self._committed[4] != self._xform_gen:
                self._arrays = self._scene.arrays()
                try:
                    it.refit_scene(self._arrays)   # moved scene objects: the BVH is refitted, not rebuilt
                except RuntimeError:
                    it.set_scene(self._arrays)     # a mirroring change reversed windings: other triangles
            it.set_option("first_hit_features", 1 if features else 0)
            it.set_frame(self.frame_desc())
            self._committed = state
            self._samples = 0
        return it

    def render(self, passes=1):
        """accumulates `passes` further sample indices into the film"""
        it = self._integrator()
        it.render(self._samples, int(passes))
        self._samples += int(passes)
        return self._samples

    def sample_count(self):
        return self._samples

    def readback(self, target=ReadbackTarget.FinalColor):
        if ReadbackTarget(int(target)) != ReadbackTarget.FinalColor:
            raise ValueError("only the FinalColor target exists")
        it = self._integrator()
        return Readback(it.final_color(self.post_params()))

    def film(self):
        return self._integrator().film()


class Resources:
    """pythonScripts/yapt/resources.py: meshes from glTF files and 2D textures from image files"""

    def __init__(self, renderer):
        self._renderer = renderer
        self._min_alpha = {}

    def load_meshes_from_path(self, mesh_path, verbose=False):
        """one entry per glTF mesh: ([Mesh with one submesh per primitive], {primitive: (0, submesh)})"""
        from . import gltf
        out = []
        base = os.path.splitext(os.path.basename(mesh_path))[0]
        for mi, prims in enumerate(gltf.mesh_primitives(mesh_path)):
            if prims:
                out.append(([_mesh_from_primitives("%s_%d" % (base, mi), prims)], {k: (0, k) for k in range(len(prims))}))
        return out

    def load_texture_2D(self, tex_path, fmt="R8G8B8A8_SRGB", verbose=False):
        """an image file as a texture in `fmt` (R8G8B8A8_SRGB or R8G8B8A8_UNORM); a KTX file keeps the
        format it was written in. An _SFLOAT format, or an .hdr / .exr / .npy file, gives a linear float
        texture, the kind World.Skycube takes for a long-lat environment."""
        from . import ktx
        low = str(tex_path).lower()
        if "SFLOAT" in str(fmt).upper() or low.endswith((".hdr", ".exr", ".npy")):
            return self._load_float(tex_path)
        if low.endswith((".ktx", ".ktx2")):
            rgba, srgb = ktx.read(tex_path)
        else:
            from PIL import Image
            rgba = np.asarray(Image.open(tex_path).convert("RGBA"), np.uint8)
            srgb = str(fmt).upper().endswith("SRGB")
        tex = Texture(os.path.basename(str(tex_path)), rgba, srgb=srgb)
        self._min_alpha[str(tex_path)] = tex.min_alpha()
        return tex

    def _load_float(self, tex_path):
        from . import hdr_images, ktx
        name = os.path.basename(str(tex_path))
        if str(tex_path).lower().endswith((".ktx", ".ktx2")):
            texels, cube = ktx.read_float(tex_path)
            return Texture(name, texels, cube=cube)
        if str(tex_path).lower().endswith((".png", ".jpg", ".jpeg")):
            from PIL import Image
            c = np.asarray(Image.open(tex_path).convert("RGBA"), np.float64) / 255.0
            c[..., :3] = np.where(c[..., :3] <= 0.04045, c[..., :3] / 12.92, ((c[..., :3] + 0.055) / 1.055) ** 2.4)
            return Texture(name, c.astype(np.float32))
        return Texture(name, hdr_images.read(tex_path))

    def load_texture_cube(self, tex_path, fmt="R16G16B16A16_SFLOAT", faces=None, verbose=False):
        """a cube map for World.Skycube: a KTX2 file with six faces, or six face images given as `faces` in
        the order +X, -X, +Y, -Y, +Z, -Z (texture_loader_ktx.py:150-175); tex_path names the texture"""
        name = os.path.basename(str(tex_path))
        if faces is None:
            from . import ktx
            texels, cube = ktx.read_float(tex_path)
            if not cube:
                raise ValueError("%s holds no cube map (faceCount 1)" % tex_path)
            return Texture(name, texels, cube=True)
        if len(faces) != 6:
            raise ValueError("a cube map takes 6 face images, got %d" % len(faces))
        imgs = [self._load_float(f).texels for f in faces]
        if any(i.shape != imgs[0].shape or i.shape[0] != i.shape[1] for i in imgs):
            raise ValueError("cube-map faces must be square and of one size")
        return Texture(name, np.stack(imgs), cube=True)

    def get_texture_min_alpha(self, tex_path):
        return self._min_alpha.get(str(tex_path))


class Application:
    """the `yapt_instance` a scene script receives (app.py): renderer, scene and resources"""

    def __init__(self, width=None, height=None, device=0):
        self._renderer = Renderer(device)
        self._scene = Scene(self._renderer)
        self._renderer._scene = self._scene
        self._resources = Resources(self._renderer)
        if width is not None:
            self._renderer.getRendererVariable("Generic.RenderResolution").setFromIntArray([width, height])

    def get_renderer(self):
        return self._renderer

    def get_scene(self):
        return self._scene

    def get_resources(self):
        return self._resources

    def run_script(self, path, replacements=None):
        """execs a scene script with `yapt_instance` bound, as the app's Run Script menu does"""
        with open(path, encoding="utf-8") as f:
            source = f.read()
        for old, new in (replacements or {}).items():
            source = source.replace(old, new)
        context = {"yapt_instance": self}
        exec(compile(source, path, "exec"), context, context)
        return context

    def shutdown(self):
        self._renderer.shutdown()


def _decompose(m):
    """translation, rotation (x, y, z, w) and scale of an affine matrix without shear"""
    m = np.asarray(m, np.float64)
    t = m[:3, 3].copy()
    a = m[:3, :3]
    s = np.linalg.norm(a, axis=0)
    if np.linalg.det(a) < 0:
        s[0] = -s[0]
    r = a / np.where(s == 0, 1.0, s)
    return t, matrix_to_quat(r), s


class SceneLoader:
    """pythonScripts/yapt/scene_loader.py over this build's glTF parser: a node tree of scene objects
    with their local transforms (a `matrix` node is decomposed into translation, rotation and
    scale, where the reference raises), one render object per mesh node with a material per
    primitive, KHR_lights_punctual lights as children of their nodes, and materials and textures
    mapped as in yapt_amd/gltf.py."""

    def __init__(self, resources, scene, renderer):
        self._resources, self._scene, self._renderer = resources, scene, renderer
        self._materials, self._doc, self._default = [], None, None

    def load_scene_gltf(self, scene_path, verbose=False):
        from . import gltf
        doc, binchunk = gltf._read_container(scene_path)
        base = os.path.dirname(os.path.abspath(scene_path))
        bufs = gltf._buffers(doc, base, binchunk)
        self._doc = doc
        images = gltf._Images(doc, bufs, base)
        self._materials = []
        for k, mdef in enumerate(doc.get("materials", [])):
            rec, slots = gltf._material(doc, mdef, images)
            mat = self._renderer.createMaterial(mdef.get("name", "material_%d" % k))
            mat._p = _lib.Material.from_buffer_copy(rec)
            self._materials.append((mat, slots))
        textures = [self._renderer.createTexture("image_%d" % i, rgba, srgb) for i, (rgba, srgb) in enumerate(images.textures)]
        for mat, slots in self._materials:
            for slot, index in slots.items():
                if index >= 0:
                    mat._bind(slot, textures[index], getattr(mat._p, slot + "_tex_scale"))
        self._meshes = []
        name = os.path.splitext(os.path.basename(scene_path))[0]
        for mi, prims in enumerate(gltf.mesh_primitives(scene_path)):
            self._meshes.append(_mesh_from_primitives("%s_%d" % (name, mi), prims) if prims else None)
        self._prims = [[p for p in m.get("primitives", []) if p.get("mode", 4) == 4] for m in doc.get("meshes", [])]
        si = doc.get("scene", 0)
        roots = doc["scenes"][si]["nodes"] if doc.get("scenes") else list(range(len(doc.get("nodes", []))))
        return [self._node_tree(r) for r in roots]

    def get_gltf(self):
        return self._doc

    def get_material(self, gltf_material_name):
        for mat, _ in self._materials:
            if mat.getName() == gltf_material_name:
                return mat
        return None

    def _default_material(self):
        if self._default is None:
            self._default = self._renderer.createMaterial("default_mat")
            self._default.setFromMaterialPreset(MaterialPreset.PLASTIC)
            self._default.setSpecularAmount(0.0)
        return self._default

    def _node_tree(self, index):
        node = self._doc["nodes"][index]
        obj = self._node_object(node, index)
        self._node_light(node, obj)
        for c in node.get("children", []):
            child = self._node_tree(c)
            child.getTransform().setParent(obj.getTransform())
        return obj

    def _node_object(self, node, index):
        name = node.get("name", "node_%d" % index)
        obj = self._scene.create_scene_object(name)
        tr = obj.getTransform()
        if "matrix" in node:
            from . import gltf
            t, q, s = _decompose(gltf.node_matrix(node))
            tr.setTranslation(t)
            tr.setOrientation(q)
            tr.setScale(s)
        else:
            if "translation" in node:
                tr.setTranslation(vec3(node["translation"]))
            if "scale" in node:
                tr.setScale(vec3(node["scale"]))
            if "rotation" in node:
                tr.setOrientation(vec4(node["rotation"]))
        mesh = self._meshes[node["mesh"]] if "mesh" in node and node["mesh"] < len(self._meshes) else None
        if mesh is not None:
            ro = self._scene.create_render_object("%s_0" % name)
            ro.setMesh(mesh)
            ro.getTransform().setParent(tr)
            for k, prim in enumerate(self._prims[node["mesh"]]):
                m = prim.get("material")
                ro.setMaterial(self._materials[m][0] if m is not None else self._default_material(), k)
        return obj

    def _node_light(self, node, obj):
        from . import gltf
        ref = (node.get("extensions") or {}).get("KHR_lights_punctual")
        defs = (self._doc.get("extensions") or {}).get("KHR_lights_punctual", {}).get("lights", [])
        if ref is None or ref.get("light") is None or ref["light"] >= len(defs):
            return None
        d = defs[ref["light"]]
        t = {"point": LightType.POINT, "spot": LightType.SPOT, "directional": LightType.DIRECTIONAL}.get(d.get("type"))
        if t is None:
            return None
        light = self._scene.create_light(d.get("name") or "%s_light" % node.get("name", "node"), t)
        light.setColor(vec3(d.get("color", [1.0, 1.0, 1.0])))
        light.setIntensity(d.get("intensity", 1.0) / gltf.GLTF_LUMENS_PER_WATT)
        if t == LightType.SPOT:
            spot = d.get("spot", {})
            light.setSpotConeAngles(spot.get("innerConeAngle", 0.0), spot.get("outerConeAngle", math.pi / 4.0))
        light.getTransform().setParent(obj.getTransform())
        return light


def _mesh_from_primitives(name, prims):
    """one Mesh with a submesh per primitive"""
    P, N, T, UV, I, subs = [], [], [], [], [], []
    nv = ni = 0
    has_n = all(p["normals"] is not None for p in prims)
    has_t = all(p["tangents"] is not None for p in prims)
    has_uv = any(p["uvs"] is not None for p in prims)
    for p in prims:
        P.append(p["positions"])
        if has_n:
            N.append(p["normals"])
        if has_t:
            T.append(p["tangents"])
        if has_uv:
            UV.append(p["uvs"] if p["uvs"] is not None else np.zeros((len(p["positions"]), 2), np.float32))
        I.append(p["indices"])
        subs.append((ni, len(p["indices"]), nv))
        nv += len(p["positions"])
        ni += len(p["indices"])
    return Mesh(name, np.concatenate(P), np.concatenate(I), np.concatenate(N) if has_n else None,
                np.concatenate(UV) if has_uv else None, np.concatenate(T) if has_t else None, subs)
