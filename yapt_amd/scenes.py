"""Synthetic scenes of the benchmark configurations, as world-space triangle soups.

* ``cornell_box``: 5 wall quads + 2 axis-aligned boxes + 1 quad light (36 triangles), diffuse
  principled materials (PLASTIC preset, specularAmount 0, roughness 1), RGB-white emitter whose
  spectrum is D65-shaped (spectralDistribution.hlsl:140-180).
* ``random_shapes``: seeded spheres / boxes / tori (1:1:1) tessellated to ~1M triangles inside a
  20-unit cube, roughness ~ U(0.05, 1), metalness ~ Bernoulli(0.3), 10% glass (GLASS preset with
  the BK7 Cauchy coefficients and dispersion), lit by quad lights above the cube. The reference's own
  scene family (pythonScripts/yapt/random_shapes.py) needs glTF assets that are not available; this
  generator follows its material families and presets instead.
Material presets follow MaterialProxy.cpp:15-52 and the defaults MaterialInternal.h:41-75.
"""
import math
import os

import numpy as np

from . import _lib
from .camera import Camera


# ------------------------------------------------------------------ materials
def blank_material():
    """MaterialParameters() defaults (MaterialInternal.h:41-75)."""
    m = _lib.Material()
    m.albedo[:] = (1.0, 1.0, 1.0)
    m.transparency = 0.0
    m.specular[:] = (1.0, 1.0, 1.0)
    m.metalness = 1.0
    m.absorption[:] = (0.0, 0.0, 0.0)
    m.dielectric_ior = 1.5
    m.emissive[:] = (0.0, 0.0, 0.0)
    m.roughness = 0.1
    m.anisotropy = 0.0
    m.anisotropy_rotation = 0.0
    m.material_mask = 0
    m.specular_amount = 1.0
    m.clear_coat_amount = 0.0
    m.clear_coat_ior = 1.5
    m.clear_coat_roughness = 0.1
    m.sheen_tint[:] = (1.0, 1.0, 1.0)
    m.sheen_roughness = 0.5
    m.sheen_amount = 0.0
    m.thin_film_thickness_nm = 0.0
    m.emission_focus = 0.0
    m.alpha_cutoff = 0.0
    m.cauchy_coeffs[:] = (1.5046, 0.00420)
    m.subsurface_radius[:] = (1.0, 0.2, 0.1)
    m.subsurface_amount = 0.0
    m.subsurface_radius_scale = 1.0
    m.albedo_tex = m.normal_tex = m.orm_tex = m.emissive_tex = -1   # TEX_UNBOUND_INDEX
    for sc in (m.albedo_tex_scale, m.normal_tex_scale, m.orm_tex_scale, m.emissive_tex_scale):
        sc[:] = (1.0, 1.0)
    m.blackbody_kelvin = 0.0                        # extensions: off
    m.sellmeier_b[:] = (0.0, 0.0, 0.0)
    m.sellmeier_c[:] = (0.0, 0.0, 0.0)
    return m


BK7_SELLMEIER = ((1.03961212, 0.231792344, 1.01046945), (0.00600069867, 0.0200179144, 103.560653))


def _pow22(c):
    return tuple(float(v) for v in np.power(np.asarray(c, np.float32), np.float32(2.2)))


PRESETS = {
    # albedo, transparency, specular, metalness, ior, roughness  (MaterialProxy.cpp:22-48)
    "METAL_GOLD": (_pow22((0.9451, 0.7294, 0.37255)), 0.0, _pow22((1.0, 0.97255, 0.73333)), 1.0, 1.5, 0.0),
    "METAL_SILVER": (_pow22((0.9607, 0.9490, 0.9176)), 0.0, _pow22((1.0, 1.0, 1.0)), 1.0, 1.5, 0.0),
    "METAL_ALUMINIUM": (_pow22((0.9137, 0.9137, 0.9137)), 0.0, _pow22((0.9686, 0.98039, 0.9882)), 1.0, 1.5, 0.0),
    "METAL_BRASS": (_pow22((0.903, 0.744, 0.473)), 0.0, _pow22((0.975, 0.958, 0.918)), 1.0, 1.5, 0.0),
    "METAL_COPPER": (_pow22((0.92549, 0.68627, 0.50196)), 0.0, _pow22((0.9960, 0.945098, 0.8196)), 1.0, 1.5, 0.0),
    "GLASS": ((1.0, 1.0, 1.0), 1.0, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0),
    "PLASTIC": ((1.0, 1.0, 1.0), 0.0, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0),
}


def preset_material(name):
    m = blank_material()
    if name == "BLANK":
        return m
    albedo, transparency, specular, metalness, ior, roughness = PRESETS[name]
    m.albedo[:] = albedo
    m.transparency = transparency
    m.specular[:] = specular
    m.metalness = metalness
    m.dielectric_ior = ior
    m.roughness = roughness
    return m


# ------------------------------------------------------------------ geometry helpers
def arbitrary_tangents(normals):
    """Unit tangents orthogonal to each normal, handedness 1 (the role of
    MeshUtility.calculate_arbitrary_tangents, pythonScripts/yapt/mesh_utils.py:69-85, which returns a
    zero vector for normals such as (-1, 0, 0); here the axis least aligned with n is crossed instead)."""
    n = np.asarray(normals, np.float64).reshape(-1, 3)
    a = np.abs(n)
    axis = np.zeros_like(n)
    k = np.argmin(a, axis=1)
    axis[np.arange(len(n)), k] = 1.0
    t = np.cross(axis, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return np.concatenate([t, np.ones((len(n), 1))], axis=1).astype(np.float32)


class MeshBuilder:
    def __init__(self):
        self.pos, self.nrm, self.idx = [], [], []
        self.inst_first, self.inst_count, self.inst_mat = [], [], []
        self.materials = []
        self.nverts = 0
        self.ntris = 0

    def material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add(self, positions, normals, indices, material):
        positions = np.asarray(positions, np.float32).reshape(-1, 3)
        normals = np.asarray(normals, np.float32).reshape(-1, 3)
        indices = np.asarray(indices, np.uint32).reshape(-1, 3)
        self.pos.append(positions)
        self.nrm.append(normals)
        self.idx.append(indices + self.nverts)
        self.inst_first.append(self.ntris)
        self.inst_count.append(len(indices))
        self.inst_mat.append(material)
        self.nverts += len(positions)
        self.ntris += len(indices)

    def add_many(self, positions, normals, indices, materials):
        """positions/normals: (n_inst, V, 3) sharing one index template (T, 3)."""
        n, V = positions.shape[0], positions.shape[1]
        T = len(indices)
        off = (self.nverts + np.arange(n, dtype=np.uint64) * V)[:, None, None]
        self.pos.append(positions.reshape(-1, 3).astype(np.float32))
        self.nrm.append(normals.reshape(-1, 3).astype(np.float32))
        self.idx.append((indices[None, :, :].astype(np.uint64) + off).reshape(-1, 3).astype(np.uint32))
        self.inst_first.extend((self.ntris + np.arange(n) * T).tolist())
        self.inst_count.extend([T] * n)
        self.inst_mat.extend([int(m) for m in materials])
        self.nverts += n * V
        self.ntris += n * T

    def quad(self, corner, u, v, material):
        """Quad corner, corner+u, corner+u+v, corner+v; front side along cross(u, v)."""
        c = np.asarray(corner, np.float32)
        u = np.asarray(u, np.float32)
        v = np.asarray(v, np.float32)
        n = np.cross(u, v)
        n = n / np.linalg.norm(n)
        p = np.stack([c, c + u, c + u + v, c + v])
        self.add(p, np.tile(n, (4, 1)), [[0, 1, 2], [0, 2, 3]], material)

    def arrays(self, lights=(), uvs=None, textures=(), environment=None):
        nrm = np.concatenate(self.nrm)
        return _lib.SceneArrays(np.concatenate(self.pos), nrm, np.concatenate(self.idx),
                                self.inst_first, self.inst_count, self.inst_mat, self.materials,
                                tangents=arbitrary_tangents(nrm), lights=lights, uvs=uvs, textures=textures,
                                environment=environment)


def box_template():
    """Unit cube [-1, 1]^3, 24 vertices (flat faces), outward winding."""
    P, N, I = [], [], []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3, np.float32)
            n[axis] = sgn
            a, b = (axis + 1) % 3, (axis + 2) % 3
            u = np.zeros(3, np.float32)
            v = np.zeros(3, np.float32)
            u[a], v[b] = 2.0, 2.0
            if np.dot(np.cross(u, v), n) < 0:
                u, v = v, u
            c = n - u / 2 - v / 2
            base = len(P)
            P += [c, c + u, c + u + v, c + v]
            N += [n] * 4
            I += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    return np.array(P, np.float32), np.array(N, np.float32), np.array(I, np.uint32)


def sphere_template(stacks=20, slices=20):
    P, N = [], []
    for i in range(stacks + 1):
        th = math.pi * i / stacks
        for j in range(slices):
            ph = 2 * math.pi * j / slices
            n = (math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph))
            P.append(n)
            N.append(n)
    I = []
    for i in range(stacks):
        for j in range(slices):
            a = i * slices + j
            b = i * slices + (j + 1) % slices
            c = (i + 1) * slices + j
            d = (i + 1) * slices + (j + 1) % slices
            if i != 0:
                I.append([a, b, c])
            if i != stacks - 1:
                I.append([b, d, c])
    P, N, I = np.array(P, np.float32), np.array(N, np.float32), np.array(I, np.uint32)
    return _fix_winding(P, N, I)


def torus_template(major=32, minor=12, R=1.0, r=0.35):
    P, N = [], []
    for i in range(major):
        u = 2 * math.pi * i / major
        for j in range(minor):
            v = 2 * math.pi * j / minor
            cx, cz = R * math.cos(u), R * math.sin(u)
            n = (math.cos(v) * math.cos(u), math.sin(v), math.cos(v) * math.sin(u))
            P.append((cx + r * n[0], r * n[1], cz + r * n[2]))
            N.append(n)
    I = []
    for i in range(major):
        for j in range(minor):
            a = i * minor + j
            b = ((i + 1) % major) * minor + j
            c = i * minor + (j + 1) % minor
            d = ((i + 1) % major) * minor + (j + 1) % minor
            I.append([a, b, d])
            I.append([a, d, c])
    P, N, I = np.array(P, np.float32), np.array(N, np.float32), np.array(I, np.uint32)
    return _fix_winding(P, N, I)


def _fix_winding(P, N, I):
    """Orient every triangle so that cross(p1 - p0, p2 - p0) agrees with the vertex normals."""
    e = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    flip = np.einsum("ij,ij->i", e, N[I].sum(axis=1)) < 0
    I = I.copy()
    I[flip] = I[flip][:, [0, 2, 1]]
    return P, N, I


def random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    return R


def hsv_to_rgb(h, s, v):
    i = int(h * 6.0) % 6
    f = h * 6.0 - int(h * 6.0)
    p, q, t = v * (1 - s), v * (1 - s * f), v * (1 - s * (1 - f))
    return [(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v), (v, p, q)][i]


# ------------------------------------------------------------------ scenes
def cornell_box():
    """36-triangle Cornell-style box (configuration: 256x256, 16 spp, depth 8)."""
    b = MeshBuilder()

    def diffuse(rgb):
        m = preset_material("PLASTIC")
        m.albedo[:] = rgb
        m.specular_amount = 0.0
        m.roughness = 1.0
        return b.material(m)

    white = diffuse((0.73, 0.73, 0.73))
    red = diffuse((0.63, 0.065, 0.05))
    green = diffuse((0.14, 0.45, 0.091))
    lm = preset_material("PLASTIC")
    lm.albedo[:] = (0.78, 0.78, 0.78)
    lm.specular_amount = 0.0
    lm.roughness = 1.0
    lm.emissive[:] = (15.0, 15.0, 15.0)
    light = b.material(lm)
    # room x in [-1, 1], y in [0, 2], z in [-1, 1]; fronts face the interior
    b.quad((-1, 0, 1), (2, 0, 0), (0, 0, -2), white)     # floor, +y
    b.quad((-1, 2, -1), (2, 0, 0), (0, 0, 2), white)     # ceiling, -y
    b.quad((-1, 0, -1), (2, 0, 0), (0, 2, 0), white)     # back, +z
    b.quad((-1, 0, 1), (0, 0, -2), (0, 2, 0), red)       # left, +x
    b.quad((1, 0, -1), (0, 0, 2), (0, 2, 0), green)      # right, -x
    P, N, I = box_template()
    for centre, half in (((-0.35, 0.6, -0.3), (0.3, 0.6, 0.3)), ((0.4, 0.3, 0.35), (0.3, 0.3, 0.3))):
        b.add(P * np.asarray(half, np.float32) + np.asarray(centre, np.float32), N, I, white)
    b.quad((-0.25, 1.98, -0.25), (0.5, 0, 0), (0, 0, 0.5), light)  # light, -y
    cam = Camera(eye=(0.0, 1.0, 3.9), target=(0.0, 1.0, 0.0), fovy=math.pi * 0.3)
    return b.arrays(), cam


def focused_emission(focus=8.0, two_sided=False):
    """The focused-emitter box of the reference's integrator comparison (Tests/IntegratorCompare/
    focused_emission_scene.py:51-77) from generated quads: a closed room of white diffuse PLASTIC
    (albedo 0.7, roughness 1, no specular) whose floor (y = 0), ceiling (y = 40) and walls (x, z = +-20)
    are 80x80 quads turned inwards, and one 12x12 emitter (emission 8, cos^n focus `focus`) 2 units
    below the ceiling facing the floor; a two-sided emitter hangs in the middle of the room instead.
    The reference's plane.glb is a 2x2 quad (load_material_showcase_scene.py:583) scaled by 40 for the
    room and 6 for the emitter. Camera at (0, 22, 18) pitched 25 degrees down, the default 54-degree
    field of view; environment scale 0, manual exposure 1."""
    b = MeshBuilder()
    room = preset_material("PLASTIC")
    room.albedo[:] = (0.7, 0.7, 0.7)
    room.roughness = 1.0
    room.specular_amount = 0.0
    rm = b.material(room)
    L, h, S = 40.0, 20.0, 80.0   # room height, half width, quad size (2 x scale 40)
    b.quad((-L, 0, L), (S, 0, 0), (0, 0, -S), rm)        # floor, +y
    b.quad((-L, L, -L), (S, 0, 0), (0, 0, S), rm)        # ceiling, -y
    b.quad((L, h - L, h), (-S, 0, 0), (0, S, 0), rm)     # wall z = +20, -z
    b.quad((h, h - L, -L), (0, 0, S), (0, S, 0), rm)     # wall x = +20, -x
    b.quad((-L, h - L, -h), (S, 0, 0), (0, S, 0), rm)    # wall z = -20, +z
    b.quad((-h, h - L, L), (0, 0, -S), (0, S, 0), rm)    # wall x = -20, +x
    light = preset_material("BLANK")
    light.emissive[:] = (8.0, 8.0, 8.0)
    light.emission_focus = float(focus)
    light.material_mask = 1 if two_sided else 0          # MaterialMask_TwoSided
    lm = b.material(light)
    y = h if two_sided else L - 2.0
    b.quad((-6.0, y, -6.0), (12.0, 0, 0), (0, 0, 12.0), lm)  # scale 6, facing -y
    pitch = math.radians(25.0)
    eye = (0.0, 22.0, h - 2.0)
    cam = Camera(eye=eye, target=(0.0, 22.0 - math.sin(pitch), h - 2.0 - math.cos(pitch)), fovy=math.pi * 0.3)
    return b.arrays(), cam


def longlat_directions(width, height):
    """unit directions of the texel centres of a long-lat environment image (H x W x 3): the inverse of
    getSkyBoxColor's mapping u = atan2(z, x) / 2pi + 1/2, v = acos(y) / pi (row 0 looks up)"""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    theta = (u - 0.5) * 2.0 * math.pi
    phi = v * math.pi
    st, ct = np.sin(phi)[:, None], np.cos(phi)[:, None]
    return np.stack(np.broadcast_arrays(st * np.cos(theta)[None, :], ct, st * np.sin(theta)[None, :]), axis=-1)


def cube_face_directions(size):
    """unit directions of the texel centres of the six faces +X, -X, +Y, -Y, +Z, -Z (6 x N x N x 3), with the
    D3D / Vulkan face orientation: for face coordinates sc = 2s - 1, tc = 2t - 1 (s along a row, t down the
    rows) the direction is +X (1, -tc, -sc), -X (-1, -tc, sc), +Y (sc, 1, tc), -Y (sc, -1, -tc),
    +Z (sc, -tc, 1), -Z (-sc, -tc, -1)"""
    c = (np.arange(size) + 0.5) / size * 2.0 - 1.0
    sc, tc = np.meshgrid(c, c)       # sc varies along a row, tc down the rows
    one = np.ones_like(sc)
    faces = [(one, -tc, -sc), (-one, -tc, sc), (sc, one, tc), (sc, -one, -tc), (sc, -tc, one), (-sc, -tc, -one)]
    d = np.stack([np.stack(f, axis=-1) for f in faces])
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


SUN_DIRECTION = np.array([0.4, 0.6, -0.5]) / np.linalg.norm([0.4, 0.6, -0.5])


def sky_radiance(d, sun=40.0):
    """A procedural HDR sky over unit directions d (..., 3), linear RGB: a blue zenith fading to a pale
    horizon above the horizon, a darker warm ground below it, and a small bright sun of peak `sun`
    towards SUN_DIRECTION. Stands in for the reference's HDR environment images, which are absent."""
    d = np.asarray(d, np.float64)
    y = d[..., 1:2]
    up = np.clip(y, 0.0, 1.0)
    sky = np.array([0.55, 0.7, 0.95]) * (1.0 - up) + np.array([0.15, 0.3, 0.8]) * up
    ground = np.array([0.25, 0.2, 0.15]) * (0.6 + 0.4 * np.clip(-y, 0.0, 1.0))
    c = np.clip(d @ SUN_DIRECTION, -1.0, 1.0)[..., None]
    glow = sun * np.exp((c - 1.0) * 600.0) * np.array([1.0, 0.9, 0.75])
    return np.where(y > 0.0, sky, ground) + glow


def longlat_environment(width=64, height=32, fn=sky_radiance):
    """(YPT_ENVIRONMENT_LONGLAT, H x W x 4 float32) with fn evaluated at the texel centres"""
    rgb = fn(longlat_directions(width, height))
    return _lib.YPT_ENVIRONMENT_LONGLAT, np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,))], -1).astype(np.float32)


def cube_environment(size=16, fn=sky_radiance):
    """(YPT_ENVIRONMENT_CUBE, 6 x N x N x 4 float32) with fn evaluated at the texel centres of each face"""
    rgb = fn(cube_face_directions(size))
    return _lib.YPT_ENVIRONMENT_CUBE, np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,))], -1).astype(np.float32)


def punctual_lights(light_type="mixed", sphere_res=(24, 24), environment=None):
    """The punctual-light test scene (Tests/IntegratorCompare/punctual_lights_scene.py:16-105) built
    from generated meshes: an 80x80 floor and back wall, diffuse, copper and plastic spheres, lit by
    punctual lights only (environment scale 0). light_type: point, spot, directional, mixed (all three
    plus an emissive sphere), mesh_control (a small emissive sphere where the point light is), env (no
    lights: the scene is lit by the environment at scale 0.5, returned as the third value). environment:
    the scene's World.Skycube texture ((type, texels), e.g. longlat_environment()), None for the
    texture-less sky.
    The reference's plane.glb (a 2x2 quad, load_material_showcase_scene.py:583) and Sphere.glb (unit
    radius) are absent: the planes here are quads of twice their render objects' scale, and the spheres
    have radius = scale. Returns (scene, camera, env_intensity_scale)."""
    b = MeshBuilder()
    diffuse = preset_material("PLASTIC")
    diffuse.albedo[:] = (0.7, 0.7, 0.7)
    diffuse.roughness = 1.0
    diffuse.specular_amount = 0.0
    d = b.material(diffuse)
    glossy = preset_material("METAL_COPPER")
    glossy.roughness = 0.35
    g = b.material(glossy)
    plastic = preset_material("PLASTIC")
    plastic.albedo[:] = (0.2, 0.4, 0.8)
    plastic.roughness = 0.25
    pl = b.material(plastic)
    b.quad((-40, 0, 40), (80, 0, 0), (0, 0, -80), d)      # floor (scale 40), +y
    b.quad((-40, -20, -20), (80, 0, 0), (0, 80, 0), d)    # back wall at z = -20 (scale 40, centre y = 20), +z
    P, N, I = sphere_template(*sphere_res)

    def sphere(centre, radius, mat):
        b.add(P * np.float32(radius) + np.asarray(centre, np.float32), N, I, mat)

    sphere((-6, 3, 0), 3, d)
    sphere((2, 3, -4), 3, g)
    sphere((7, 2, 3), 2, pl)

    def light(t, position, target, color, intensity):
        l = _lib.Light(t)
        l.lookAt(position, target)
        l.setColor(color)
        l.setIntensity(intensity)
        return l

    lights = []
    if light_type in ("point", "mixed"):
        lights.append(light(_lib.Light.POINT, (-2, 12, 6), (0, 0, 0), (1.0, 0.9, 0.8), 150.0))
    if light_type in ("spot", "mixed"):
        spot = light(_lib.Light.SPOT, (10, 16, 10), (2, 0, -2), (0.8, 0.9, 1.0), 600.0)
        spot.setSpotConeAngles(0.3, 0.5)
        lights.append(spot)
    if light_type in ("directional", "mixed"):
        lights.append(light(_lib.Light.DIRECTIONAL, (-6, 12, 8), (0, 0, 0), (1.0, 1.0, 1.0), 1.0))
    if light_type == "mesh_control":
        m = blank_material()
        m.emissive[:] = (150.0, 135.0, 120.0)
        sphere((-2, 12, 6), 0.5, b.material(m))
    if light_type == "mixed":
        m = blank_material()
        m.emissive[:] = (6.0, 3.0, 1.0)
        sphere((-10, 6, -10), 1.5, b.material(m))
    ang = math.radians(25.0)
    eye = (0.0, 14.0, 26.0)
    cam = Camera(eye=eye, target=(0.0, 14.0 - math.sin(ang), 26.0 - math.cos(ang)), fovy=math.pi * 0.3)
    return b.arrays(lights, environment=environment), cam, (0.5 if light_type == "env" else 0.0)


def blackbody_rgb(kelvin):
    """Linear sRGB of a blackbody (Planck's law over 360-830 nm against the CIE 1931 observer), scaled
    so that its largest component is 1."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "cie_1931_2deg_d65.npz"))
    lam = d["lambda"].astype(np.float64) * 1e-9
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    planck = 2 * h * c * c / lam ** 5 / np.expm1(h * c / (lam * k * kelvin))
    xyz = np.array([np.sum(planck * d[a]) for a in ("x", "y", "z")])
    rgb = np.maximum(d["srgb_xyz_to_rgb"] @ xyz, 0.0)
    return rgb / rgb.max()


def random_shapes(n_shapes=2000, seed=2, cube=20.0, n_lights=4, light_emission=80.0, sphere_res=(20, 20),
                  torus_res=(32, 12), emissive_triangles=0, triangle_emission=200.0, spectral_extensions=False):
    """~1M-triangle random-shapes scene (n_shapes=2000, seed 2). emissive_triangles > 0 adds that many
    seeded two-sided emissive triangles (edge ~1.5) coloured as blackbodies of 2700-9000 K.
    spectral_extensions: the glass disperses by BK7's Sellmeier coefficients instead of Cauchy's, the
    emissive triangles emit Planck spectra of their temperatures and the area lights one of 6500 K
    (the benchmark text's materials; the reference's own are the default)."""
    rng = np.random.default_rng(seed)
    b = MeshBuilder()
    templates = [sphere_template(*sphere_res), box_template(), torus_template(*torus_res)]
    kinds = np.arange(n_shapes) % 3
    centres = rng.uniform(-cube / 2, cube / 2, size=(n_shapes, 3)).astype(np.float32)
    scales = rng.uniform(0.4, 1.1, size=n_shapes).astype(np.float32)
    rots = random_rotations(rng, n_shapes).astype(np.float32)
    mats = []
    for _ in range(n_shapes):
        u = rng.random()
        if u < 0.1:
            m = preset_material("GLASS")
            m.material_mask = _lib.YPT_MATERIAL_DISPERSION
            m.roughness = float(rng.uniform(0.0, 0.3))
            if spectral_extensions:
                m.sellmeier_b[:], m.sellmeier_c[:] = BK7_SELLMEIER
        else:
            metal = rng.random() < 0.3
            m = preset_material("METAL_SILVER" if metal else "PLASTIC")
            rgb = hsv_to_rgb(rng.random(), rng.uniform(0.3, 0.9), rng.uniform(0.4, 0.95))
            m.albedo[:] = tuple(round(float(c) * 63.0) / 63.0 for c in rgb)  # on the colour-table grid
            m.roughness = float(rng.uniform(0.05, 1.0))
        mats.append(b.material(m))
    mats = np.array(mats)
    for k, (P, N, I) in enumerate(templates):
        sel = np.nonzero(kinds == k)[0]
        pos = np.einsum("nij,vj->nvi", rots[sel], P) * scales[sel, None, None] + centres[sel, None, :]
        nrm = np.einsum("nij,vj->nvi", rots[sel], N)
        b.add_many(pos, nrm, I, mats[sel])
    fm = preset_material("PLASTIC")
    fm.albedo[:] = (0.6, 0.6, 0.6)
    fm.roughness = 0.8
    floor = b.material(fm)
    ext = cube * 3.0
    b.quad((-ext / 2, -cube / 2 - 1.0, ext / 2), (ext, 0, 0), (0, 0, -ext), floor)  # floor, +y
    lm = preset_material("PLASTIC")
    lm.albedo[:] = (0.5, 0.5, 0.5)
    lm.specular_amount = 0.0
    lm.roughness = 1.0
    lm.emissive[:] = (light_emission,) * 3
    if spectral_extensions:
        lm.blackbody_kelvin = 6500.0
    light = b.material(lm)
    side = cube * 0.4
    for i in range(n_lights):
        cx = (-0.5 + (i % 2)) * cube * 0.5 - side / 2
        cz = (-0.5 + (i // 2)) * cube * 0.5 - side / 2
        b.quad((cx, cube * 0.6, cz), (side, 0, 0), (0, 0, side), light)  # facing -y
    if emissive_triangles:
        er = np.random.default_rng(seed + 1000)
        for _ in range(emissive_triangles):
            em = preset_material("PLASTIC")
            em.albedo[:] = (0.5, 0.5, 0.5)
            em.specular_amount = 0.0
            em.roughness = 1.0
            em.material_mask = _lib.YPT_MATERIAL_TWO_SIDED
            kelvin = er.uniform(2700.0, 9000.0)
            em.emissive[:] = tuple(float(v) for v in blackbody_rgb(kelvin) * triangle_emission)
            if spectral_extensions:
                em.blackbody_kelvin = float(kelvin)
            mid = b.material(em)
            c = er.uniform(-cube / 2, cube / 2, size=3)
            e1, e2 = er.normal(size=3), er.normal(size=3)
            e1 *= 1.5 / np.linalg.norm(e1)
            e2 -= e1 * (e2 @ e1) / (e1 @ e1)
            e2 *= 1.5 / np.linalg.norm(e2)
            P = np.array([c, c + e1, c + e2], np.float32)
            n = np.cross(e1, e2)
            n /= np.linalg.norm(n)
            b.add(P, np.tile(n, (3, 1)).astype(np.float32), np.array([[0, 1, 2]]), mid)
    cam = Camera(eye=(0.0, cube * 0.35, cube * 1.35), target=(0.0, -cube * 0.1, 0.0), fovy=math.pi * 0.3)
    return b.arrays(), cam
