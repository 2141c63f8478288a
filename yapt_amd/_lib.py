"""ctypes binding of the C ABI declared in include/yapt_pt.h.

The shared library ``libyapt_pt.so`` is built in-tree by ``yapt_amd.build``. Loading fails loudly
when it is missing: there is no CPU fallback for the product path.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyapt_pt.so")

YPT_INTEGRATOR_BPT = 0
YPT_INTEGRATOR_NEE = 1
YPT_INTEGRATOR_BDPT = 2
YPT_FLAG_DISABLE_JITTER = 1 << 1
YPT_FLAG_WHITE_FURNACE = 1 << 8
YPT_FLAG_BDPT_NODE_RANGES = 1 << 9
YPT_MATERIAL_TWO_SIDED = 1 << 0
YPT_MATERIAL_DISPERSION = 1 << 1
YPT_MATERIAL_ALPHA_BLEND = 1 << 2

F3 = ctypes.c_float * 3
F16 = ctypes.c_float * 16


class Material(ctypes.Structure):
    """ypt_material: the reference's MaterialParameters (MaterialInternal.h:41-125)."""
    _fields_ = [
        ("albedo", F3), ("transparency", ctypes.c_float),
        ("specular", F3), ("metalness", ctypes.c_float),
        ("absorption", F3), ("dielectric_ior", ctypes.c_float),
        ("emissive", F3), ("roughness", ctypes.c_float),
        ("anisotropy", ctypes.c_float), ("anisotropy_rotation", ctypes.c_float),
        ("material_mask", ctypes.c_uint32), ("specular_amount", ctypes.c_float),
        ("clear_coat_amount", ctypes.c_float), ("clear_coat_ior", ctypes.c_float),
        ("clear_coat_roughness", ctypes.c_float), ("sheen_tint", F3),
        ("sheen_roughness", ctypes.c_float), ("sheen_amount", ctypes.c_float),
        ("thin_film_thickness_nm", ctypes.c_float), ("emission_focus", ctypes.c_float),
        ("alpha_cutoff", ctypes.c_float), ("cauchy_coeffs", ctypes.c_float * 2),
        ("subsurface_radius", F3), ("subsurface_amount", ctypes.c_float),
        ("subsurface_radius_scale", ctypes.c_float),
        ("albedo_tex", ctypes.c_int32), ("normal_tex", ctypes.c_int32), ("orm_tex", ctypes.c_int32),
        ("emissive_tex", ctypes.c_int32),
        ("albedo_tex_scale", ctypes.c_float * 2), ("normal_tex_scale", ctypes.c_float * 2),
        ("orm_tex_scale", ctypes.c_float * 2), ("emissive_tex_scale", ctypes.c_float * 2),
        ("blackbody_kelvin", ctypes.c_float), ("sellmeier_b", F3), ("sellmeier_c", F3),
    ]


class Texture(ctypes.Structure):
    """ypt_texture: RGBA8 texels, srgb = colour data decoded before filtering"""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("srgb", ctypes.c_uint32),
                ("rgba8", ctypes.POINTER(ctypes.c_uint8))]


YPT_ENVIRONMENT_NONE, YPT_ENVIRONMENT_CUBE, YPT_ENVIRONMENT_LONGLAT = 0, 1, 2


class Environment(ctypes.Structure):
    """ypt_environment: the World.Skycube texture, linear RGBA float texels (six square faces +X, -X, +Y, -Y,
    +Z, -Z for a cube map, one image for long-lat)"""
    _fields_ = [("type", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("rgba", ctypes.POINTER(ctypes.c_float))]


def environment_texels(kind, texels):
    """validated float32 texels of an environment: H x W x 4 (or x 3, alpha 1 added) for long-lat, 6 x N x N x 4
    (or x 3) for a cube map"""
    a = np.asarray(texels, dtype=np.float32)
    if a.ndim not in (3, 4) or a.shape[-1] not in (3, 4):
        raise ValueError("environment texels are H x W x 4 (long-lat) or 6 x N x N x 4 (cube)")
    if a.shape[-1] == 3:
        a = np.concatenate([a, np.ones(a.shape[:-1] + (1,), np.float32)], axis=-1)
    if kind == YPT_ENVIRONMENT_CUBE and (a.ndim != 4 or a.shape[0] != 6 or a.shape[1] != a.shape[2]):
        raise ValueError("a cube map is 6 square faces: 6 x N x N x 4")
    if kind == YPT_ENVIRONMENT_LONGLAT and a.ndim != 3:
        raise ValueError("a long-lat environment is one H x W x 4 image")
    if kind not in (YPT_ENVIRONMENT_CUBE, YPT_ENVIRONMENT_LONGLAT):
        raise ValueError("unknown environment type %r" % (kind,))
    return np.ascontiguousarray(a)


YPT_LIGHT_POINT, YPT_LIGHT_SPOT, YPT_LIGHT_DIRECTIONAL = 1, 2, 3


class PunctualLight(ctypes.Structure):
    """ypt_punctual_light: PunctualLightParameters (LightManager.h:15-23)."""
    _fields_ = [
        ("type", ctypes.c_uint32), ("position", ctypes.c_float * 3), ("direction", ctypes.c_float * 3),
        ("intensity", ctypes.c_float * 3), ("cos_inner", ctypes.c_float), ("cos_outer", ctypes.c_float),
    ]


class Light:
    """Host-side light with the py_yapt Light surface (PyLight.cpp:17-35, LightProxy.cpp): type, color,
    intensity, spot cone angles (outer clamped to [0, pi/2], inner to [0, outer]) and a look-at
    transform. params() packs it as LightManager::fillParameters does (LightManager.cpp:97-107):
    direction = transform * (0, 0, -1), intensity = color * max(0, intensity)."""
    POINT, SPOT, DIRECTIONAL = YPT_LIGHT_POINT, YPT_LIGHT_SPOT, YPT_LIGHT_DIRECTIONAL

    def __init__(self, light_type=YPT_LIGHT_POINT):
        self.type = int(light_type)
        self.color = np.ones(3, np.float32)
        self.intensity = 1.0
        self.spot_inner = 0.0
        self.spot_outer = math.pi * 0.25
        self.position = np.zeros(3, np.float32)
        self.forward = np.array([0.0, 0.0, -1.0], np.float32)

    def setType(self, t):
        self.type = int(t)

    def setColor(self, rgb):
        self.color = np.asarray(rgb, np.float32)

    def setIntensity(self, v):
        self.intensity = float(v)

    def setSpotConeAngles(self, inner, outer):
        self.spot_outer = min(max(float(outer), 0.0), math.pi * 0.5)
        self.spot_inner = min(max(float(inner), 0.0), self.spot_outer)

    def lookAt(self, position, target):
        """the light's -z axis towards target (Transform::lookAt)"""
        self.position = np.asarray(position, np.float32)
        d = np.asarray(target, np.float32) - self.position
        n = float(np.linalg.norm(d))
        self.forward = d / n if n > 0 else np.array([0.0, 0.0, -1.0], np.float32)

    def params(self):
        p = PunctualLight()
        p.type = self.type
        p.position[:] = [float(x) for x in self.position]
        p.direction[:] = [float(x) for x in self.forward]
        p.intensity[:] = [float(c) * max(0.0, self.intensity) for c in self.color]
        p.cos_inner = math.cos(self.spot_inner)
        p.cos_outer = math.cos(self.spot_outer)
        return p


class PostParams(ctypes.Structure):
    """ypt_post_params: the denoise and tonemap renderer variables (RendererVarsList.h:24-54)."""
    _fields_ = [
        ("denoise", ctypes.c_uint32), ("prefilter_features", ctypes.c_uint32),
        ("depth_sigma_scale", ctypes.c_float), ("normal_sigma", ctypes.c_float), ("material_sigma", ctypes.c_float),
        ("color_sigma", ctypes.c_float), ("feature_variance_scale", ctypes.c_float),
        ("tonemap", ctypes.c_uint32), ("exposure", ctypes.c_float),
    ]

    @classmethod
    def defaults(cls, **kw):
        p = cls(0, 1, 0.02, 0.06, 0.05, 3.0, 1.0, 1, 1.0)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.POINTER(ctypes.c_float)), ("normals", ctypes.POINTER(ctypes.c_float)),
        ("tangents", ctypes.POINTER(ctypes.c_float)), ("vertex_count", ctypes.c_uint32), ("indices", ctypes.POINTER(ctypes.c_uint32)),
        ("triangle_count", ctypes.c_uint32), ("instance_first_triangle", ctypes.POINTER(ctypes.c_uint32)),
        ("instance_triangle_count", ctypes.POINTER(ctypes.c_uint32)),
        ("instance_material", ctypes.POINTER(ctypes.c_uint32)), ("instance_count", ctypes.c_uint32),
        ("materials", ctypes.POINTER(Material)), ("material_count", ctypes.c_uint32),
        ("punctual_lights", ctypes.POINTER(PunctualLight)), ("punctual_light_count", ctypes.c_uint32),
        ("uvs", ctypes.POINTER(ctypes.c_float)), ("textures", ctypes.POINTER(Texture)), ("texture_count", ctypes.c_uint32),
        ("environment", ctypes.POINTER(Environment)),
    ]


class FrameDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
        ("uv_to_view", F16), ("view_to_uv", F16), ("world_to_view", F16), ("view_to_world", F16),
        ("aperture_radius", ctypes.c_float), ("focus_distance", ctypes.c_float),
        ("forward_z_sign", ctypes.c_float), ("env_intensity_scale", ctypes.c_float),
        ("max_depth", ctypes.c_uint32), ("regularize_after", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("firefly_clamp_scale", ctypes.c_float), ("firefly_clamp_relax_power", ctypes.c_float),
        ("kill_nan_samples", ctypes.c_uint32), ("light_path_vertices", ctypes.c_uint32),
        ("bdpt_node_ranges", ctypes.c_uint32 * 4),
    ]


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _uptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class SceneArrays:
    """Keeps the numpy arrays a SceneDesc points into alive."""

    def __init__(self, positions, normals, indices, inst_first, inst_count, inst_material, materials, tangents=None,
                 lights=(), uvs=None, textures=(), environment=None):
        """textures: (H x W x 4 uint8 array, srgb) pairs; uvs: V x 2 floats; environment: None or
        (YPT_ENVIRONMENT_CUBE or YPT_ENVIRONMENT_LONGLAT, float texels as environment_texels takes them)"""
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        self.tangents = None if tangents is None else np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 4)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.inst_first = np.ascontiguousarray(inst_first, dtype=np.uint32)
        self.inst_count = np.ascontiguousarray(inst_count, dtype=np.uint32)
        self.inst_material = np.ascontiguousarray(inst_material, dtype=np.uint32)
        self.materials = (Material * len(materials))(*materials)
        self.lights = [l if isinstance(l, PunctualLight) else l.params() for l in lights]
        self.light_array = (PunctualLight * max(1, len(self.lights)))(*self.lights)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        self.texel_arrays = [np.ascontiguousarray(t, dtype=np.uint8) for t, _ in textures]
        texs = [Texture(a.shape[1], a.shape[0], 1 if srgb else 0, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
                for a, (_, srgb) in zip(self.texel_arrays, textures)]
        for a in self.texel_arrays:
            if a.ndim != 3 or a.shape[2] != 4:
                raise ValueError("textures must be H x W x 4 uint8")
        self.texture_array = (Texture * max(1, len(texs)))(*texs)
        self.desc = SceneDesc(
            _fptr(self.positions), None if self.normals is None else _fptr(self.normals),
            None if self.tangents is None else _fptr(self.tangents), len(self.positions),
            _uptr(self.indices), len(self.indices), _uptr(self.inst_first), _uptr(self.inst_count),
            _uptr(self.inst_material), len(self.inst_first), self.materials, len(materials),
            self.light_array, len(self.lights), None if self.uvs is None else _fptr(self.uvs), self.texture_array,
            len(self.texel_arrays))
        self.set_environment(environment)

    def set_environment(self, environment):
        """the World.Skycube texture of the scene (None: the texture-less |direction| sky)"""
        self.env_texels, self.env = None, None
        if environment is not None:
            kind, texels = environment
            self.env_texels = environment_texels(int(kind), texels)
            h, w = self.env_texels.shape[-3], self.env_texels.shape[-2]
            self.env = Environment(int(kind), w, h, _fptr(self.env_texels))
        self.desc.environment = ctypes.pointer(self.env) if self.env is not None else None

    @property
    def triangle_count(self):
        return len(self.indices)


YPT_E_INVALID, YPT_E_DEVICE, YPT_E_MEMORY, YPT_E_NOSCENE, YPT_E_CANCELLED, YPT_E_TIMEOUT = -1, -2, -3, -4, -5, -6


class YptError(RuntimeError):
    """an error status of the library; `code` is the YPT_E_* value"""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


_LIB = None


def share_hip_runtime():
    """PyTorch's wheel carries its own HIP runtime under the soname the library links
    (libamdhip64.so.7), and a process serves both from whichever copy was loaded first; if the system's
    copy comes first, torch later finds no GPU. Importing torch (without initialising its device) before
    the library is opened makes one runtime serve both, so device tensors and streams pass between them."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Loads libyapt_pt.so (built by yapt_amd.build). Raises when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise YptError("libyapt_pt.so is not built; run yapt_amd.build.build() (no CPU fallback exists)")
    share_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.ypt_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    lib.ypt_destroy.argtypes = [P]
    lib.ypt_set_scene.argtypes = [P, ctypes.POINTER(SceneDesc)]
    lib.ypt_set_frame.argtypes = [P, ctypes.POINTER(FrameDesc)]
    lib.ypt_clear_film.argtypes = [P]
    lib.ypt_render.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, P]
    lib.ypt_synchronize.argtypes = [P]
    lib.ypt_get_film.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)]
    lib.ypt_get_last_samples.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
    lib.ypt_get_first_hit_features.argtypes = [P, ctypes.POINTER(ctypes.c_float)]
    lib.ypt_get_stats.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.ypt_set_option.argtypes = [P, ctypes.c_char_p, ctypes.c_int64]
    lib.ypt_trace_closest.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_float)]
    lib.ypt_sample_dims.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]
    lib.ypt_sobol_point.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    lib.ypt_last_error.restype = ctypes.c_char_p
    lib.ypt_version.restype = ctypes.c_char_p
    lib.ypt_scene_create.argtypes = [ctypes.POINTER(SceneDesc), ctypes.c_int, ctypes.POINTER(P)]
    lib.ypt_scene_destroy.argtypes = [P]
    lib.ypt_scene_info.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]
    lib.ypt_set_resident_scene.argtypes = [P, P]
    lib.ypt_server_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(P)]
    lib.ypt_server_destroy.argtypes = [P]
    lib.ypt_server_submit.argtypes = [P, P, ctypes.c_int, ctypes.POINTER(FrameDesc), ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_uint64)]
    lib.ypt_server_wait.argtypes = [P, ctypes.c_uint64, ctypes.c_double]
    lib.ypt_server_set_option.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64]
    lib.ypt_final_color.argtypes = [P, ctypes.POINTER(PostParams), ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_float)]
    lib.ypt_post_params_default.argtypes = [ctypes.POINTER(PostParams)]
    lib.ypt_post_params_default.restype = None
    lib.ypt_server_take.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    lib.ypt_server_stats.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.ypt_server_cancel.argtypes = [P, ctypes.c_uint64]
    lib.ypt_server_set_deadline.argtypes = [P, ctypes.c_uint64, ctypes.c_double]
    lib.ypt_server_fused_stats.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    lib.ypt_refit_scene.argtypes = [P, ctypes.POINTER(SceneDesc)]
    lib.ypt_scene_bvh_counts.argtypes = [P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    lib.ypt_clock.restype = ctypes.c_double
    _LIB = lib
    return lib


EXPORTED_SYMBOLS = [
    "ypt_create", "ypt_destroy", "ypt_set_scene", "ypt_set_frame", "ypt_clear_film", "ypt_render",
    "ypt_synchronize", "ypt_get_film", "ypt_get_last_samples", "ypt_get_first_hit_features", "ypt_get_stats",
    "ypt_set_option", "ypt_trace_closest", "ypt_sample_dims", "ypt_sobol_point", "ypt_last_error", "ypt_version",
    "ypt_scene_create", "ypt_scene_destroy", "ypt_scene_info", "ypt_set_resident_scene",
    "ypt_server_create", "ypt_server_destroy", "ypt_server_submit", "ypt_server_wait", "ypt_server_take",
    "ypt_server_stats", "ypt_clock", "ypt_server_set_option", "ypt_final_color", "ypt_post_params_default",
    "ypt_get_feature_moments", "ypt_get_film_variance", "ypt_trace_closest_noalloc", "ypt_server_cancel",
    "ypt_server_set_deadline", "ypt_server_fused_stats", "ypt_refit_scene", "ypt_scene_bvh_counts",
]


def check(rc):
    if rc != 0:
        raise YptError("yapt_pt error %d: %s" % (rc, load().ypt_last_error().decode()), rc)


class Integrator:
    """One path-integrator instance on one GPU (the role of PathIntegratorSubStage)."""

    def __init__(self, integrator_type=YPT_INTEGRATOR_NEE, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.ypt_create(int(integrator_type), int(device), ctypes.byref(h)))
        self.h = h
        self.width = self.height = 0
        self._scene = None

    def close(self):
        if self.h:
            self.lib.ypt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, name, value):
        check(self.lib.ypt_set_option(self.h, name.encode(), int(value)))

    def set_scene(self, scene):
        if isinstance(scene, ResidentScene):
            check(self.lib.ypt_set_resident_scene(self.h, scene.h))
        else:
            check(self.lib.ypt_set_scene(self.h, ctypes.byref(scene.desc)))
        self._scene = scene

    def set_frame(self, frame):
        check(self.lib.ypt_set_frame(self.h, ctypes.byref(frame)))
        self.width, self.height = frame.width, frame.height

    def refit_scene(self, scene):
        """moved vertex data for the scene's triangles: refits the BVH instead of rebuilding it"""
        check(self.lib.ypt_refit_scene(self.h, ctypes.byref(scene.desc)))
        self._scene = scene

    def bvh_counts(self):
        """(builds, refits) of the handle's scene"""
        b, r = ctypes.c_uint32(), ctypes.c_uint32()
        check(self.lib.ypt_scene_bvh_counts(self.h, ctypes.byref(b), ctypes.byref(r)))
        return b.value, r.value

    def clear_film(self):
        check(self.lib.ypt_clear_film(self.h))

    def render(self, sample_begin, sample_count, stream=None):
        check(self.lib.ypt_render(self.h, int(sample_begin), int(sample_count), stream))

    def synchronize(self):
        check(self.lib.ypt_synchronize(self.h))

    def film(self, with_counts=False):
        out = np.zeros((self.height, self.width, 4), np.float32)
        cnt = np.zeros((self.height, self.width), np.uint32) if with_counts else None
        check(self.lib.ypt_get_film(self.h, _fptr(out), None if cnt is None else _uptr(cnt)))
        return (out, cnt) if with_counts else out

    def last_samples(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(self.lib.ypt_get_last_samples(self.h, _fptr(out)))
        return out

    def final_color(self, params=None, linear=False):
        """FinalColor: H x W x 4 uint8 (display encoded); with linear=True also the H x W x 4 float32
        linear sRGB image before exposure (the denoised film)."""
        p = params if params is not None else PostParams.defaults()
        rgba = np.zeros((self.height, self.width, 4), np.uint8)
        lin = np.zeros((self.height, self.width, 4), np.float32) if linear else None
        check(self.lib.ypt_final_color(self.h, ctypes.byref(p), rgba.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                       None if lin is None else _fptr(lin)))
        return (rgba, lin) if linear else rgba

    def feature_moments(self):
        out = np.zeros((4, self.height, self.width, 4), np.float32)
        check(self.lib.ypt_get_feature_moments(self.h, _fptr(out)))
        return out

    def film_variance(self):
        out = np.zeros((self.height, self.width), np.float32)
        check(self.lib.ypt_get_film_variance(self.h, _fptr(out)))
        return out

    def first_hit_features(self):
        out = np.zeros((self.height, self.width, 12), np.float32)
        check(self.lib.ypt_get_first_hit_features(self.h, _fptr(out)))
        return out

    def stats(self):
        out = np.zeros(16, np.float64)
        check(self.lib.ypt_get_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        t = np.zeros(n, np.float32)
        tri = np.zeros(n, np.uint32)
        b = np.zeros((n, 2), np.float32)
        check(self.lib.ypt_trace_closest(self.h, _fptr(rays), n, _fptr(t), _uptr(tri), _fptr(b)))
        return t, tri, b

    def sample_dims(self, sample_index, seeds, dims):
        seeds = np.ascontiguousarray(seeds, np.uint32)
        dims = np.ascontiguousarray(dims, np.uint32)
        out = np.zeros(len(seeds), np.float32)
        check(self.lib.ypt_sample_dims(self.h, int(sample_index), _uptr(seeds), _uptr(dims), len(seeds), _fptr(out)))
        return out


def native():
    """the pybind11 binding of the C ABI (yapt_amd/_yapt_native, built by yapt_amd.build), imported after
    the library is loaded so that both use one copy of it and of the HIP runtime. Raises when it is
    missing."""
    load()
    try:
        from . import _yapt_native
    except ImportError as e:
        raise YptError("the pybind11 module _yapt_native is not built; run yapt_amd.build.build() (%s)" % e)
    return _yapt_native


def sobol_point(index):
    out = np.zeros(256, np.uint32)
    check(load().ypt_sobol_point(int(index), _uptr(out)))
    return out


class ResidentScene:
    """A scene built once on a device (BVH and buffers) and shared by integrators and render jobs."""

    def __init__(self, scene, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.ypt_scene_create(ctypes.byref(scene.desc), int(device), ctypes.byref(h)))
        self.h = h
        self.device = device
        nbytes, ntris = ctypes.c_uint64(), ctypes.c_uint32()
        check(self.lib.ypt_scene_info(self.h, ctypes.byref(nbytes), ctypes.byref(ntris)))
        self.device_bytes, self.triangle_count = nbytes.value, ntris.value

    def close(self):
        if self.h:
            self.lib.ypt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Render server: continuous batching of progressive jobs over resident scenes (C ABI ypt_server_*)."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, max_active_jobs=8):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.ypt_server_create(int(device), int(hbm_budget_bytes), int(max_active_jobs), ctypes.byref(h)))
        self.h = h
        self._frames = {}

    def submit(self, scene, integrator, frame, spp):
        jid = ctypes.c_uint64()
        check(self.lib.ypt_server_submit(self.h, scene.h, int(integrator), ctypes.byref(frame), int(spp), ctypes.byref(jid)))
        self._frames[jid.value] = (frame.width, frame.height)
        return jid.value

    def set_option(self, name, value, integrator=-1):
        check(self.lib.ypt_server_set_option(self.h, int(integrator), name.encode(), int(value)))

    def wait(self, job, timeout=-1.0):
        """True when the job has finished, False when `timeout` seconds passed first; raises YptError for a
        job that failed, was cancelled (code YPT_E_CANCELLED) or passed its deadline (YPT_E_TIMEOUT)."""
        rc = self.lib.ypt_server_wait(self.h, int(job), float(timeout))
        if rc == 1:
            return False
        check(rc)
        return True

    def cancel(self, job):
        check(self.lib.ypt_server_cancel(self.h, int(job)))

    def set_deadline(self, job, seconds):
        check(self.lib.ypt_server_set_deadline(self.h, int(job), float(seconds)))

    def take(self, job):
        """Film (H x W x 4) and (submit, admit, finish) times of a finished job."""
        w, h = self._frames.pop(job)
        film = np.zeros((h, w, 4), np.float32)
        times = np.zeros(3, np.float64)
        check(self.lib.ypt_server_take(self.h, int(job), _fptr(film), times.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return film, times

    def forget(self, job):
        """drops a job that ended without a film (cancelled, timed out or failed); returns its status code"""
        self._frames.pop(job, None)
        return self.lib.ypt_server_take(self.h, int(job), None, None)

    def stats(self):
        out = np.zeros(6, np.float64)
        check(self.lib.ypt_server_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return {"queued": int(out[0]), "active": int(out[1]), "completed": int(out[2]), "reserved_bytes": out[3],
                "peak_active": int(out[4]), "budget_bytes": out[5]}

    def fused_stats(self):
        """fused steps issued so far and the job chunks they carried"""
        steps, chunks = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.lib.ypt_server_fused_stats(self.h, ctypes.byref(steps), ctypes.byref(chunks)))
        return {"steps": steps.value, "job_chunks": chunks.value}

    def clock(self):
        return self.lib.ypt_clock()

    def close(self):
        if self.h:
            self.lib.ypt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
