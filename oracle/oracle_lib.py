"""ctypes binding of the CPU checker (liboracle.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("liboracle.so is not built; run oracle/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    FP = ctypes.POINTER(ctypes.c_float)
    UP = ctypes.POINTER(ctypes.c_uint32)
    lib.oracle_render.argtypes = [P, P, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_uint32, FP, UP, FP, FP, ctypes.POINTER(ctypes.c_double)]
    lib.oracle_trace_closest.argtypes = [P, FP, ctypes.c_uint32, ctypes.c_int, FP, UP, FP]
    lib.oracle_sample_dims.argtypes = [ctypes.c_uint32, UP, UP, ctypes.c_uint32, FP]
    lib.oracle_sobol_point.argtypes = [ctypes.c_uint32, UP]
    lib.oracle_get_ms_luts.argtypes = [FP]
    lib.oracle_rgb_coeffs.argtypes = [FP, FP]
    lib.oracle_last_error.restype = ctypes.c_char_p
    _LIB = lib
    return lib


def _f(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d: %s" % (rc, load().oracle_last_error().decode()))


def render(scene, frame, integrator, sample_begin, sample_count, threads=0, use_bvh=True, pixel_stride=1,
           features=False):
    lib = load()
    W, H = frame.width, frame.height
    film = np.zeros((H, W, 4), np.float32)
    counts = np.zeros((H, W), np.uint32)
    last = np.zeros((H, W, 4), np.float32)
    feat = np.zeros((H, W, 12), np.float32) if features else None
    stats = np.zeros(8, np.float64)
    _check(lib.oracle_render(ctypes.addressof(scene.desc), ctypes.addressof(frame), int(integrator), int(sample_begin),
                             int(sample_count), int(threads), int(bool(use_bvh)), int(pixel_stride), _f(film), _u(counts),
                             _f(last), None if feat is None else _f(feat),
                             stats.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    out = dict(film=film, counts=counts, last=last, stats=stats)
    if features:
        out["features"] = feat
    return out


def trace_closest(scene, rays, use_bvh=False):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(rays)
    t = np.zeros(n, np.float32)
    tri = np.zeros(n, np.uint32)
    b = np.zeros((n, 2), np.float32)
    _check(load().oracle_trace_closest(ctypes.addressof(scene.desc), _f(rays), n, int(bool(use_bvh)), _f(t), _u(tri), _f(b)))
    return t, tri, b


def sample_dims(sample_index, seeds, dims):
    seeds = np.ascontiguousarray(seeds, np.uint32)
    dims = np.ascontiguousarray(dims, np.uint32)
    out = np.zeros(len(seeds), np.float32)
    _check(load().oracle_sample_dims(int(sample_index), _u(seeds), _u(dims), len(seeds), _f(out)))
    return out


def sobol_point(index):
    out = np.zeros(256, np.uint32)
    _check(load().oracle_sobol_point(int(index), _u(out)))
    return out


def ms_luts():
    out = np.zeros(16384, np.float32)
    n = load().oracle_get_ms_luts(_f(out))
    return out[:n]


def rgb_coeffs(rgb):
    rgb = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(3, np.float32)
    _check(load().oracle_rgb_coeffs(_f(rgb), _f(out)))
    return out


FN = {"d_ggx": 0, "pdf_wm_ggx": 1, "sample_wm_ggx": 2, "fresnel_dielectric": 3, "fresnel_conductor": 4, "sin": 5,
      "cos": 6, "exp": 7, "log": 8, "pow": 9, "sample_cosine_hemisphere": 10, "sigmoid": 11, "sigmoid_inv": 12,
      "hash_pair": 13, "owen_scramble": 14, "uint_to_float01": 15, "ss_dir": 16, "ss_avg": 17, "ssms_dir": 18,
      "sheen_dir": 19, "tr_dir": 20, "emission_profile": 21, "f32_to_f16": 22, "f16_to_f32": 23, "accumulate_ri": 24,
      "pdf_ratio": 25, "burley_profile_area": 26, "burley_sample_radius": 27, "burley_d": 28,
      "pdf_subsurface_exit_plane": 29, "sellmeier_ior": 30, "planck_relative": 31, "avg_fresnel": 32,
      "glambda_sheen": 33, "d_sheen": 34, "cauchy_ior": 35, "apply_thin_film": 36, "ssms_avg": 37, "tr_avg": 38,
      "atan2": 39, "acos": 40, "asin": 41}


def evaluate(name, args):
    """Evaluates one shared-core function over rows of arguments (up to 8 floats per row)."""
    lib = load()
    lib.oracle_eval.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]
    a = np.zeros((len(args), 8), np.float32)
    arr = np.asarray(args)
    if arr.dtype == np.uint32:
        arr = arr.view(np.float32)
    arr = arr.astype(np.float32, copy=False).reshape(len(args), -1)
    a[:, :arr.shape[1]] = arr
    out = np.zeros((len(args), 4), np.float32)
    _check(lib.oracle_eval(FN[name], _f(a), len(a), _f(out)))
    return out


def rgb_reproduce(coeffs):
    lib = load()
    lib.oracle_rgb_reproduce.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    c = np.ascontiguousarray(coeffs, np.float32)
    out = np.zeros(3, np.float32)
    _check(lib.oracle_rgb_reproduce(_f(c), _f(out)))
    return out


EVAL_NONE, EVAL_PDF_ONLY, EVAL_LIGHT_PATH, EVAL_TREAT_AS_DELTA, EVAL_DONT_PROJECT = 1, 2, 4, 8, 16


def env_color(environment, dirs):
    """getSkyBoxColor of the shared core (RGB per direction) for an environment given as (type, texels) the
    way yapt_amd._lib.SceneArrays takes it, or None for the texture-less sky"""
    from yapt_amd import _lib as ylib
    lib = load()
    lib.oracle_env_color.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_float)]
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros((len(d), 3), np.float32)
    env = None
    if environment is not None:
        texels = ylib.environment_texels(int(environment[0]), environment[1])
        env = ylib.Environment(int(environment[0]), texels.shape[-2], texels.shape[-3], _f(texels))
    _check(lib.oracle_env_color(None if env is None else ctypes.addressof(env), _f(d), len(d), _f(out)))
    return out


def eval_surface(material, wo, wi, hero, flags=EVAL_NONE, reciprocal=False):
    """The layered BSDF of one untextured material (a yapt_amd._lib.Material) at a point whose normal
    is +y and tangent +x, seen from outside: (weight per hero lane (4), pdf, transmission flags)."""
    lib = load()
    lib.oracle_eval_surface.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.c_float, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    o = np.ascontiguousarray(wo, np.float32)
    i = np.ascontiguousarray(wi, np.float32)
    out = np.zeros(6, np.float32)
    _check(lib.oracle_eval_surface(ctypes.addressof(material), _f(o), _f(i), float(hero), int(flags), int(bool(reciprocal)),
                                   _f(out)))
    return out[:4].copy(), float(out[4]), int(out[5])
