"""Wave-per-ray traversal (one 64-lane wavefront per ray) against the checker and against the
thread-per-ray launches.

The closest hit is the (t, triangle id) minimum, so spreading one ray over the lanes of a wave must
give the same triangle, t and barycentrics to the bit, and a render whose short bounce queues are
traced wave-per-ray must give the same per-sample radiance and film as the checker.
"""
import numpy as np
import pytest

from test_gpu_parity import _compare_render, _rays

pytestmark = pytest.mark.gpu

ALL_WAVE = 1 << 30  # every queue is shorter: all rays wave-per-ray


@pytest.mark.parametrize("width", [2, 8])
def test_wave_closest_hits_bitexact(ylib, oracle, cornell, shapes_small, width):
    for k, (scene, lo, hi, n) in enumerate([(cornell[0], -0.99, 1.99, 6000), (shapes_small[0], -5.0, 5.0, 4000)]):
        it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
        it.set_option("bvh_width", width)
        it.set_option("wave_ray_below", ALL_WAVE)
        it.set_scene(scene)
        rays = _rays(np.random.default_rng(20 + k), n, lo, hi, scene.positions)
        # a bounded segment on some rays, so the box cut-off starts below infinity
        rays[::3, 7] = 4.0
        t, tri, b = it.trace_closest(rays)
        it.close()
        t2, tri2, b2 = oracle.trace_closest(scene, rays, use_bvh=False)
        assert (tri != 0xFFFFFFFF).mean() > 0.3
        assert np.array_equal(tri, tri2), (width, k)
        assert np.array_equal(t.view(np.uint32), t2.view(np.uint32)), (width, k)
        assert np.array_equal(b.view(np.uint32), b2.view(np.uint32)), (width, k)


@pytest.mark.parametrize("integrator", [0, 1])
@pytest.mark.parametrize("width", [2, 8])
def test_wave_render_parity_shapes(ylib, oracle, shapes_small, integrator, width):
    # every extension and shadow queue of every bounce traced wave-per-ray
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 48, 40, max_depth=12)
    _compare_render(ylib, oracle, scene, frame, integrator, passes=2, begin=3, width=width,
                    options={"wave_ray_below": ALL_WAVE})


@pytest.mark.parametrize("width", [2, 8])
def test_wave_bdpt_parity(ylib, oracle, shapes_small, width):
    # the bidirectional camera wavefront's extension and connection queues wave-per-ray
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 48, 40, max_depth=8, light_path_vertices=6)
    _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=2, begin=1, width=width,
                    options={"wave_ray_below": ALL_WAVE, "bdpt_wavefront": 1})


def test_wave_render_parity_punctual_lights(ylib, oracle):
    # occlusion rays without a target triangle (punctual lights) and with one (the area light)
    from yapt_amd import camera, scenes
    scene, cam, env = scenes.punctual_lights("mixed", sphere_res=(12, 12))
    frame = camera.make_frame(cam, 48, 32, max_depth=6, env_intensity_scale=env)
    _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_NEE, passes=2, options={"wave_ray_below": ALL_WAVE})


@pytest.mark.parametrize("opts", [
    {"wave_ray_below": 3000},
    {"wave_ray_below": 3000, "persistent": 0},
    {"wave_ray_below": ALL_WAVE, "count_traversal": 1},
    {"wave_ray_below": 800, "graph": 1},
])
def test_wave_and_thread_launches_give_the_same_film(ylib, shapes_small, opts):
    # long queues thread-per-ray, short ones wave-per-ray, within one render
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 83, 61, max_depth=12)

    def film(o, width):
        it = ylib.Integrator(1, 0)
        it.set_option("bvh_width", width)
        for k, v in o.items():
            it.set_option(k, v)
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(2, 3)
        it.render(5, 2)
        f = it.film()
        it.close()
        return f

    for width in (2, 8):
        base = film({}, width)
        assert np.array_equal(film(opts, width).view(np.uint32), base.view(np.uint32)), (width, opts)


def test_wave_ray_below_is_validated(ylib):
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    with pytest.raises(Exception):
        it.set_option("wave_ray_below", -1)
    it.set_option("wave_ray_below", 0)
    it.close()
