"""Edge cases of the render path on the GPU against the checker: frames of one pixel and ragged sizes
(partial 8x8 BDPT tiles, partial workgroups), a view where every ray misses (with and without the
sky), a scene of one emitter triangle, the depth limits, a large first sample index, the most passes
in flight, and the inputs the library refuses."""
import numpy as np
import pytest

from test_gpu_parity import _compare_render

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size", [(1, 1), (1, 7), (7, 1), (13, 11), (65, 9)])
@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_ragged_frames(ylib, oracle, cornell, size, integrator):
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, size[0], size[1], max_depth=6, light_path_vertices=5)
    _compare_render(ylib, oracle, scene, frame, integrator, passes=3)


@pytest.mark.parametrize("env", [0.0, 1.5])
def test_every_ray_misses(ylib, oracle, cornell, env):
    # the camera looks away from the box: the film is the sky (zero at scale 0) in every integrator
    from yapt_amd import camera
    scene, _ = cornell
    away = camera.Camera(eye=(0.5, 0.5, -3.0), target=(0.5, 0.5, -6.0), fovy=0.6)
    frame = camera.make_frame(away, 24, 16, max_depth=4, env_intensity_scale=env, light_path_vertices=4)
    for integrator in (0, 1, 2):
        film = _compare_render(ylib, oracle, scene, frame, integrator, passes=2)
        if env == 0.0:
            assert not film[..., :3].any()
        else:
            assert film[..., 1].min() > 0


def test_one_emitter_triangle(ylib, oracle):
    from yapt_amd import camera, scenes
    b = scenes.MeshBuilder()
    e = scenes.blank_material()
    e.emissive[:] = (3.0, 2.0, 1.0)
    b.add([[-1, 0, -1], [0, 0, 1], [1, 0, -1]], [[0, 1, 0]] * 3, [[0, 1, 2]], b.material(e))   # facing +y
    scene = b.arrays()
    assert scene.triangle_count == 1
    cam = camera.Camera(eye=(0, 3, 0.5), target=(0, 0, 0), fovy=0.9)
    frame = camera.make_frame(cam, 20, 20, max_depth=4, light_path_vertices=4)
    for integrator in (0, 1, 2):
        film = _compare_render(ylib, oracle, scene, frame, integrator, passes=2)
        assert film[..., 1].max() > 0


@pytest.mark.parametrize("depth", [1, 32])
def test_depth_limits(ylib, oracle, shapes_small, depth):
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 24, 20, max_depth=depth, light_path_vertices=min(depth + 1, 16))
    for integrator in (1, 2):
        _compare_render(ylib, oracle, scene, frame, integrator, passes=2)


def test_large_first_sample_index(ylib, oracle, cornell):
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 20, 16, max_depth=5)
    _compare_render(ylib, oracle, scene, frame, 1, passes=2, begin=(1 << 20) - 1)


def test_most_passes_in_flight(ylib, oracle, cornell):
    # 130 passes of a tiny frame: one batch of 65 passes (the limit is 128) and the film against the
    # checker, as every batch size gives the same film
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 9, 7, max_depth=5)
    _compare_render(ylib, oracle, scene, frame, 1, passes=130, options={"passes_in_flight": 128})


def test_refused_inputs(ylib, cornell):
    from yapt_amd import camera
    scene, cam = cornell
    it = ylib.Integrator(1, 0)
    with pytest.raises(ylib.YptError):
        it.render(0, 1)                                   # no scene, no frame
    it.set_scene(scene)
    empty_frame = camera.make_frame(cam, 4, 4)
    empty_frame.width = 0
    with pytest.raises(ylib.YptError):
        it.set_frame(empty_frame)                         # empty frame
    with pytest.raises(ylib.YptError):
        it.set_option("passes_in_flight", 129)
    empty = ylib.SceneArrays(np.zeros((0, 3), np.float32), None, np.zeros((0, 3), np.uint32), np.zeros(0, np.uint32),
                             np.zeros(0, np.uint32), np.zeros(0, np.uint32), list(scene.materials))
    with pytest.raises(ylib.YptError):
        it.set_scene(empty)                               # no triangles
    it.close()
