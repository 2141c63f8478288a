"""The benchmark workloads at full size on the GPU, checked against the checker where it is cheap: the
1M-triangle random-shapes scene (seed 2) with closest hits of 50,000 rays bit for bit through both
BVH widths, NEE at 1920x1080 and depth 12 with the per-sample radiance of every 97th pixel of the
last pass, and BDPT at 1280x720 with 8 emissive triangles and light subpaths of 9 vertices (camera and
light subpaths of depth 8, as benchmarked), checked on every 7th camera workgroup. Wavelengths must
match bit for bit and radiance within 1e-4 (in practice every bit matches)."""
import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shapes_1m():
    return scenes.random_shapes(n_shapes=2000, seed=2)


@pytest.fixture(scope="module")
def shapes_1m_bdpt():
    return scenes.random_shapes(n_shapes=2000, seed=2, emissive_triangles=8)


@pytest.mark.parametrize("width", [2, 8])
def test_full_size_closest_hits_are_bit_exact(ylib, oracle, shapes_1m, width):
    scene, cam = shapes_1m
    rng = np.random.default_rng(11)
    n = 50000
    P = scene.positions[rng.integers(0, len(scene.positions), n)]
    o = rng.uniform(-14, 14, (n, 3)).astype(np.float32)
    d = P + rng.normal(0, 0.02, (n, 3)).astype(np.float32) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 1e-4, np.float32), d, np.full((n, 1), 1e30, np.float32)], 1)
    it = ylib.Integrator(_lib.YPT_INTEGRATOR_NEE, 0)
    it.set_option("bvh_width", width)
    it.set_scene(scene)
    t, tri, b = it.trace_closest(rays)
    it.close()
    rt, rtri, rb = oracle.trace_closest(scene, rays, use_bvh=True)
    assert (tri != 0xFFFFFFFF).mean() > 0.9
    assert np.array_equal(tri, rtri)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))


def _compare_last(g_last, ref_last):
    traced = ref_last[..., 3] != 0
    assert traced.sum() > 1000
    g, r = g_last[traced], ref_last[traced]
    assert np.array_equal(g[:, 3].view(np.uint32), r[:, 3].view(np.uint32))
    d = np.abs(g[:, :3].astype(np.float64) - r[:, :3])
    assert np.all((d <= 1e-4 * np.abs(r[:, :3])) | (d == 0)), d.max()
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_full_size_nee_samples_match_the_checker(ylib, oracle, shapes_1m):
    scene, cam = shapes_1m
    frame = camera.make_frame(cam, 1920, 1080, max_depth=12)
    it = ylib.Integrator(_lib.YPT_INTEGRATOR_NEE, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(5, 1)
    g_last = it.last_samples()
    it.close()
    ref = oracle.render(scene, frame, _lib.YPT_INTEGRATOR_NEE, 5, 1, pixel_stride=97)
    _compare_last(g_last, ref["last"])


def test_full_size_bdpt_samples_match_the_checker(ylib, oracle, shapes_1m_bdpt):
    scene, cam = shapes_1m_bdpt
    frame = camera.make_frame(cam, 1280, 720, max_depth=8, light_path_vertices=9)
    it = ylib.Integrator(_lib.YPT_INTEGRATOR_BDPT, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(3, 1)
    g_last = it.last_samples()
    it.close()
    ref = oracle.render(scene, frame, _lib.YPT_INTEGRATOR_BDPT, 3, 1, pixel_stride=7)
    _compare_last(g_last, ref["last"])
