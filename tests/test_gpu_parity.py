"""GPU (HIP library through the C ABI) against the CPU checker.

Bars: Sobol dimensions, hero wavelengths and closest-hit triangle ids bit-exact; per-sample spectral
radiance (XYZ of each path) within 1e-4 relative; merged film within 1e-4 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo, hi, scene_pos):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    # aim at random vertices so that most rays hit something, plus fully random directions
    tgt = scene_pos[rng.integers(0, len(scene_pos), size=n)]
    d = tgt - o
    d[n // 2:] = rng.normal(size=(n - n // 2, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = 5e-4
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


@pytest.fixture(scope="module")
def gpu_nee(ylib):
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    yield it
    it.close()


def test_sampler_dims_bitexact(ylib, oracle, gpu_nee, cornell):
    gpu_nee.set_scene(cornell[0])
    rng = np.random.default_rng(1)
    n = 20000
    seeds = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    seeds[:1000] = np.arange(1000)
    dims = rng.integers(0, 300, size=n).astype(np.uint32)  # includes dimensions past the 256-entry point
    for s in (0, 1, 7, 63, 1000):
        g = gpu_nee.sample_dims(s, seeds, dims)
        c = oracle.sample_dims(s, seeds, dims)
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), "sample %d" % s


@pytest.mark.parametrize("width", [2, 8])
def test_closest_hit_bitexact_cornell(ylib, oracle, gpu_nee, cornell, width):
    scene = cornell[0]
    gpu_nee.set_scene(scene)
    gpu_nee.set_option("bvh_width", width)
    rays = _rays(np.random.default_rng(2), 20000, -0.99, 1.99, scene.positions)
    t, tri, b = gpu_nee.trace_closest(rays)
    t2, tri2, b2 = oracle.trace_closest(scene, rays, use_bvh=False)
    gpu_nee.set_option("bvh_width", 2)
    assert np.array_equal(tri, tri2)
    assert np.array_equal(t.view(np.uint32), t2.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))


@pytest.mark.parametrize("width", [2, 8])
def test_closest_hit_bitexact_shapes(ylib, oracle, gpu_nee, shapes_small, width):
    scene = shapes_small[0]
    gpu_nee.set_scene(scene)
    gpu_nee.set_option("bvh_width", width)
    rays = _rays(np.random.default_rng(3), 4000, -5.0, 5.0, scene.positions)
    t, tri, b = gpu_nee.trace_closest(rays)
    t2, tri2, b2 = oracle.trace_closest(scene, rays, use_bvh=False)
    gpu_nee.set_option("bvh_width", 2)
    assert (tri != 0xFFFFFFFF).mean() > 0.3
    assert np.array_equal(tri, tri2)
    assert np.array_equal(t.view(np.uint32), t2.view(np.uint32))


def _compare_render(ylib, oracle, scene, frame, integrator, passes, begin=0, width=2, options=None):
    it = ylib.Integrator(integrator, 0)
    it.set_option("bvh_width", width)
    for k, v in (options or {}).items():
        it.set_option(k, v)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(begin, passes)
    g_last, g_film = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(scene, frame, integrator, begin, passes, use_bvh=True)
    c_last, c_film = ref["last"], ref["film"]
    assert np.array_equal(g_last[..., 3].view(np.uint32), c_last[..., 3].view(np.uint32)), "hero wavelengths"
    d = np.abs(g_last[..., :3].astype(np.float64) - c_last[..., :3])
    # every per-sample value within 1e-4 relative (the bar), with no exempt fraction; in practice the
    # shared FP32 core, the same hits and one fixed order of every sum give the same bits
    rel = d / np.maximum(np.abs(c_last[..., :3]), 1e-30)
    assert np.all((rel <= 1e-4) | (d <= 1e-12)), "per-sample radiance: max rel %.3g" % rel.max()
    assert np.array_equal(g_last.view(np.uint32), c_last.view(np.uint32)), "per-sample XYZ bits (max |d| %.3g)" % d.max()
    assert np.array_equal(g_film.view(np.uint32), c_film.view(np.uint32)), "film bits"
    return g_film


@pytest.mark.parametrize("integrator", [0, 1])
def test_render_parity_cornell(ylib, oracle, cornell, integrator):
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 96, 96, max_depth=8)
    _compare_render(ylib, oracle, scene, frame, integrator, passes=4)


def test_render_parity_cornell_configured_size(ylib, oracle, cornell):
    # the Cornell configuration as stated: NEE, 256x256, 16 spp, depth 8
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 256, 256, max_depth=8)
    film = _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_NEE, passes=16)
    assert film[..., 1].mean() > 0.02


def test_render_parity_shapes_nee(ylib, oracle, shapes_small):
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 80, 60, max_depth=12)
    _compare_render(ylib, oracle, scene, frame, 1, passes=2, begin=5)
    _compare_render(ylib, oracle, scene, frame, 1, passes=1, begin=9, width=8)


@pytest.mark.parametrize("opts", [
    {"persistent": 0},
    {"persistent": 1, "persistent_waves": 1, "refill_threshold": 1},
    {"persistent": 1, "persistent_waves": 2, "refill_threshold": 64, "block_size": 192},
    {"persistent": 1, "persistent_waves": 32, "refill_threshold": 16, "passes_in_flight": 3},
    {"persistent": 1, "leaf_share": 1},
    {"persistent": 1, "leaf_share": 64, "refill_threshold": 8},
    {"persistent": 1, "block_size": 512},
])
def test_launch_shapes_give_the_same_film(ylib, oracle, shapes_small, opts):
    # persistent or one-workgroup-per-entry launches, any refill threshold, leaf share, block size or
    # number of passes in flight: every path touches only its own slot, so the films agree bit for bit
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
        it.render(2, 5)
        f = it.film()
        it.close()
        return f

    for width in (2, 8):
        base = film({"persistent": 0, "passes_in_flight": 1}, width)
        assert np.array_equal(film(opts, width).view(np.uint32), base.view(np.uint32)), (width, opts)


@pytest.mark.parametrize("size,width", [((64, 64), 2), ((83, 61), 2), ((64, 48), 8)])
def test_bdpt_parity_cornell(ylib, oracle, cornell, size, width):
    # light subpaths, camera subpaths with every connection, light tracing splats summed per texel
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, size[0], size[1], max_depth=8)
    _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=3, width=width)


def test_bdpt_parity_shapes(ylib, oracle, shapes_small):
    # glass with dispersion, conductors and rough dielectrics on both subpaths
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 80, 56, max_depth=8, light_path_vertices=8)
    _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=2, begin=4)


@pytest.mark.parametrize("light_type,integrator", [("mixed", 0), ("mixed", 1), ("spot", 1), ("directional", 0)])
def test_render_parity_punctual_lights(ylib, oracle, light_type, integrator):
    # punctual lights sampled explicitly (BPT: punctual table only; NEE: all lights), occlusion rays
    # without a target triangle
    from yapt_amd import camera, scenes
    scene, cam, env = scenes.punctual_lights(light_type, sphere_res=(12, 12))
    frame = camera.make_frame(cam, 72, 44, max_depth=6, env_intensity_scale=env)
    film = _compare_render(ylib, oracle, scene, frame, integrator, passes=3)
    assert film[..., 1].max() > 0


@pytest.mark.parametrize("light_type", ["mixed", "directional", "env"])
def test_bdpt_parity_punctual_and_environment_lights(ylib, oracle, light_type):
    # light subpaths from point, spot and directional lights, or from the environment when the scene
    # has no lights; connections to origins without geometry
    from yapt_amd import camera, scenes
    scene, cam, env = scenes.punctual_lights(light_type, sphere_res=(12, 12))
    frame = camera.make_frame(cam, 72, 48, max_depth=6, env_intensity_scale=env, light_path_vertices=6)
    for wf in (1, 0):
        _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=2, options={"bdpt_wavefront": wf})


def test_bdpt_first_hit_features_match_nee(ylib, cornell):
    # the first camera vertex is the same surface point in both integrators (same texel jitter)
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 40, 32, max_depth=4)
    out = []
    for integ in (ylib.YPT_INTEGRATOR_NEE, ylib.YPT_INTEGRATOR_BDPT):
        it = ylib.Integrator(integ, 0)
        it.set_option("first_hit_features", 1)
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(0, 1)
        out.append(it.first_hit_features())
        it.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("integrator", [0, 1])
def test_graph_batches_match_eager(ylib, shapes_small, integrator):
    # captured batches replayed several times, with eager renders in between, give the eager film
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 64, 48, max_depth=10)

    def handle(graph):
        it = ylib.Integrator(integrator, 0)
        it.set_option("passes_in_flight", 8)
        it.set_option("graph", graph)
        it.set_scene(scene)
        it.set_frame(frame)
        return it

    a = handle(0)
    a.render(0, 20)
    want, want_last = a.film(), a.last_samples()
    a.close()
    b = handle(1)
    b.render(0, 8)
    b.set_option("graph", 0)
    b.render(8, 4)
    b.set_option("graph", 1)
    b.render(12, 8)
    got = b.film()
    b.close()
    c = handle(1)
    c.render(0, 20)             # batches of 8, 8 and 4: two graphs
    got_c, got_c_last = c.film(), c.last_samples()
    c.close()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(got_c.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(got_c_last.view(np.uint32), want_last.view(np.uint32))


def test_bdpt_wavefront_matches_megakernel(ylib, shapes_small):
    # the same camera subpaths traced as one kernel per texel or as queues
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 72, 48, max_depth=8, light_path_vertices=8)
    out = []
    for wf in (0, 1):
        it = ylib.Integrator(ylib.YPT_INTEGRATOR_BDPT, 0)
        it.set_option("bdpt_wavefront", wf)
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(3, 2)
        out.append((it.last_samples(), it.film()))
        it.close()
    (la, fa), (lb, fb) = out
    # connection records added in queue order: the wavefront adds in the megakernel's order
    assert np.array_equal(la.view(np.uint32), lb.view(np.uint32))
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


@pytest.mark.parametrize("split", [1, 2, 4])
def test_bdpt_split_connections_match(ylib, oracle, shapes_small, split):
    # connections of each camera vertex in a launch of their own, one thread per light vertex: the
    # checker's per-sample radiance, and the film of the in-loop form bit for bit (records of
    # connections that carry nothing add +0 between the others)
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 72, 48, max_depth=8, light_path_vertices=8)
    _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=2, begin=1,
                    options={"bdpt_wavefront": 1, "bdpt_split_connections": split})
    out = []
    for opts in ({"bdpt_split_connections": 0}, {"bdpt_split_connections": split},
                 {"bdpt_split_connections": split, "bdpt_waves": 2}):
        it = ylib.Integrator(ylib.YPT_INTEGRATOR_BDPT, 0)
        it.set_option("bdpt_wavefront", 1)
        for k, v in opts.items():
            it.set_option(k, v)
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(3, 2)
        out.append((it.last_samples(), it.film()))
        it.close()
    for la, fa in out[1:]:
        assert np.array_equal(la.view(np.uint32), out[0][0].view(np.uint32))
        assert np.array_equal(fa.view(np.uint32), out[0][1].view(np.uint32))


def test_bdpt_split_connections_is_validated(ylib):
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_BDPT, 0)
    with pytest.raises(Exception):
        it.set_option("bdpt_split_connections", 3)
    it.set_option("bdpt_split_connections", 0)
    it.close()


@pytest.mark.gpu
def test_bdpt_light_stage_overlap_matches_one_stream(ylib, shapes_small):
    # light subpaths of the next pass on a second stream (default) or everything on one stream
    # (profile mode): with the per-texel camera kernel every sum has one order, so the films match bit
    # for bit over several passes that reuse both light-subpath sets
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 64, 40, max_depth=6, light_path_vertices=6)
    films = []
    for prof in (0, 1):
        it = ylib.Integrator(ylib.YPT_INTEGRATOR_BDPT, 0)
        it.set_option("bdpt_wavefront", 0)
        it.set_option("profile", prof)
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(0, 3)
        it.render(3, 2)
        films.append(it.film())
        it.close()
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))


@pytest.mark.parametrize("integrator", [1, 2])
def test_render_on_a_caller_stream(ylib, shapes_small, integrator):
    # passes rendered on a caller's stream, then on the library's stream, give the film of a handle
    # that used its own stream throughout; destroying the handle leaves the caller's stream usable
    import ctypes
    import torch
    from yapt_amd import camera
    scene, cam = shapes_small
    frame = camera.make_frame(cam, 64, 48, max_depth=8, light_path_vertices=6)

    def handle():
        it = ylib.Integrator(integrator, 0)
        if integrator == 2:
            it.set_option("bdpt_wavefront", 0)  # one fixed order of every sum: bit-comparable films
        it.set_scene(scene)
        it.set_frame(frame)
        return it

    a = handle()
    a.render(0, 6)
    want = a.film()
    a.close()
    caller = torch.cuda.Stream()
    b = handle()
    b.render(0, 2, stream=ctypes.c_void_p(caller.cuda_stream))
    b.render(2, 2)
    b.render(4, 2, stream=ctypes.c_void_p(caller.cuda_stream))
    got = b.film()
    b.close()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    with torch.cuda.stream(caller):
        x = torch.arange(1024, device="cuda", dtype=torch.float32) * 2.0
    caller.synchronize()
    assert float(x[-1]) == 2046.0


@pytest.mark.parametrize("kind,integrator,lights", [("longlat", 0, "env"), ("longlat", 1, "env"), ("cube", 1, "env"),
                                                    ("longlat", 2, "env"), ("cube", 2, "env"), ("cube", 1, "mixed"),
                                                    ("longlat", 2, "mixed")])
def test_render_parity_environment_texture(ylib, oracle, kind, integrator, lights):
    # escaping rays read the sky texture (long-lat with its wrap, a cube map by face); with no lights BDPT
    # starts light subpaths at the textured sky
    from yapt_amd import camera, scenes
    env = scenes.longlat_environment(64, 32) if kind == "longlat" else scenes.cube_environment(24)
    scene, cam, _ = scenes.punctual_lights(lights, sphere_res=(12, 12), environment=env)
    frame = camera.make_frame(cam, 72, 44, max_depth=6, env_intensity_scale=0.5, light_path_vertices=6)
    for opts in ([{"bdpt_wavefront": 1}, {"bdpt_wavefront": 0}] if integrator == 2 else [{}]):
        film = _compare_render(ylib, oracle, scene, frame, integrator, passes=2, options=opts)
        assert film[..., 1].max() > 0


def test_bdpt_parity_strategy_ranges(ylib, oracle, cornell):
    # the debug strategy filter (light vertices 1..3, camera vertices 1..4) in both camera-stage forms
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 64, 48, max_depth=8, bdpt_light_nodes=(1, 3), bdpt_camera_nodes=(1, 4))
    for opts in ({"bdpt_wavefront": 1}, {"bdpt_wavefront": 0}):
        film = _compare_render(ylib, oracle, scene, frame, ylib.YPT_INTEGRATOR_BDPT, passes=2, options=opts)
        assert film[..., 1].max() > 0
