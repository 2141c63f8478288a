"""Moved geometry without a BVH rebuild: ypt_refit_scene keeps the tree and refits its boxes to new
vertex data for the same triangles. Closest hits after a refit are the same bits as brute force on
the moved scene, renders match the checker on the moved scene, and the py_yapt surface refits when
only render-object transforms change."""
import numpy as np
import pytest

from test_gpu_parity import _rays

pytestmark = pytest.mark.gpu


def _moved(ylib, scene, shift=(0.75, -0.5, 2.25), k=0, angle=0.4):
    """the scene with instance k translated and instance k + 1 rotated about its centre (same triangles)"""
    pos = scene.positions.copy()
    nrm = scene.normals.copy()
    tan = None if scene.tangents is None else scene.tangents.copy()
    verts = lambda i: np.unique(scene.indices[scene.inst_first[i]:scene.inst_first[i] + scene.inst_count[i]])
    v0 = verts(k)
    pos[v0] += np.asarray(shift, np.float32)
    v1 = verts(k + 1)
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)
    ctr = pos[v1].mean(0)
    pos[v1] = (pos[v1] - ctr) @ R.T + ctr
    nrm[v1] = nrm[v1] @ R.T
    if tan is not None:
        tan[v1, :3] = tan[v1, :3] @ R.T
    return ylib.SceneArrays(pos, nrm, scene.indices, scene.inst_first, scene.inst_count, scene.inst_material,
                            list(scene.materials), tangents=tan, lights=scene.lights)


@pytest.mark.parametrize("width", [2, 8])
def test_refit_hits_are_brute_force_hits_of_the_moved_scene(ylib, oracle, shapes_small, width):
    scene = shapes_small[0]
    moved = _moved(ylib, scene)
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    it.set_option("bvh_width", width)
    it.set_scene(scene)
    rays = _rays(np.random.default_rng(31), 6000, -5.0, 5.0, moved.positions)
    it.refit_scene(moved)
    assert it.bvh_counts() == (1, 1)
    for wb in (0, 1 << 30):     # thread-per-ray and wave-per-ray traversal of the refitted tree
        it.set_option("wave_ray_below", wb)
        t, tri, b = it.trace_closest(rays)
        t2, tri2, b2 = oracle.trace_closest(moved, rays, use_bvh=False)
        assert (tri != 0xFFFFFFFF).mean() > 0.3
        assert np.array_equal(tri, tri2) and np.array_equal(t.view(np.uint32), t2.view(np.uint32))
        assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))
    it.close()


@pytest.mark.parametrize("integrator", [1, 2])
def test_refit_render_matches_the_checker_on_the_moved_scene(ylib, oracle, shapes_small, integrator):
    from yapt_amd import camera
    from test_gpu_parity import _compare_render
    scene, cam = shapes_small
    moved = _moved(ylib, scene)
    frame = camera.make_frame(cam, 48, 36, max_depth=8, light_path_vertices=6)
    it = ylib.Integrator(integrator, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 1)
    it.refit_scene(moved)      # clears the film: the next render starts over
    it.render(2, 2)
    g_last, g_film = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(moved, frame, integrator, 2, 2, use_bvh=True)
    assert np.array_equal(g_last.view(np.uint32), ref["last"].view(np.uint32))
    assert np.array_equal(g_film.view(np.uint32), ref["film"].view(np.uint32))


def test_refit_refuses_other_triangles_and_shared_scenes(ylib, shapes_small, cornell):
    scene = shapes_small[0]
    it = ylib.Integrator(1, 0)
    it.set_scene(scene)
    with pytest.raises(ylib.YptError):
        it.refit_scene(cornell[0])               # other triangles
    rs = ylib.ResidentScene(scene)
    it2 = ylib.Integrator(1, 0)
    it2.set_scene(rs)
    with pytest.raises(ylib.YptError):
        it2.refit_scene(_moved(ylib, scene))     # shared with the resident-scene handle
    it.close()
    it2.close()
    rs.close()


def test_py_yapt_refits_when_only_transforms_change(ylib, oracle):
    from yapt_amd import camera, py_yapt, scenes
    app = py_yapt.Application(width=48, height=32)
    renderer, scene = app.get_renderer(), app.get_scene()
    box = scenes.box_template()
    mesh = renderer.createMesh("box", box[0], box[2], normals=box[1])
    mat = renderer.createMaterial("m")
    mat.setFromMaterialPreset(py_yapt.Material.MaterialPreset.PLASTIC)
    lightm = renderer.createMaterial("l")
    lightm.setEmission(py_yapt.vec3([6.0, 6.0, 6.0]))
    objs = []
    for k, (m, t) in enumerate([(mat, (0, 0, 0)), (mat, (1.6, 0.2, -0.5)), (lightm, (0, 3.0, 0))]):
        o = scene.create_render_object("o%d" % k)
        o.setMesh(mesh)
        o.setMaterial(m, 0)
        o.getTransform().setTranslation(py_yapt.vec3(list(t)))
        objs.append(o)
    scene.get_main_camera().getTransform().lookAt(py_yapt.vec3([0, 2, 6]), py_yapt.vec3([0, 0.5, 0]), py_yapt.vec3([0, 1, 0]))
    renderer.getRendererVariable("RenderPipeline").setSelectedOption(1)
    renderer.render(1)
    objs[1].getTransform().setTranslation(py_yapt.vec3([-1.2, 0.4, 0.3]))   # a move: refit
    renderer.render(2)
    it = renderer._integrators[py_yapt.PIPELINES["BPT_WITH_NEE"]]
    assert it.bvh_counts() == (1, 1)
    film = renderer.film()
    ref = oracle.render(renderer._arrays, renderer.frame_desc(), 1, 0, 2, use_bvh=True)
    assert np.array_equal(film.view(np.uint32), ref["film"].view(np.uint32))
