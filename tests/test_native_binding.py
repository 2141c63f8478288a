"""The pybind11 binding of the C ABI (yapt_amd/_yapt_native): built in-tree, the record sizes it was
compiled with equal the ctypes mirror's, and on the GPU its integrator gives the same bits as the
ctypes path for films, closest hits and the FinalColor readback (py_yapt's renderer runs on it)."""
import ctypes

import numpy as np
import pytest


def test_native_module_loads_and_agrees_on_record_sizes(ylib):
    m = ylib.native()
    assert m.FRAME_DESC_BYTES == ctypes.sizeof(ylib.FrameDesc)
    assert m.MATERIAL_BYTES == ctypes.sizeof(ylib.Material)
    assert m.version() == ylib.load().ypt_version().decode()
    for k in (0, 1, 5, 1000):
        assert np.array_equal(m.sobol_point(k), ylib.sobol_point(k))
    for name in ("Integrator", "ResidentScene"):
        assert hasattr(m, name)


def test_native_scene_description_reads_the_scene_arrays(ylib, cornell):
    from yapt_amd import scenes
    m = ylib.native()
    scene = cornell[0]
    c = m.scene_desc_counts(scene)
    assert c["vertices"] == len(scene.positions) and c["triangles"] == scene.triangle_count
    assert c["instances"] == len(scene.inst_first) and c["materials"] == len(scene.materials)
    assert c["normals"] and c["tangents"] and not c["environment"] and c["textures"] == 0
    punct, _, _ = scenes.punctual_lights("mixed", sphere_res=(8, 8))
    assert m.scene_desc_counts(punct)["punctual_lights"] == len(punct.lights) > 0
    bad = ylib.SceneArrays(scene.positions.astype(np.float64), scene.normals, scene.indices, scene.inst_first, scene.inst_count,
                           scene.inst_material, list(scene.materials))
    bad.positions = scene.positions.astype(np.float64)   # not the ABI's element type
    with pytest.raises(ValueError):
        m.scene_desc_counts(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_native_integrator_matches_the_ctypes_path(ylib, cornell, integrator):
    from yapt_amd import camera
    scene, cam = cornell
    frame = camera.make_frame(cam, 48, 40, max_depth=6, light_path_vertices=6)
    m = ylib.native()
    out = []
    for make in (lambda: ylib.Integrator(integrator, 0), lambda: m.Integrator(integrator, 0)):
        it = make()
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(0, 3)
        rays = np.zeros((64, 8), np.float32)
        rays[:, :3] = (0.5, 0.5, 1.5)
        rays[:, 3] = 1e-4
        d = np.random.default_rng(1).normal(size=(64, 3)).astype(np.float32)
        rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
        rays[:, 7] = np.inf
        out.append((it.film(), it.last_samples(), it.final_color(), it.trace_closest(rays)))
        it.close()
    (f0, l0, c0, t0), (f1, l1, c1, t1) = out
    assert np.array_equal(f0.view(np.uint32), f1.view(np.uint32))
    assert np.array_equal(l0.view(np.uint32), l1.view(np.uint32))
    assert np.array_equal(c0, c1)
    for a, b in zip(t0, t1):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.gpu
def test_native_resident_scene_and_errors(ylib, shapes_small):
    from yapt_amd import camera
    scene, cam = shapes_small
    m = ylib.native()
    rs = m.ResidentScene(scene, 0)
    assert rs.triangle_count == scene.triangle_count and rs.device_bytes > 0
    it = m.Integrator(1, 0)
    it.set_scene(rs)
    it.set_frame(camera.make_frame(cam, 32, 24, max_depth=4))
    it.render(0, 2)
    ref = ylib.Integrator(1, 0)
    ref.set_scene(scene)
    ref.set_frame(camera.make_frame(cam, 32, 24, max_depth=4))
    ref.render(0, 2)
    assert np.array_equal(it.film().view(np.uint32), ref.film().view(np.uint32))
    with pytest.raises(RuntimeError):
        it.set_option("bvh_width", 3)
    it.close()
    with pytest.raises(RuntimeError):
        it.render(0, 1)
    ref.close()
    rs.close()
