"""ypt_trace_closest_noalloc (device buffers, traversal-stack overflow in the library's own device
scratch, no allocation) against ypt_trace_closest on the same rays: bit-exact hits, two calls in a row
on one handle, the second smaller, with launches split into several chunks of the scratch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(rng, n, scene_pos):
    o = rng.uniform(-5.0, 5.0, size=(n, 3)).astype(np.float32)
    d = scene_pos[rng.integers(0, len(scene_pos), size=n)] - o
    d[n // 3:] = rng.normal(size=(n - n // 3, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = 5e-4
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


@pytest.mark.parametrize("width", [2, 8])
def test_noalloc_trace_matches_the_allocating_entry_point(ylib, shapes_small, width):
    import torch
    lib = ylib.load()
    P = ctypes.c_void_p
    lib.ypt_trace_closest_noalloc.argtypes = [P, P, ctypes.c_uint32, P, P, P, P]
    scene = shapes_small[0]
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    it.set_option("bvh_width", width)
    it.set_option("lds_stack_depth", 1)   # a deep overflow stack: several launches per call
    it.set_scene(scene)
    rng = np.random.default_rng(17)
    n_max = 300000
    d_rays = torch.empty((n_max, 8), dtype=torch.float32, device="cuda")
    d_t = torch.full((n_max,), -7.0, dtype=torch.float32, device="cuda")
    d_tri = torch.full((n_max,), 12345, dtype=torch.int32, device="cuda")
    d_b = torch.full((n_max, 2), -7.0, dtype=torch.float32, device="cuda")
    first = None
    for n in (n_max, 5000):   # two calls in a row, the second with a smaller input
        rays = _rays(rng, n, scene.positions)
        want_t, want_tri, want_b = it.trace_closest(rays)
        d_rays[:n].copy_(torch.from_numpy(rays))
        torch.cuda.synchronize()
        rc = lib.ypt_trace_closest_noalloc(it.h, P(d_rays.data_ptr()), n, P(d_t.data_ptr()), P(d_tri.data_ptr()),
                                           P(d_b.data_ptr()), None)
        assert rc == 0, lib.ypt_last_error()
        it.synchronize()
        t = d_t[:n].cpu().numpy()
        tri = d_tri[:n].cpu().numpy().view(np.uint32)
        b = d_b[:n].cpu().numpy()
        assert (want_tri != 0xFFFFFFFF).mean() > 0.3
        assert np.array_equal(tri, want_tri)
        assert np.array_equal(t.view(np.uint32), want_t.view(np.uint32))
        assert np.array_equal(b.view(np.uint32), want_b.view(np.uint32))
        if first is None:
            first = t.copy()
    # the second call wrote only its own rays' outputs
    assert np.array_equal(d_t[5000:].cpu().numpy().view(np.uint32), first[5000:].view(np.uint32))
    it.close()


def test_noalloc_trace_refuses_null_and_sceneless_calls(ylib):
    lib = ylib.load()
    P = ctypes.c_void_p
    lib.ypt_trace_closest_noalloc.argtypes = [P, P, ctypes.c_uint32, P, P, P, P]
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    p = P(256)
    assert lib.ypt_trace_closest_noalloc(it.h, p, 10, p, p, None, None) == -1
    assert lib.ypt_trace_closest_noalloc(it.h, p, 10, p, p, p, None) == -4     # no scene yet
    it.close()
