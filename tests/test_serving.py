"""Render server: the request generator on the CPU; on the GPU, results that do not depend on the
mix of jobs around them, admission by memory budget and active-job limit, rejection of a job over the
budget, cancellation and deadlines."""
import numpy as np
import pytest

from yapt_amd import serving


def test_requests_are_seeded_poisson_with_the_stated_mix():
    a = serving.make_requests(4000, rate=40.0, seed=3)
    b = serving.make_requests(4000, rate=40.0, seed=3)
    assert a == b
    t = np.array([r["t"] for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 40.0) < 0.05 / 40.0          # exponential gaps, mean 1/rate
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.06        # coefficient of variation 1
    integ = np.array([r["integrator"] for r in a])
    frac = [np.mean(integ == k) for k in (0, 1, 2)]
    assert abs(frac[0] - 0.25) < 0.03 and abs(frac[1] - 0.5) < 0.03 and abs(frac[2] - 0.25) < 0.03
    assert {r["spp"] for r in a} == {1, 4, 16, 64}
    assert {r["res"] for r in a} == set(serving.RESOLUTIONS)


def test_percentiles():
    p = serving.percentiles(np.arange(1, 101))
    assert abs(p["p50"] - 50.5) < 1e-9 and abs(p["p99"] - 99.01) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("fuse", [1, 0])
def test_job_result_is_the_same_alone_and_in_the_mix(ylib, cornell, shapes_small, fuse):
    from yapt_amd import camera
    sc_a = ylib.ResidentScene(cornell[0])
    sc_b = ylib.ResidentScene(shapes_small[0])
    fr_a = camera.make_frame(cornell[1], 64, 48, max_depth=8)
    srv = ylib.Server(0, 8 << 30, 4)
    srv.set_option("fuse_jobs", fuse)
    try:
        ja = srv.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8)
        srv.wait(ja)
        alone, _ = srv.take(ja)
        mix = []
        for k, (w, h, integ, spp) in enumerate([(96, 64, 2, 4), (128, 96, 0, 16), (80, 56, 1, 8), (64, 64, 2, 2)]):
            fr = camera.make_frame(shapes_small[1] if k % 2 else cornell[1], w, h, max_depth=8)
            mix.append(srv.submit(sc_b if k % 2 else sc_a, integ, fr, spp))
        jm = srv.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8)
        mix.append(srv.submit(sc_b, ylib.YPT_INTEGRATOR_BDPT, camera.make_frame(shapes_small[1], 72, 40, max_depth=6), 3))
        srv.wait(jm)
        in_mix, times = srv.take(jm)
        for j in mix:
            srv.wait(j)
            f, _ = srv.take(j)
            assert np.isfinite(f).all()
        assert times[0] <= times[1] <= times[2]
        assert np.array_equal(alone.view(np.uint32), in_mix.view(np.uint32))
        # and the same as a plain integrator handle rendering the same samples
        it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
        it.set_scene(sc_a)
        it.set_frame(fr_a)
        it.render(0, 8)
        direct = it.film()
        it.close()
        assert np.array_equal(direct.view(np.uint32), alone.view(np.uint32))
        assert srv.stats()["peak_active"] >= 2
    finally:
        srv.close()
        sc_a.close()
        sc_b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("max_slots", [16 << 20, 20000])
def test_fused_steps_give_each_job_its_own_film(ylib, cornell, shapes_small, max_slots):
    # BPT and NEE jobs over four scenes (an alpha-tested one, punctual lights with BPT shadow rays)
    # traced together in fused steps; every film equals a plain handle's render bit for bit. A small
    # slot limit splits the steps (a chunk larger than the limit runs alone).
    from yapt_amd import camera, scenes
    from test_textures import _alpha_render_scene
    alpha_scene, alpha_cam = _alpha_render_scene()
    punct, pcam, env = scenes.punctual_lights("mixed", sphere_res=(12, 12))
    scs = [(cornell[0], cornell[1], 0.0), (shapes_small[0], shapes_small[1], 0.0), (alpha_scene, alpha_cam, 0.0),
           (punct, pcam, env)]
    res = [ylib.ResidentScene(s) for s, _, _ in scs]
    spec = [(0, 64, 48, 1, 4, 8), (1, 96, 64, 0, 3, 12), (2, 48, 40, 1, 5, 5), (3, 72, 44, 0, 2, 6),
            (3, 40, 40, 1, 6, 6), (1, 128, 72, 1, 2, 12), (0, 32, 32, 0, 9, 8), (2, 80, 48, 0, 2, 5),
            (1, 56, 56, 1, 17, 10), (0, 100, 60, 1, 1, 3)]
    srv = ylib.Server(0, 16 << 30, 12)
    srv.set_option("fuse_max_slots", max_slots)
    try:
        frames = [camera.make_frame(scs[k][1], w, h, max_depth=d, env_intensity_scale=scs[k][2]) for k, w, h, _, _, d in spec]
        ids = [srv.submit(res[k], integ, frames[i], spp) for i, (k, w, h, integ, spp, d) in enumerate(spec)]
        films = []
        for j in ids:
            srv.wait(j)
            films.append(srv.take(j)[0])
        fs = srv.fused_stats()
        assert fs["steps"] > 0
        if max_slots > 1 << 20:
            assert fs["job_chunks"] > fs["steps"]      # steps carried several jobs
        for i, (k, w, h, integ, spp, d) in enumerate(spec):
            it = ylib.Integrator(integ, 0)
            it.set_scene(res[k])
            it.set_frame(frames[i])
            it.render(0, spp)
            direct = it.film()
            it.close()
            assert films[i][..., 1].max() > 0, i
            assert np.array_equal(films[i].view(np.uint32), direct.view(np.uint32)), (i, spec[i])
    finally:
        srv.close()
        for r in res:
            r.close()


@pytest.mark.gpu
def test_fuse_settings_are_validated(ylib):
    srv = ylib.Server(0, 1 << 30, 2)
    try:
        with pytest.raises(Exception):
            srv.set_option("fuse_max_slots", 0)
        srv.set_option("fuse_jobs", 0)
        assert srv.fused_stats() == {"steps": 0, "job_chunks": 0}
    finally:
        srv.close()


@pytest.mark.gpu
def test_admission_respects_the_active_limit(ylib, cornell):
    from yapt_amd import camera
    sc = ylib.ResidentScene(cornell[0])
    srv = ylib.Server(0, 4 << 30, 1)
    try:
        jobs = [srv.submit(sc, k % 3, camera.make_frame(cornell[1], 32 + 8 * k, 32, max_depth=4), 2) for k in range(6)]
        times = []
        for j in jobs:
            srv.wait(j)
            times.append(srv.take(j)[1])
        st = srv.stats()
        assert st["peak_active"] == 1 and st["completed"] == 6
        # one at a time, in arrival order: each job is admitted after the previous one finished
        for a, b in zip(times, times[1:]):
            assert b[1] >= a[2]
    finally:
        srv.close()
        sc.close()


@pytest.mark.gpu
def test_server_launch_options_change_no_result_bit(ylib, shapes_small):
    from yapt_amd import camera
    sc = ylib.ResidentScene(shapes_small[0])
    fr = camera.make_frame(shapes_small[1], 72, 40, max_depth=6)
    films = []
    for opts in ({}, {"bvh_width": 8, "persistent": 1, "persistent_waves": 8, "refill_threshold": 8,
                      "shade_reg_waves": 2, "bdpt_wavefront": 0, "bdpt_waves": 4}):
        srv = ylib.Server(0, 4 << 30, 2)
        try:
            for k, v in opts.items():
                srv.set_option(k, v)
            jobs = [srv.submit(sc, integ, fr, 3) for integ in (0, 1, 2)]
            out = []
            for j in jobs:
                srv.wait(j)
                out.append(srv.take(j)[0])
            films.append(out)
        finally:
            srv.close()
    sc.close()
    (bpt_a, nee_a, bdpt_a), (bpt_b, nee_b, bdpt_b) = films
    assert np.array_equal(bpt_a.view(np.uint32), bpt_b.view(np.uint32))
    assert np.array_equal(nee_a.view(np.uint32), nee_b.view(np.uint32))
    # bidirectional: the megakernel and the queued camera stage fold connections in the same order
    assert np.array_equal(bdpt_a.view(np.uint32), bdpt_b.view(np.uint32))


@pytest.mark.gpu
def test_job_over_the_memory_budget_is_rejected(ylib, cornell):
    from yapt_amd import camera
    sc = ylib.ResidentScene(cornell[0])
    srv = ylib.Server(0, 1 << 20, 2)
    try:
        with pytest.raises(ylib.YptError) as e:
            srv.submit(sc, ylib.YPT_INTEGRATOR_NEE, camera.make_frame(cornell[1], 256, 256, max_depth=4), 4)
        assert e.value.code == ylib.YPT_E_MEMORY
        assert srv.stats()["queued"] == 0
    finally:
        srv.close()
        sc.close()


def _alone(ylib, scene, integ, frame, spp):
    it = ylib.Integrator(integ, 0)
    try:
        it.set_scene(scene)
        it.set_frame(frame)
        it.render(0, spp)
        return it.film()
    finally:
        it.close()


@pytest.mark.gpu
def test_cancelled_jobs_end_and_leave_their_neighbours_unchanged(ylib, cornell, shapes_small):
    from yapt_amd import camera
    sc_a = ylib.ResidentScene(cornell[0])
    sc_b = ylib.ResidentScene(shapes_small[0])
    fr_long = camera.make_frame(cornell[1], 512, 512, max_depth=8)
    fr_a = camera.make_frame(cornell[1], 64, 48, max_depth=8)
    fr_b = camera.make_frame(shapes_small[1], 72, 40, max_depth=6)
    srv = ylib.Server(0, 8 << 30, 3)
    try:
        long_job = srv.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_long, 4096)   # seconds of work, in chunks
        ja = srv.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8)
        jb = srv.submit(sc_b, ylib.YPT_INTEGRATOR_BDPT, fr_b, 3)
        assert not srv.wait(long_job, 0.2)
        srv.cancel(long_job)
        with pytest.raises(ylib.YptError) as e:
            srv.wait(long_job)
        assert e.value.code == ylib.YPT_E_CANCELLED
        assert srv.forget(long_job) == ylib.YPT_E_CANCELLED
        # a job cancelled while it waits in the queue (one active slot, held by a long job)
        srv2 = ylib.Server(0, 8 << 30, 1)
        try:
            blocker = srv2.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_long, 4096)
            queued = srv2.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8)
            srv2.cancel(queued)
            with pytest.raises(ylib.YptError) as e:
                srv2.wait(queued)
            assert e.value.code == ylib.YPT_E_CANCELLED
            srv2.cancel(blocker)
            with pytest.raises(ylib.YptError):
                srv2.wait(blocker)
            after = srv2.submit(sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8)   # the pooled handle renders cleanly
            srv2.wait(after)
            film_after, _ = srv2.take(after)
        finally:
            srv2.close()
        for j, sc, integ, fr, spp in ((ja, sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8),
                                      (jb, sc_b, ylib.YPT_INTEGRATOR_BDPT, fr_b, 3)):
            srv.wait(j)
            film, _ = srv.take(j)
            assert np.array_equal(film.view(np.uint32), _alone(ylib, sc, integ, fr, spp).view(np.uint32))
        assert np.array_equal(film_after.view(np.uint32), _alone(ylib, sc_a, ylib.YPT_INTEGRATOR_NEE, fr_a, 8).view(np.uint32))
        # a cancelled job can be cancelled no more
        with pytest.raises(ylib.YptError):
            srv.cancel(long_job)
    finally:
        srv.close()
        sc_a.close()
        sc_b.close()


@pytest.mark.gpu
def test_deadline_stops_a_job_and_spares_one_that_meets_it(ylib, cornell):
    from yapt_amd import camera
    sc = ylib.ResidentScene(cornell[0])
    fr_long = camera.make_frame(cornell[1], 512, 512, max_depth=8)
    fr = camera.make_frame(cornell[1], 64, 48, max_depth=8)
    srv = ylib.Server(0, 8 << 30, 2)
    try:
        slow = srv.submit(sc, ylib.YPT_INTEGRATOR_NEE, fr_long, 4096)
        srv.set_deadline(slow, 0.3)
        quick = srv.submit(sc, ylib.YPT_INTEGRATOR_NEE, fr, 8)
        srv.set_deadline(quick, 60.0)
        with pytest.raises(ylib.YptError) as e:
            srv.wait(slow)
        assert e.value.code == ylib.YPT_E_TIMEOUT
        t_end = srv.clock()
        assert srv.forget(slow) == ylib.YPT_E_TIMEOUT
        srv.wait(quick)
        film, times = srv.take(quick)
        assert t_end - times[0] < 30.0
        assert np.array_equal(film.view(np.uint32), _alone(ylib, sc, ylib.YPT_INTEGRATOR_NEE, fr, 8).view(np.uint32))
        with pytest.raises(ylib.YptError):
            srv.set_deadline(slow, 1.0)     # forgotten
    finally:
        srv.close()
        sc.close()


@pytest.mark.gpu
def test_poisson_trace_over_resident_scenes_matches_jobs_rendered_alone(ylib):
    """A short seeded Poisson trace (12 requests at 40 per second) over resident scenes of ~10k, ~100k, ~1M and
    ~5M triangles: a sampled job of each integrator gives the same film bits as that job rendered alone."""
    from yapt_amd import camera, scenes as S
    shapes = serving.SCENE_SHAPES
    built, cams = [], []
    try:
        for k, n in enumerate(shapes):
            sc, cam = S.random_shapes(n_shapes=n, seed=2 + k, emissive_triangles=8)
            built.append(ylib.ResidentScene(sc))
            cams.append(cam)
        reqs = serving.make_requests(12, 40.0, seed=7, n_scenes=len(shapes))
        sample = {}
        for i, r in enumerate(reqs):       # the cheapest request of each integrator on the two largest scenes
            if r["scene"] >= 2:
                cost = r["res"][0] * r["res"][1] * r["spp"]
                if r["integrator"] not in sample or cost < sample[r["integrator"]][1]:
                    sample[r["integrator"]] = (i, cost)
        assert sample
        srv = ylib.Server(0, 96 << 30, 8)
        try:
            recs = serving.run(srv, built, cams, reqs, depth=8, keep_films={i for i, _ in sample.values()})
            peak = srv.stats()["peak_active"]
        finally:
            srv.close()
        assert all(r["ok"] for r in recs) and peak >= 2
        for i, _ in sample.values():
            r = reqs[i]
            fr = camera.make_frame(cams[r["scene"]], r["res"][0], r["res"][1], max_depth=8)
            alone = _alone(ylib, built[r["scene"]], r["integrator"], fr, r["spp"])
            assert np.array_equal(recs[i]["film"].view(np.uint32), alone.view(np.uint32)), r
    finally:
        for b in built:
            b.close()
