"""Render server: the request generator on the CPU; on the GPU, results that do not depend on the
mix of jobs around them, admission by memory budget and active-job limit."""
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
def test_job_result_is_the_same_alone_and_in_the_mix(ylib, cornell, shapes_small):
    from yapt_amd import camera
    sc_a = ylib.ResidentScene(cornell[0])
    sc_b = ylib.ResidentScene(shapes_small[0])
    fr_a = camera.make_frame(cornell[1], 64, 48, max_depth=8)
    srv = ylib.Server(0, 8 << 30, 4)
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
    # bidirectional: the megakernel and the queued camera stage add visible connections in another order
    assert np.allclose(bdpt_a, bdpt_b, rtol=1e-5, atol=1e-7)
