"""The reference's published agreement figures (Tests/IntegratorCompare/README.md:58-109) on the
focused-emitter box (focused_emission_scene.py:51-77, yapt_amd.scenes.focused_emission), rendered
by the oracle at reduced size and taken through the FinalColor steps (ACES, manual exposure 1,
8-bit sRGB) of the NumPy checker, at the reference's path depth of 32:

* focus sweep (focus_sweep.py, NEE): capture means 56.9 / 64.0 / 67.5 / 65.2 for n = 0 / 2 / 8 / 32,
  the peak (p99) climbing from about 68 to 192 and the floor's edges falling below their n = 0 level
  while its centre rises: the (n + 2) / 2 normalisation redistributes light and keeps the power;
* integrator agreement: BPT and BPT_WITH_NEE within well under 1%, BDPT about 1% brighter.

The capture here is 64x48 rather than the reference's window (about 900x700), so the means carry a
little sampling noise and the tolerances below say how much."""
import numpy as np
import pytest

README_MEANS = {0.0: 56.9, 2.0: 64.0, 8.0: 67.5, 32.0: 65.2}


def _capture(oracle, focus, integrator, spp, begin=1000, size=(64, 48)):
    from oracle import post_ref as R
    from yapt_amd import camera, scenes
    scene, cam = scenes.focused_emission(focus)
    frame = camera.make_frame(cam, size[0], size[1], max_depth=32)
    film = oracle.render(scene, frame, integrator, begin, spp)["film"].astype(np.float64)
    return film, R.final_color(R.xyz_to_srgb(film), 1.0, True)[..., :3]


@pytest.fixture(scope="module")
def sweep(oracle):
    return {n: _capture(oracle, n, 1, 256)[1] for n in README_MEANS}


def test_focus_sweep_means_match_the_reference(sweep):
    for n, want in README_MEANS.items():
        got = float(sweep[n].mean())
        assert abs(got / want - 1) < 0.04, (n, got, want)


def test_focus_sweep_redistributes_rather_than_dims(sweep):
    from yapt_amd.compare import floor_profile
    p99 = {n: float(np.percentile(img, 99)) for n, img in sweep.items()}
    assert 55 < p99[0.0] < 80 and p99[32.0] > 2.3 * p99[0.0], p99           # README: 68 -> 192
    m0, m32 = sweep[0.0].mean(), sweep[32.0].mean()
    assert m32 > 0.9 * m0                                                    # without (n + 2) / 2: x 2/34
    f0, f32 = np.array(floor_profile(sweep[0.0])), np.array(floor_profile(sweep[32.0]))
    assert f32[5:7].mean() > 1.8 * f0[5:7].mean()                           # the centre climbs
    assert f32[[0, 11]].mean() < f0[[0, 11]].mean()                          # the edges fall below n = 0


def test_bpt_nee_and_bdpt_agree_as_the_reference_reports(oracle):
    y = {}
    for integ, spp in ((0, 512), (1, 256), (2, 64)):
        film, _ = _capture(oracle, 0.0, integ, spp, size=(48, 36))
        y[integ] = film[..., 1].mean()
    bpt_gap = 100 * (y[0] / y[1] - 1)
    bdpt_gap = 100 * (y[2] / y[1] - 1)
    assert abs(bpt_gap) < 1.2, bpt_gap          # README: 0.35% (rel|d| 16x); sampling noise here ~0.4%
    assert 0.0 < bdpt_gap < 2.0, bdpt_gap       # README: about 0.9% brighter
