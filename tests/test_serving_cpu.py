"""Host side of the serving workload: the request mix generator."""
import numpy as np

from tinytrace_amd.serving import PASSES, RESOLUTIONS, orbit_camera, request_mix


def test_request_mix_distribution():
    specs = request_mix(4000, rate_per_s=10.0, scenes=4, seed=0)
    t = np.array([s.arrival_s for s in specs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 0.1) < 0.01 and abs(gaps.std() - 0.1) < 0.015  # exponential: std = mean
    res = np.array([RESOLUTIONS.index((s.width, s.height)) for s in specs])
    np.testing.assert_allclose(np.bincount(res, minlength=4) / len(specs), [0.4, 0.3, 0.2, 0.1], atol=0.03)
    passes = np.array([PASSES.index(s.passes) for s in specs])
    np.testing.assert_allclose(np.bincount(passes, minlength=4) / len(specs), [0.25] * 4, atol=0.03)
    assert set(s.scene for s in specs) == {0, 1, 2, 3}
    assert all(s.seed == s.scene for s in specs)


def test_orbit_cameras_look_at_the_centre():
    rng = np.random.default_rng(0)
    for _ in range(100):
        cam = orbit_camera(rng)
        p = np.array(cam.position)
        assert abs(np.linalg.norm(p - [0.0, 1.0, 0.0]) - 3.5) < 1e-4
        # The view direction is -forward() and points at the centre.
        d = -cam.forward()
        to_centre = (np.array([0.0, 1.0, 0.0]) - p) / 3.5
        assert np.dot(d, to_centre) > 0.999
