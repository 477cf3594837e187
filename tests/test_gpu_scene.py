"""Device scene rasterisation (hz_rasterize_materials behind
physics.rasterize_materials) against the oracle's restatement of the
reference's sphere SDF, smoothing and blend, and against closed forms for
the box and ellipsoid distances.  FP64; sin is the only operation whose last
bit may differ from the host, so materials agree to 1e-15 relative."""
import numpy as np
import pytest

from oracle import materials as omat

pytestmark = pytest.mark.gpu


def test_sphere_materials_match_oracle():
    from ultrasound_solver_amd import physics
    shape, dx = (20, 18, 16), 1.07e-3
    center, radius = (10*dx, 9*dx, 7*dx), 5*dx
    rho, c = physics.rasterize_materials(shape, dx, [("sphere", center, radius)], 1.2, 343.0)
    r2, c2 = omat.sphere_materials(shape, dx, 1.2, 343.0, center, radius)
    np.testing.assert_allclose(rho, r2, rtol=1e-15, atol=0)
    np.testing.assert_allclose(c, c2, rtol=1e-15, atol=0)
    # End values: sin(-pi) rounds to -1.2e-16, not 0, so the obstacle's
    # density is 1250 to 2e-14 (the reference formula does the same).
    assert rho.max() == pytest.approx(1250.0, rel=1e-13) and rho.min() == pytest.approx(1.2, rel=1e-13)


def test_sharp_step_and_medium_fields():
    from ultrasound_solver_amd import physics
    shape, dx = (12, 11, 10), 1e-3
    rng = np.random.default_rng(4)
    cm = 343.0*(1 + 0.05*rng.standard_normal(shape))
    rm = 1.2*(1 + 0.1*rng.random(shape))
    rho, c, sdf = physics.rasterize_materials(shape, dx, [("sphere", (5e-3, 5e-3, 4e-3), 3e-3)], 1.2, 343.0,
                                              smoothing_cells=0, rho_medium=rm, c_medium=cm, want_sdf=True)
    h = (sdf > 0).astype(float)
    r2, c2 = omat.blend(h, 1250.0, 2200.0, rm, cm)
    np.testing.assert_array_equal(rho, r2)
    np.testing.assert_array_equal(c, c2)
    # No obstacle: the medium itself, bit for bit.
    rho0, c0 = physics.rasterize_materials(shape, dx, [], 1.2, 343.0, rho_medium=rm, c_medium=cm)
    np.testing.assert_array_equal(rho0, rm)
    np.testing.assert_array_equal(c0, cm)


def test_box_ellipsoid_and_union_closed_forms():
    from ultrasound_solver_amd import physics
    _, _, b = physics.rasterize_materials((11, 11, 11), 1.0, [("box", (5, 5, 5), (2, 3, 4))], 1.2, 343.0,
                                          want_sdf=True)
    assert b[5, 5, 5] == -2 and b[8, 5, 5] == 1 and b[5, 5, 10] == 1 and b[5, 9, 5] == 1
    assert b[9, 10, 5] == pytest.approx(np.sqrt(2.0**2 + 2.0**2), rel=1e-15)
    _, _, e = physics.rasterize_materials((21, 21, 21), 1.0, [("ellipsoid", (10, 10, 10), (5, 5, 5))], 1.2, 343.0,
                                          want_sdf=True)
    g = np.ogrid[:21, :21, :21]
    sph = np.sqrt(sum((g[a] - 10.0)**2 for a in range(3))) - 5
    np.testing.assert_allclose(e, sph, atol=1e-12)
    assert e[10, 10, 10] == -5
    prims = [("sphere", (4, 4, 4), 3.0), ("box", (12, 10, 8), (3, 2, 4)), ("ellipsoid", (6, 14, 12), (4, 2, 3))]
    _, _, u = physics.rasterize_materials((17, 18, 19), 1.0, prims, 1.2, 343.0, want_sdf=True)
    parts = [physics.rasterize_materials((17, 18, 19), 1.0, [p], 1.2, 343.0, want_sdf=True)[2] for p in prims]
    np.testing.assert_array_equal(u, np.minimum(np.minimum(parts[0], parts[1]), parts[2]))


@pytest.mark.parametrize("shape,corr", [((24, 20, 18), 5.0), ((64, 64, 64), 5.0), ((9, 30, 7), 2.5)])
def test_gaussian_random_field_matches_scipy(shape, corr):
    """The device filter against scipy.ndimage.gaussian_filter (mode "wrap",
    truncate 4) on the same seeded noise, including edges shorter than the
    41-tap kernel (several periodic wraps)."""
    from scipy.ndimage import gaussian_filter
    from ultrasound_solver_amd import scenes
    g = scenes.gaussian_random_field(shape, corr, 7)
    ref = gaussian_filter(np.random.default_rng(7).standard_normal(shape), corr, mode="wrap")
    ref = np.clip((ref - ref.mean())/ref.std(), -3.0, 3.0)
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-13)
