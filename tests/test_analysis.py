"""Analysis metrics against closed forms: an ellipsoid mesh recovers its
semi-axes (so D = (a - c)/(a + c)); a sphere has D = 0; a linear velocity
field gives a linear binned profile; uniform random vertices give g(r) = 1."""
import numpy as np
import pytest

from gpucell_amd import analysis, casegen


@pytest.mark.parametrize("axes", [(1.0, 1.0, 1.0), (3.0, 2.0, 1.0), (5.0, 1.5, 1.5)])
def test_ellipsoid_axes_and_deformation_index(axes):
    V, F = casegen.icosphere(5)
    X = V * np.array(axes) + np.array([3.0, -2.0, 7.0])
    s = analysis.ellipsoid_axes(X, F)
    assert np.allclose(s, sorted(axes, reverse=True), rtol=2e-3)
    a, c = max(axes), min(axes)
    assert analysis.deformation_index(X, F) == pytest.approx((a - c) / (a + c), abs=2e-3)


def test_rotation_invariance_of_deformation_index():
    V, F = casegen.icosphere(4)
    X = V * np.array([2.0, 1.2, 0.8])
    R = casegen.random_rotation(np.random.default_rng(3))
    assert analysis.deformation_index(X @ R.T, F) == pytest.approx(
        analysis.deformation_index(X, F), abs=1e-12)


def test_velocity_profile_linear_field():
    rng = np.random.default_rng(0)
    pos = rng.uniform(-5, 5, size=(200000, 3))
    vel = np.zeros_like(pos)
    vel[:, 0] = 0.3 * pos[:, 1] + 1.0
    c, m, n = analysis.velocity_profile(pos, vel, axis=1, comp=0, bins=10, lo=-5, hi=5)
    assert np.allclose(m, 0.3 * c + 1.0, atol=5e-3)
    assert n.sum() == pos.shape[0]


def test_rdf_of_uniform_points_is_one():
    rng = np.random.default_rng(1)
    L = np.array([20.0, 20.0, 20.0])
    pos = rng.uniform(0, 20, size=(20000, 3))
    cell = rng.integers(0, 50, size=20000)
    r, g = analysis.membrane_rdf(pos, cell, L, (1, 1, 1), rmax=3.0, bins=6)
    assert np.allclose(g[1:], 1.0, atol=0.05)
