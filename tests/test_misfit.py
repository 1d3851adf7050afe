"""Misfit and adjoint-source helpers (CPU)."""
import numpy as np

from ses3d_amd import misfit as MF


def test_taper_vanishes_at_both_ends_and_is_one_inside():
    w = MF.cosine_taper(401, 0.25, 20.0)
    assert w[0] == 0.0 and abs(w[-1]) < 1e-15
    assert np.all(w[100:300] == 1.0)
    assert np.all(np.diff(w[:81]) >= 0)


def test_adjoint_rows_are_time_reversed_negative_derivative():
    nt, dt = 50, 0.5
    t = np.arange(nt) * dt
    u = np.zeros((3, 2, nt)); d = np.zeros((3, 2, nt))
    u[0, 1] = t ** 2
    rows = MF.l2_adjoint_rows(u, d, dt)
    assert rows.shape == (2, nt, 3)
    a = rows[1, ::-1, 0]                     # back to forward time
    np.testing.assert_allclose(a[:-1], -(2 * t[:-1] + dt), rtol=1e-6)   # -d/dt by forward difference
    assert a[-1] == 0.0
    assert np.all(rows[0] == 0) and np.all(rows[1, :, 1:] == 0)


def test_misfit_matches_definition():
    rng = np.random.default_rng(0)
    u, d = rng.normal(size=(3, 4, 30)), rng.normal(size=(3, 4, 30))
    assert np.isclose(MF.l2_misfit(u, d, 0.1), 0.5 * np.sum((u - d) ** 2) * 0.1)
    w = MF.cosine_taper(30, 0.1, 1.0)
    assert np.isclose(MF.l2_misfit(u, d, 0.1, w), 0.5 * np.sum(((u - d) * w) ** 2) * 0.1)
