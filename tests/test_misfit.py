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


def test_block_and_projection_files_round_trip(tmp_path):
    """MODELS_3D/block_* and projected gradient files (the layouts
    TOOLS/SOURCE/project_kernel.f90 reads and writes) round-trip."""
    from ses3d_amd import formats as F
    edges = [np.linspace(40.0, 60.0, 11), np.array([45.0, 47.5, 50.0])]
    F.write_block_file(str(tmp_path / "block_x"), edges)
    back = F.read_block_file(str(tmp_path / "block_x"))
    assert len(back) == 2 and all(np.array_equal(a, b) for a, b in zip(back, edges))
    vals = [np.array([1.5, -2.25e-30, 3.0e12], np.float32)]
    F.write_projected(str(tmp_path / "gradient_csh"), vals)
    got = F.read_projected(str(tmp_path / "gradient_csh"))
    assert np.array_equal(got[0], vals[0])


def test_problem_measurement_hook():
    """Problem(measure=fn) takes the event's misfit and adjoint rows from fn;
    without it, the L2 waveform misfit and its rows."""
    from ses3d_amd import configs
    from ses3d_amd.autograd import Problem
    w = configs.Workload("m", (2, 2, 2), False, 1, 0.5, 40, 3, "none")
    run = configs.build(w, seed=0)
    rng = np.random.default_rng(1)
    u, d = rng.normal(size=(3, 3, 40)), rng.normal(size=(3, 3, 40))
    rows = rng.normal(size=(3, 40, 3))
    calls = []

    def fn(ev, uu, dt):
        calls.append((ev, dt))
        return 2.5, rows

    p = Problem(run, [], measure=fn)
    chi, r = p._measure(0, u, True)
    assert chi == 2.5 and calls == [(0, 0.5)]
    assert r.dtype == np.float32 and r.flags.c_contiguous and np.array_equal(r, rows.astype(np.float32))
    assert p._measure(0, u, False)[1] is None
    q = Problem(run, [d], taper_width=5.0)
    chi, r = q._measure(0, u, True)
    d32 = d.astype(np.float32)          # observed data are held in float32
    assert chi == MF.l2_misfit(u, d32, 0.5, q.taper)
    assert np.array_equal(r, MF.l2_adjoint_rows(u, d32, 0.5, taper=q.taper))
