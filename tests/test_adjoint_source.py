"""Adjoint-source generation (ses3d_amd.adjoint_source) against what fixes it
independently of its own formulas.

The reference's tools (TOOLS/adjoint_source.py, TOOLS/seismograms.py,
TOOLS/prolate_spheroidal.py) are Python 2 and import obspy, which this image
lacks, so they cannot run here; they are pinned by
  * the reference's own example adjoint-source file (ADJOINT/1/ad_src_1),
    which the writer reproduces byte for byte from its values;
  * scipy's discrete prolate spheroidal sequences and their concentrations
    (scipy.signal.windows.dpss), through Slepian's identities: the sequences
    normalised to one over [-W, W] are v/sqrt(lambda) and the band-limited
    index-domain functions are lambda times them, so the multitaper spectral
    estimate is sum_k (sum v_k)(sum_n d v_k e^{-2 pi i n nu})/lambda_k;
  * the measurements themselves: each adjoint source a satisfies
    sum_n a[n] delta_u[n] dt = delta(measurement) (finite differences of the
    phase or log-amplitude of that estimate, of the log energy, and the delay
    of a shifted trace for the cross-correlation);
  * exact summation by parts for the velocity-seismogram convention;
  * hand-derived windows, including Python 2's rounding of halves.
"""
import os

import numpy as np
import pytest
from scipy.signal.windows import dpss

from ses3d_amd import adjoint_source as A, formats as F

REF_AD = "/root/reference/ADJOINT/1/ad_src_1"


def _seis(traces, dt, integrate):
    """A seismogram holding exactly ``traces`` as its (displacement or velocity) traces."""
    s = A.Seismogram.from_traces(traces, dt, rec=(50.36, 27.88, 0.0))
    s.trace_x, s.trace_y, s.trace_z = (np.array(v, np.float64) for v in traces)
    s.integrate = integrate
    return s


def _wavelet(t, t0, f=0.2, width=1.5):
    return np.exp(-((t - t0) / width) ** 2) * np.sin(2 * np.pi * f * (t - t0))


def _source(traces, dt, integrate=True):
    a = A.AdjointSource()
    a.fetch_seismogram(_seis(traces, dt, integrate))
    return a


# ---------------------------------------------------------------------------
# Python 2 numerics and the file format
# ---------------------------------------------------------------------------

def test_python2_rounding_and_str():
    assert [A.py2_round(x) for x in (2.5, 3.5, -2.5, 0.49999999999999994, 1.4999, 0.0)] == [3, 4, -3, 0, 1, 0]
    # Python 2.7's str(float): '%.12g', '.0' appended to integral digits
    cases = [(0.1, "0.1"), (1.0 / 3.0, "0.333333333333"), (1e20, "1e+20"), (123456789.0, "123456789.0"),
             (-1175.088271054321, "-1175.08827105"), (1e-5, "1e-05"), (0.0, "0.0"), (-0.0, "-0.0"),
             (1234567890123.0, "1.23456789012e+12"), (50.3600006, "50.3600006")]
    for v, s in cases:
        assert A.py2_str(v) == s, (v, A.py2_str(v), s)


@pytest.mark.skipif(not os.path.exists(REF_AD), reason="reference tree absent")
def test_writer_reproduces_the_reference_example_file(tmp_path):
    """ADJOINT/1/ad_src_1 as the reference's writer left it: reading its
    header and values and writing them again gives the same bytes."""
    text = open(REF_AD).read()
    lines = text.splitlines()
    rows = np.array([[float(v) for v in l.split()] for l in lines[4:]])
    assert rows.shape == (4000, 3)
    s = A.Seismogram.from_traces(np.zeros((3, 4000)), 0.13, rec=[float(v) for v in lines[2].split()])
    a = A.AdjointSource()
    a.fetch_seismogram(s)
    a.adsrc_x, a.adsrc_y, a.adsrc_z = rows[::-1].T.copy()
    a.write(str(tmp_path) + "/", "ad_src_1")
    assert (tmp_path / "ad_src_1").read_text() == text


def test_written_files_read_back_through_the_solver_readers(tmp_path):
    rng = np.random.default_rng(3)
    srcs = []
    for k in range(2):
        s = A.Seismogram.from_traces(rng.normal(size=(3, 300)), 0.1, rec=(40.0 + k, 20.5, 0.0))
        a = A.AdjointSource()
        a.fetch_seismogram(s)
        a.make_adsrc_cc("y", 5.0, 25.0, 2.0)
        a.make_adsrc_L2amp("z", 3.0, 20.0, 1.0)
        srcs.append(a)
    d = str(tmp_path / "ADJOINT" / "1")
    A.write_event_adjoint(d, srcs)
    locs = F.read_ad_srcfile(os.path.join(d, "ad_srcfile"))
    np.testing.assert_array_equal(locs[:, 2], [0.0, 0.0])
    np.testing.assert_array_equal(locs[:, 0], [F.deg2rad32(np.float32(40.0)), F.deg2rad32(np.float32(41.0))])
    for k, a in enumerate(srcs, 1):
        rows = F.read_ad_src(os.path.join(d, f"ad_src_{k}"), 300)
        # the file holds 12 significant digits; the solver reads them into float32
        want = np.array([[np.float32(float(A.py2_str(v))) for v in r] for r in a.forward_time()[:, ::-1].T])
        np.testing.assert_array_equal(rows, want)
        np.testing.assert_allclose(rows, a.rows(), rtol=1e-6)
        assert not rows[:, 0].any() and rows[:, 1].any() and rows[:, 2].any()


# ---------------------------------------------------------------------------
# seismogram reader
# ---------------------------------------------------------------------------

def test_seismogram_reader(tmp_path):
    rng = np.random.default_rng(0)
    tr = rng.normal(size=(3, 50)).astype(np.float32)
    F.write_seismogram(str(tmp_path), "ADVT_.KO.___", tr, 50, 0.13, (49.5668, 29.7383, 0.0), (50.74, 41.04, 5000.0))
    s = A.Seismogram()
    s.read(str(tmp_path) + "/", "ADVT_.KO.___")
    # numbers as the header prints them (9 significant digits of the float32 value)
    txt = lambda v: float(F._fortran_real(v))
    assert (s.nt, s.dt) == (50, txt(0.13))
    assert (s.rx, s.ry, s.rz) == (txt(49.5668), txt(29.7383), 0.0)
    assert (s.sx, s.sy, s.sz) == (txt(50.74), txt(41.04), 5000.0)
    want = np.vectorize(txt)(tr)
    np.testing.assert_array_equal(np.stack([s.trace_x, s.trace_y, s.trace_z]), want)
    assert np.array_equal(want.astype(np.float32), tr)
    # the time axis spans nt*dt over nt samples
    assert s.t[0] == 0.0 and s.t[-1] == 50 * s.dt and np.allclose(np.diff(s.t), 50 * s.dt / 49)
    d = A.Seismogram()
    d.read(str(tmp_path) + "/", "ADVT_.KO.___", integrate=True)
    np.testing.assert_array_equal(d.trace_z, np.cumsum(want[2]) * s.dt)
    assert d.integrate and not s.integrate


# ---------------------------------------------------------------------------
# windows
# ---------------------------------------------------------------------------

def test_window_by_hand():
    """nt = 11, dt = 1: t = 0, 1.1, ..., 11.  Window (2, 8) s with 2 s tapers:
    plateau on t = 2.2 .. 7.7 (indices 2..7); the samples with t <= 2 are
    t = 0 and 1.1, so the rising ramp covers indices 2, 3 and the falling one
    indices 6, 7 (ending before index 8)."""
    t = np.linspace(0, 11.0, 11)
    w = A.cosine_window(t, 1.0, 2.0, 8.0, 2.0)
    up, down = 0.5 * (1 - np.cos(np.pi * 1.1 / 2)), 0.5 * (1 + np.cos(np.pi * 1.1 / 2))
    np.testing.assert_allclose(w, [0, 0, 0, up, 1, 1, 1, down, 0, 0, 0], atol=1e-15)
    # win_left/dt = 2.5 rounds to 3 (Python 2), moving the rising ramp
    w = A.cosine_window(t, 1.0, 2.5, 8.0, 2.0)
    np.testing.assert_allclose(w, [0, 0, 0, 0, up, 1, 1, down, 0, 0, 0], atol=1e-15)
    with pytest.raises(ValueError):
        A.cosine_window(t, 1.0, 8.0, 12.0, 2.0)


# ---------------------------------------------------------------------------
# multitaper measurements
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("N", [25, 101, 401])
def test_prolate_sequences_against_scipy(N):
    W = 6.0 / N
    ev, c, psi, P0 = A.prolate_sequences(N, W)
    v, lam = dpss(N, N * W, 5, return_ratios=True)
    np.testing.assert_allclose(ev, lam, rtol=0, atol=1e-14)
    for k in range(5):
        sgn = np.sign(c[:, k] @ v[k])
        np.testing.assert_allclose(sgn * c[:, k], v[k] / np.sqrt(lam[k]), atol=1e-13)
        np.testing.assert_allclose(psi[:, k], lam[k] * c[:, k], atol=1e-13)
        assert np.isclose(P0[k], c[:, k].sum(), rtol=1e-14)


def _dpss_estimate(d, n0, N, nu_n):
    """The multitaper estimate from scipy's sequences."""
    v, lam = dpss(N, N * 6.0 / N, 5, return_ratios=True)
    n = np.arange(n0, n0 + N)
    e = np.exp(-2j * np.pi * n * nu_n)
    g = (v.sum(axis=1) / lam)[:, None] * v       # (5, N)
    return np.sum(g @ (d[n] * e)), (g.sum(axis=0)) * e


def _mt_case():
    dt, nt = 0.05, 1200
    t = np.arange(nt) * dt
    rng = np.random.default_rng(11)
    d = _wavelet(t, 30.0) + 0.05 * rng.normal(size=nt)
    return dt, nt, t, d, 20.0, 40.0, 0.2


def test_multitaper_sources_against_scipy_sequences():
    dt, nt, t, d, wl, wr, f = _mt_case()
    n0, N, nu_n = A.py2_round(wl / dt), 2 * A.py2_round((A.py2_round(wr / dt) - A.py2_round(wl / dt)) / 2.0) + 1, f * dt
    U, kern = _dpss_estimate(d, n0, N, nu_n)
    a = _source([d, d, d], dt)
    a.make_adsrc_mttime("x", wl, wr, f)
    a.make_adsrc_mtamp("y", wl, wr, f)
    want_t, want_a = np.zeros(nt), np.zeros(nt)
    want_t[n0:n0 + N] = np.imag(kern / (2 * np.pi * dt * nu_n * U))
    want_a[n0:n0 + N] = np.real(kern / (dt * U))
    scale_t, scale_a = np.abs(want_t).max(), np.abs(want_a).max()
    np.testing.assert_allclose(a.adsrc_x, want_t, rtol=0, atol=1e-10 * scale_t)
    np.testing.assert_allclose(a.adsrc_y, want_a, rtol=0, atol=1e-10 * scale_a)
    assert not a.adsrc_z.any()


def test_multitaper_sources_are_derivatives_of_phase_and_log_amplitude():
    dt, nt, t, d, wl, wr, f = _mt_case()
    n0, N, nu_n = 400, 401, f * dt
    a = _source([d, d, d], dt)
    a.make_adsrc_mttime("x", wl, wr, f)
    a.make_adsrc_mtamp("y", wl, wr, f)
    rng = np.random.default_rng(5)
    for _ in range(3):
        du = rng.normal(size=nt)
        eps = 1e-6
        Up, _ = _dpss_estimate(d + eps * du, n0, N, nu_n)
        Um, _ = _dpss_estimate(d - eps * du, n0, N, nu_n)
        dphase = np.angle(Up / Um) / (2 * eps)
        dlogamp = (np.log(abs(Up)) - np.log(abs(Um))) / (2 * eps)
        assert np.isclose(np.sum(a.adsrc_x * du) * dt, dphase / (2 * np.pi * nu_n), rtol=1e-6)
        assert np.isclose(np.sum(a.adsrc_y * du) * dt, dlogamp, rtol=1e-6)
    # units: a trace delayed by tau gives about -tau/dt (the phase in samples)
    tau = 0.01
    ds = _wavelet(t, 30.0 + tau) + (d - _wavelet(t, 30.0))
    got = np.sum(a.adsrc_x * (ds - d)) * dt
    assert -1.1 * tau / dt < got < -0.9 * tau / dt


def test_multitaper_velocity_convention():
    """On velocity seismograms the window's samples n0..n0+N-2 are
    differentiated, sample n0+N-1 keeps its value, sample n0+N is zero and
    the rest stays zero (TOOLS/adjoint_source.py:194-197)."""
    dt, nt, t, d, wl, wr, f = _mt_case()
    n0, N = 400, 401
    disp, vel = _source([d, d, d], dt, True), _source([d, d, d], dt, False)
    for a in (disp, vel):
        a.make_adsrc_mttime("x", wl, wr, f)
        a.make_adsrc_mtamp("z", wl, wr, f)
    for c in ("x", "z"):
        p, q = getattr(disp, "adsrc_" + c), getattr(vel, "adsrc_" + c)
        np.testing.assert_allclose(q[n0:n0 + N - 1], -np.diff(p[n0:n0 + N]) / dt, rtol=1e-12,
                                   atol=1e-12 * np.abs(q).max())
        assert q[n0 + N - 1] == p[n0 + N - 1] != 0.0
        assert not q[:n0].any() and not q[n0 + N:].any()


# ---------------------------------------------------------------------------
# L2 amplitude and cross-correlation
# ---------------------------------------------------------------------------

def test_l2amp_is_derivative_of_log_energy():
    """Inside the window's plateau, a = 2 d / (sum d^2 dt) is the gradient of
    ln(sum d^2 dt)."""
    dt, nt = 0.05, 1000
    t = np.linspace(0, nt * dt, nt)
    d = _wavelet(t, 25.0)
    a = _source([d, d, d], dt)
    a.make_adsrc_L2amp("z", 10.0, 40.0, 2.0)
    rng = np.random.default_rng(2)
    du = rng.normal(size=nt) * ((t > 13.0) & (t < 37.0))    # on the plateau only
    eps = 1e-6
    E = lambda x: np.log(np.sum(x * x) * dt)
    fd = (E(d + eps * du) - E(d - eps * du)) / (2 * eps)
    assert np.isclose(np.sum(a.adsrc_z * du) * dt, fd, rtol=1e-7)
    # 2 ln(1+eps) for a scaled trace, and 1/alpha for a trace alpha times larger
    assert np.isclose(np.sum(a.adsrc_z * 1e-6 * d) * dt, 2e-6, rtol=1e-9)
    b = _source([3 * d] * 3, dt)
    b.make_adsrc_L2amp("z", 10.0, 40.0, 2.0)
    np.testing.assert_allclose(b.adsrc_z, a.adsrc_z / 3, rtol=1e-13)


@pytest.mark.parametrize("tau", [0.005, -0.003])
def test_cc_gives_the_delay_of_a_shifted_trace(tau):
    dt, nt = 0.01, 6000
    t = np.linspace(0, nt * dt, nt)
    d = _wavelet(t, 30.0)
    a = _source([d, d, d], dt)
    a.make_adsrc_cc("x", 20.0, 40.0, 2.0)
    ds = _wavelet(t, 30.0 + tau)
    assert np.isclose(np.sum(a.adsrc_x * (ds - d)) * dt, tau, rtol=0.02)


@pytest.mark.parametrize("kind", ["L2amp", "cc"])
def test_velocity_convention_is_summation_by_parts(kind):
    """With f the displacement-form source of a measurement on the velocity
    trace and a its differentiated form, sum a du dt = sum f dv dt exactly for
    dv[n] = (du[n] - du[n-1])/dt, du[-1] = 0, when f ends at zero."""
    dt, nt = 0.05, 800
    t = np.linspace(0, nt * dt, nt)
    v = _wavelet(t, 20.0)
    f, a = _source([v] * 3, dt, True), _source([v] * 3, dt, False)
    for s in (f, a):
        getattr(s, "make_adsrc_" + kind)("y", 10.0, 30.0, 2.0)
    assert f.adsrc_y[-1] == 0.0 and a.adsrc_y[-1] == 0.0
    du = np.random.default_rng(9).normal(size=nt)
    dv = np.diff(np.concatenate([[0.0], du])) / dt
    lhs, rhs = np.sum(a.adsrc_y * du) * dt, np.sum(f.adsrc_y * dv) * dt
    assert np.isclose(lhs, rhs, rtol=1e-10)


def test_bad_component_and_window_raise():
    a = _source([np.ones(100)] * 3, 0.1)
    with pytest.raises(ValueError):
        a.make_adsrc_L2amp("r", 1.0, 5.0, 0.5)
    with pytest.raises(ValueError):
        a.make_adsrc_mttime("x", 5.0, 20.0, 0.5)          # window past the record
