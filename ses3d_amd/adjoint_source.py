"""Adjoint sources from SES3D seismograms, as the reference's tools make them.

The reference makes adjoint sources with two Python 2 tools: a seismogram
reader (TOOLS/seismograms.py:37-98) and an adjoint-source class
(TOOLS/adjoint_source.py) with four measurements on one component of one
seismogram, each restricted to a time window,

  make_adsrc_mttime  multitaper phase delay at one frequency    (:99-207)
  make_adsrc_mtamp   multitaper amplitude at one frequency      (:211-318)
  make_adsrc_L2amp   L2 amplitude of the windowed trace         (:323-394)
  make_adsrc_cc      cross-correlation time shift               (:400-471)

and a writer that stores the three components time-reversed under a short
header (:76-93), the file the adjoint solver reads
(SOURCE/ses3d_init.f90:475-492).  This module offers the same names,
arguments and results, in Python 3 with the reference's numerics:

* seismograms are velocity; ``integrate=True`` turns them into displacement
  by a running sum times dt (TOOLS/seismograms.py:87-90), and the time axis is
  ``linspace(0, nt*dt, nt)`` (spacing nt*dt/(nt-1), :83), which places the
  windows;
* every measurement is derived for displacement; on velocity seismograms the
  adjoint source is then differentiated, -(a[n+1]-a[n])/dt, so that it acts
  on displacement (adjoint_source.py:194-197, 305-308, 383-387, 460-462);
* window indices use Python 2's round (halves away from zero);
* the writer prints numbers as Python 2's ``str`` does (12 significant
  digits), so it writes the reference's bytes.

The multitaper measurements use the first five prolate spheroidal sequences
of the window (bandwidth W = 6/N), computed as TOOLS/prolate_spheroidal.py
does: eigenvectors of Slepian's tridiagonal matrix, normalised to one over
[-W, W], concentrations from the sinc kernel, index-domain sequences by the
sinc sum, sorted by concentration.

Where the reference only prints a message (a component other than x, y, z) or
fails inside numpy (a window outside the record), this module raises
``ValueError``.
"""
from __future__ import annotations

import math
import os
from typing import Tuple

import numpy as np
from scipy import linalg

from . import formats as F

K_MAX = 4          # multitaper: sequences 0..K_MAX (adjoint_source.py:132, :244)


def py2_round(x: float) -> int:
    """Python 2's round(x) to an integer: halves go away from zero."""
    a = abs(float(x))
    f = math.floor(a)
    r = f + 1.0 if a - f >= 0.5 else f
    return int(math.copysign(r, x)) if r else 0


def py2_str(v: float) -> str:
    """A float as Python 2's str() prints it: 12 significant digits, with
    '.0' added when the digits alone would read as an integer."""
    s = "%.12g" % float(v)
    if not any(ch in s for ch in ".eni"):
        s += ".0"
    return s


# ---------------------------------------------------------------------------
# seismograms
# ---------------------------------------------------------------------------

class Seismogram:
    """One station's three-component seismogram (TOOLS/seismograms.py:11-31)."""

    def __init__(self):
        self.nt = 10
        self.dt = 1.0
        self.t = np.linspace(0, 10, 10)
        self.rx = self.ry = self.rz = 0.0
        self.sx = self.sy = self.sz = 0.0
        self.trace_x = np.zeros(10)
        self.trace_y = np.zeros(10)
        self.trace_z = np.zeros(10)
        self.integrate = False

    def _set(self, traces, nt: int, dt: float, integrate: bool) -> None:
        self.nt, self.dt, self.integrate = int(nt), float(dt), bool(integrate)
        tr = np.array(traces, np.float64).reshape(3, -1)[:, :self.nt]
        if tr.shape[1] != self.nt:
            raise ValueError(f"traces hold {tr.shape[1]} samples, nt is {self.nt}")
        self.t = np.linspace(0, self.nt * self.dt, self.nt)
        if self.integrate:
            # running sum in sample order, times dt
            tr = np.cumsum(tr, axis=1) * self.dt
        self.trace_x, self.trace_y, self.trace_z = tr[0].copy(), tr[1].copy(), tr[2].copy()

    def read(self, directory: str, station_name: str, integrate: bool = False) -> None:
        """Read ``directory + station_name + '.x/.y/.z'`` (the text files the
        solver writes, SOURCE/ses3d_output.f90:398-451).  ``integrate`` turns
        the velocity traces into displacement."""
        base = directory + station_name
        head, traces = None, []
        for ext in "xyz":
            with open(base + "." + ext) as fh:
                lines = fh.read().splitlines()
            if head is None:
                head = lines
            nt = int(head[1].strip().split("=")[1])
            traces.append([float(v) for v in lines[7:7 + nt]])
        nt = int(head[1].strip().split("=")[1])
        dt = float(head[2].strip().split("=")[1])
        rec = head[4].strip().split("=")
        src = head[6].strip().split("=")
        self.rx, self.ry, self.rz = float(rec[1].split("y")[0]), float(rec[2].split("z")[0]), float(rec[3])
        self.sx, self.sy, self.sz = float(src[1].split("y")[0]), float(src[2].split("z")[0]), float(src[3])
        self._set(traces, nt, dt, integrate)

    @classmethod
    def from_traces(cls, traces, dt: float, rec=(0.0, 0.0, 0.0), src=(0.0, 0.0, 0.0),
                    integrate: bool = False) -> "Seismogram":
        """A seismogram from velocity traces (3, nt) held in memory, e.g. one
        receiver of a solver run; ``rec``/``src`` are (colat deg, lon deg, depth m)."""
        s = cls()
        s.rx, s.ry, s.rz = (float(v) for v in rec)
        s.sx, s.sy, s.sz = (float(v) for v in src)
        tr = np.asarray(traces, np.float64)
        s._set(tr, tr.shape[-1], dt, integrate)
        return s

    def trace(self, component: str) -> np.ndarray:
        if component not in ("x", "y", "z"):
            raise ValueError("component must be 'x', 'y' or 'z' (theta, phi or r)")
        return getattr(self, "trace_" + component)


# ---------------------------------------------------------------------------
# windows and prolate spheroidal sequences
# ---------------------------------------------------------------------------

def cosine_window(t: np.ndarray, dt: float, win_left: float, win_right: float, taper_width: float) -> np.ndarray:
    """The window of the L2-amplitude and cross-correlation measurements
    (adjoint_source.py:344-360): one for win_left < t < win_right, then a
    rising half-cosine over the samples with t <= taper_width written from
    index round(win_left/dt) on, and the falling one written to end just
    before index round(win_right/dt)."""
    nt = len(t)
    n0 = py2_round(win_left / dt)
    n = py2_round(win_right / dt) - n0
    w = np.zeros(nt)
    w[(t > win_left) & (t < win_right)] = 1.0
    tr = t[t <= taper_width]
    m = len(tr)
    lo, hi = n0 + n - m, n0 + n
    if n0 < 0 or n0 + m > nt or lo < 0 or hi > nt:
        raise ValueError(f"window [{win_left}, {win_right}] s with {taper_width} s tapers does not fit "
                         f"a record of {nt} samples")
    w[n0:n0 + m] = 0.5 * (1.0 - np.cos(np.pi * tr / taper_width))
    w[lo:hi] = 0.5 * (1.0 + np.cos(np.pi * tr / taper_width))
    return w


def slepian_matrix(N: int, W: float) -> Tuple[np.ndarray, np.ndarray]:
    """Diagonal and off-diagonal of Slepian's symmetric tridiagonal matrix
    whose eigenvectors are the prolate spheroidal sequences of length N and
    bandwidth W (TOOLS/prolate_spheroidal.py:9-38)."""
    i = np.arange(N, dtype=np.float64)
    diag = ((0.5 * (N - 1) - i) ** 2) * np.cos(2 * np.pi * W)
    off = 0.5 * (i[:-1] + 1) * (N - i[:-1] - 1)
    return diag, off


def sinc_kernel(N: int, W: float) -> np.ndarray:
    """2W sinc(2W(n-m)) for n, m in 0..N-1: the band-limiting operator whose
    eigenvalues are the sequences' concentrations in [-W, W]."""
    k = np.arange(N, dtype=np.float64)
    return 2.0 * W * np.sinc(2.0 * W * (k[:, None] - k[None, :]))


def prolate_sequences(N: int, W: float, k_max: int = K_MAX):
    """-> (eigval[k_max+1], c[N, k_max+1], psi[N, k_max+1], Psi0[k_max+1]).

    c are the sequences normalised to one over [-W, W] (c^T K c = 1 with K
    the sinc kernel), eigval their concentrations |K c|^2 / c^T K c, sorted in
    descending order (TOOLS/prolate_spheroidal.py:42-82); psi = K c are the
    index-domain functions on the window's N samples
    (prolate_spheroidal.py:117-144 on n = -(N-1)/2 .. (N-1)/2), Psi0 the
    frequency-domain functions at zero frequency, sum_n c[n]
    (prolate_spheroidal.py:86-113 at nu = 0)."""
    if k_max >= N:
        raise ValueError(f"k_max={k_max} needs a window of more than {k_max} samples")
    diag, off = slepian_matrix(N, W)
    # the most concentrated sequences belong to the matrix's largest
    # eigenvalues; the concentration sort runs over a margin of candidates
    m = min(N, 4 * (k_max + 1) + 8)
    if N > 1:
        _, c = linalg.eigh_tridiagonal(diag, off, select="i", select_range=(N - m, N - 1))
    else:
        c = np.ones((1, 1))
    K = sinc_kernel(N, W)
    Kc = K @ c
    ctkc = np.einsum("nk,nk->k", c, Kc)
    eigval = np.einsum("nk,nk->k", Kc, Kc) / ctkc
    order = np.argsort(-eigval, kind="stable")[:k_max + 1]
    scale = np.sqrt(ctkc[order])
    c = c[:, order] / scale
    psi = Kc[:, order] / scale
    return eigval[order], c, psi, c.sum(axis=0)


def _multitaper(d: np.ndarray, dt: float, win_left: float, win_right: float, nu: float):
    """Shared part of the multitaper measurements (adjoint_source.py:122-187):
    -> (n0, N, u, a) with u the spectral estimate of the windowed trace at
    frequency nu and a[n0:n0+N] the complex adjoint kernel before its
    measurement-specific scaling."""
    nt = len(d)
    n0 = py2_round(win_left / dt)
    N = int(2 * py2_round((py2_round(win_right / dt) - n0) / 2.0) + 1)
    if n0 < 0 or n0 + N > nt:
        raise ValueError(f"window [{win_left}, {win_right}] s does not fit a record of {nt} samples")
    nu_n = nu * dt                       # frequency in cycles per sample
    W = 3.0 * K_MAX / (2.0 * N)
    eigval, _, psi, Psi0 = prolate_sequences(N, W, K_MAX)
    n = np.arange(n0, n0 + N)
    e = np.exp(-2j * np.pi * n * nu_n)
    eta = (d[n] * e) @ psi
    wgt = Psi0 / eigval
    u = np.sum(eta * wgt)
    a = np.zeros(nt, complex)
    a[n] = (psi @ wgt) * e
    return n0, N, nu_n, u, a


def _velocity_window(a: np.ndarray, n0: int, N: int, dt: float) -> None:
    """On velocity seismograms the multitaper adjoint sources are
    differentiated inside the window: samples n0..n0+N-2 become
    -(a[n+1]-a[n])/dt, sample n0+N-1 keeps its value, sample n0+N is zeroed
    (adjoint_source.py:194-197, 305-308)."""
    if n0 + N >= len(a):
        raise ValueError("on velocity seismograms the multitaper window must end before the last sample")
    a[n0:n0 + N - 1] = -np.diff(a[n0:n0 + N]) / dt
    a[n0 + N] = 0.0


def _velocity(a: np.ndarray, dt: float) -> np.ndarray:
    """-(a[n+1]-a[n])/dt with the last sample zero (adjoint_source.py:383-387)."""
    out = np.empty_like(a)
    out[:-1] = -np.diff(a) / dt
    out[-1] = 0.0
    return out


# ---------------------------------------------------------------------------
# adjoint sources
# ---------------------------------------------------------------------------

class AdjointSource:
    """Adjoint source of one receiver (TOOLS/adjoint_source.py:11-27): fetch
    a seismogram, make one measurement's adjoint source per component, write
    the file the adjoint solver reads."""

    def __init__(self):
        self.s = Seismogram()
        self.adsrc_x = np.zeros(10)
        self.adsrc_y = np.zeros(10)
        self.adsrc_z = np.zeros(10)

    def fetch_seismogram(self, seismogram: Seismogram) -> None:
        self.s = seismogram
        self.adsrc_x = np.zeros(seismogram.nt)
        self.adsrc_y = np.zeros(seismogram.nt)
        self.adsrc_z = np.zeros(seismogram.nt)

    def _store(self, component: str, a: np.ndarray) -> None:
        setattr(self, "adsrc_" + component, a)

    def make_adsrc_mttime(self, component: str, win_left: float, win_right: float, nu: float) -> None:
        """Multitaper phase delay at frequency ``nu`` [Hz] in the window
        [win_left, win_right] [s]: a = Im(kernel / (2 pi dt nu' u)), nu' = nu dt,
        so that sum_n a[n] delta_u[n] dt = delta(arg u) / (2 pi nu')."""
        d = self.s.trace(component)
        n0, N, nu_n, u, a = _multitaper(d, self.s.dt, win_left, win_right, nu)
        a = a / (2 * np.pi * self.s.dt * nu_n * u)
        if not self.s.integrate:
            _velocity_window(a, n0, N, self.s.dt)
        self._store(component, np.imag(a))

    def make_adsrc_mtamp(self, component: str, win_left: float, win_right: float, nu: float) -> None:
        """Multitaper amplitude at frequency ``nu`` [Hz]: a = Re(kernel / (dt u)),
        so that sum_n a[n] delta_u[n] dt = delta ln|u|."""
        d = self.s.trace(component)
        n0, N, nu_n, u, a = _multitaper(d, self.s.dt, win_left, win_right, nu)
        a = a / (self.s.dt * u)
        if not self.s.integrate:
            _velocity_window(a, n0, N, self.s.dt)
        self._store(component, np.real(a))

    def make_adsrc_L2amp(self, component: str, win_left: float, win_right: float, taper_width: float) -> None:
        """L2 amplitude of the windowed trace w*d: a = 2 w d / (sum (w d)^2 dt)."""
        d = self.s.trace(component) * cosine_window(self.s.t, self.s.dt, win_left, win_right, taper_width)
        a = 2.0 * d / (np.sum(d * d) * self.s.dt)
        self._store(component, a if self.s.integrate else _velocity(a, self.s.dt))

    def make_adsrc_cc(self, component: str, win_left: float, win_right: float, taper_width: float) -> None:
        """Cross-correlation time shift: with g the forward difference of the
        windowed trace (last sample 0), a = -g / (sum g^2 dt), the sum taken
        in sample order, so that sum_n a[n] delta_u[n] dt is the delay of a
        slightly shifted trace."""
        d = self.s.trace(component) * cosine_window(self.s.t, self.s.dt, win_left, win_right, taper_width)
        g = np.empty_like(d)
        g[:-1] = np.diff(d) / self.s.dt
        g[-1] = 0.0
        a = -g / (np.cumsum(g * g)[-1] * self.s.dt)
        self._store(component, a if self.s.integrate else _velocity(a, self.s.dt))

    def forward_time(self) -> np.ndarray:
        """(3, nt) components x, y, z in forward time."""
        return np.stack([self.adsrc_x, self.adsrc_y, self.adsrc_z]).astype(np.float64)

    def rows(self) -> np.ndarray:
        """(nt, 3) float32 rows in the order the file stores them (time-reversed)."""
        return np.ascontiguousarray(self.forward_time()[:, ::-1].T, np.float32)

    def write(self, directory: str, filename: str) -> None:
        """Write ``directory + filename``: header, the receiver's coordinates
        as the source location, then the samples from last to first
        (adjoint_source.py:76-93)."""
        a = self.forward_time()
        with open(directory + filename, "w") as fh:
            fh.write("-- adjoint source ------------------\n")
            fh.write("-- source coordinates (colat,lon,depth)\n")
            fh.write(f"{py2_str(self.s.rx)} {py2_str(self.s.ry)} {py2_str(self.s.rz)}\n")
            fh.write("-- source time function (x, y, z) --\n")
            fh.write("".join(f"{py2_str(a[0, k])} {py2_str(a[1, k])} {py2_str(a[2, k])}\n"
                             for k in range(self.s.nt - 1, -1, -1)))


def write_event_adjoint(directory: str, sources) -> None:
    """ADJOINT/<event>/ files for a list of AdjointSource: ad_srcfile with the
    receivers' coordinates (SOURCE/ses3d_input.f90:592-611) and
    ad_src_1..ad_src_n in the same order."""
    os.makedirs(directory, exist_ok=True)
    F.write_ad_srcfile(os.path.join(directory, "ad_srcfile"),
                       np.array([[a.s.rx, a.s.ry, a.s.rz] for a in sources], np.float64))
    for k, a in enumerate(sources, 1):
        a.write(os.path.join(directory, ""), f"ad_src_{k}")


def event_rows(u: np.ndarray, dt: float, make, rec_deg=None) -> np.ndarray:
    """Adjoint-source rows (nr, nt, 3) float32 of one event, time-reversed as
    the solver takes them: for each receiver r, ``make(a, r)`` calls the
    make_adsrc_* methods of an AdjointSource that holds receiver r's velocity
    seismogram ``u[:, r]`` (u is (3, nr, nt), as the solver returns it).
    ``rec_deg`` (nr, 3), optional: the receivers' (colat, lon, depth)."""
    u = np.asarray(u, np.float64)
    out = np.zeros((u.shape[1], u.shape[2], 3), np.float32)
    for r in range(u.shape[1]):
        a = AdjointSource()
        a.fetch_seismogram(Seismogram.from_traces(u[:, r], dt, rec=rec_deg[r] if rec_deg is not None else (0, 0, 0)))
        make(a, r)
        out[r] = a.rows()
    return out


# the reference tools' class names (TOOLS/seismograms.py:11, TOOLS/adjoint_source.py:11)
ses3d_seismogram = Seismogram
adjoint_source = AdjointSource
