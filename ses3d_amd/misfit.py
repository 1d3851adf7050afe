"""L2 waveform misfit and its adjoint sources in SES3D's conventions.

The reference's Python tools make adjoint sources from velocity seismograms
(TOOLS/adjoint_source.py:323-394) and write them time-reversed
(TOOLS/adjoint_source.py:84-91); the adjoint solver reads the rows in file
order (SOURCE/ses3d_init.f90:486-490).  This module builds the same rows in
memory for the plain waveform misfit

    chi = 1/2 * sum_receivers sum_components sum_t (u(t) - d(t))^2 * dt

of velocity seismograms u (the reference samples seismograms as velocity,
SOURCE/ses3d_output.f90:485-487).
"""
from __future__ import annotations

import numpy as np


def cosine_taper(nt: int, dt: float, width: float) -> np.ndarray:
    """Window that rises and falls with half-cosines of ``width`` seconds at
    both ends of the record (the taper of TOOLS/adjoint_source.py:344-360)."""
    t = np.arange(nt) * dt
    T = (nt - 1) * dt
    w = np.ones(nt)
    if width > 0:
        a = t < width
        w[a] = 0.5 * (1.0 - np.cos(np.pi * t[a] / width))
        b = t > T - width
        w[b] = 0.5 * (1.0 - np.cos(np.pi * (T - t[b]) / width))
    return w


def l2_misfit(u: np.ndarray, d: np.ndarray, dt: float, taper: np.ndarray = None) -> float:
    """chi for seismograms (3, nr, nt), optionally windowed by ``taper`` (nt,)."""
    r = np.asarray(u, np.float64) - np.asarray(d, np.float64)
    if taper is not None:
        r = r * taper
    return 0.5 * float(np.sum(r * r)) * dt


def l2_adjoint_rows(u: np.ndarray, d: np.ndarray, dt: float, scale: float = 1.0,
                    taper: np.ndarray = None) -> np.ndarray:
    """Adjoint-source rows (nr, nt, 3) float32, time-reversed as stored in
    ADJOINT/<ev>/ad_src_<k>, for the residual u - d of velocity seismograms.

    The (windowed) residual w^2 (u - d) is differentiated in time (-d/dt,
    last sample 0) because the measurement is on velocity seismograms, the
    convention of TOOLS/adjoint_source.py:383-387.  A window that vanishes at
    the end of the record removes the boundary term of that integration by
    parts, which otherwise biases the gradient."""
    r = (np.asarray(u, np.float64) - np.asarray(d, np.float64)) * scale   # (3, nr, nt)
    if taper is not None:
        r = r * (taper * taper)
    a = np.zeros_like(r)
    a[..., :-1] = -np.diff(r, axis=-1) / dt
    rows = a[:, :, ::-1].transpose(1, 2, 0)
    return np.ascontiguousarray(rows, np.float32)
