"""Earth models for SES3D run trees, generated as the reference's model tools do.

* :func:`generate` -- MODELS/SOURCE/generate_models.f90: the rank boxes of a
  px x py x pz decomposition of the section in ``INPUT/setup`` (:193-282),
  the GLL coordinates of each box (:296-330) and one of the 1D earth models
  of MODELS/MODELS_1D on them, chosen by ``model_type`` (:361-392):
  1 homogeneous, 2 isotropic PREM (prem_iso.f90), 3 homogeneous with the
  smooth QL6 Q profile (homogeneous_plus_Q.f90), 4 the European background
  model (eumod_bg.f90), 7 ak135 (ak135.f90).
* :func:`perturb` -- MODELS/SOURCE/add_perturbation.f90: shear, P-wave and
  density perturbations given on blocks (``MODELS/MODELS_3D``) added to a
  box's model, which is then re-expressed as rhoinv, mu, lambda and B.
* :func:`write_tree` / :func:`perturb_tree` -- the same on a run tree's
  files, as ``generate_models.exe`` and ``add_perturbation.exe`` run from
  ``MODELS/MAIN`` write them.

Every value is float32 and computed in the tools' operation order, so the
model files equal the reference tools' bit for bit.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import formats as F

f32 = np.float32
# pi of MODELS/SOURCE/modules_gm.f90:38 (the same float32 as SOURCE/ses3d_modules.f90:25)
PI = f32(3.14159265358979)

# GLL knots as the tools set them (generate_models.f90:124-190), single-precision literals
_KNOTS = {
    0: [0.0], 1: [-1.0, 1.0], 2: [-1.0, 0.0, 1.0],
    3: [-1.0, -0.4472135954999579, 0.4472135954999579, 1.0],
    4: [-1.0, -0.6546536707079772, 0.0, 0.6546536707079772, 1.0],
    5: [-1.0, -0.7650553239294647, -0.2852315164806451, 0.2852315164806451, 0.7650553239294647, 1.0],
    6: [-1.0, -0.8302238962785670, -0.4688487934707142, 0.0, 0.4688487934707142, 0.8302238962785670, 1.0],
    7: [-1.0, -0.8717401485096066, -0.5917001814331423, -0.2092992179024789, 0.2092992179024789,
        0.5917001814331423, 0.8717401485096066, 1.0],
}


def knots(lpd: int) -> np.ndarray:
    return np.array(_KNOTS[lpd], np.float32)


def _half(lpd: int) -> np.ndarray:
    """0.5*(1+knots(n)) in float32, the GLL node offsets in element widths."""
    return f32(0.5) * (f32(1.0) + knots(lpd))


# ---------------------------------------------------------------------------
# rank boxes and coordinates
# ---------------------------------------------------------------------------

@dataclass
class GridSpacing:
    """Element widths of the whole section (generate_models.f90:198-200)."""
    dx: np.float32
    dy: np.float32
    dz: np.float32


def boxes(setup: F.Setup) -> Tuple[List[F.Box], GridSpacing]:
    """The rank boxes of generate_models.f90:193-282 with float32 bounds."""
    px, py, pz = setup.px, setup.py, setup.pz
    nx, ny, nz = setup.nx_global, setup.ny_global, setup.nz_global
    x_min = f32(setup.theta_min) * PI / f32(180)
    x_max = f32(setup.theta_max) * PI / f32(180)
    y_min = f32(setup.phi_min) * PI / f32(180)
    y_max = f32(setup.phi_max) * PI / f32(180)
    z_min, z_max = f32(setup.z_min), f32(setup.z_max)
    g = GridSpacing((x_max - x_min) / f32(nx + px), (y_max - y_min) / f32(ny + py), (z_max - z_min) / f32(nz + pz))
    bnx, bny, bnz = (nx + 1) // px, (ny + 1) // py, (nz + 1) // pz
    out = []
    for k in range(1, pz + 1):
        for j in range(1, py + 1):
            for i in range(1, px + 1):
                mi = (k - 1) * py * px + (j - 1) * px + i
                lo = ((i - 1) * bnx, (j - 1) * bny, (k - 1) * bnz)
                hi = (i * bnx if i != px else nx, j * bny if j != py else ny, k * bnz if k != pz else nz)
                b = [x_min + f32(lo[0] + i - 1) * g.dx, x_min + f32(hi[0] + i) * g.dx,
                     y_min + f32(lo[1] + j - 1) * g.dy, y_min + f32(hi[1] + j) * g.dy,
                     z_min + f32(lo[2] + k - 1) * g.dz, z_min + f32(hi[2] + k) * g.dz]
                out.append(F.Box(mi, (i, j, k), lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                                 *[float(v) for v in b]))
    return out, g


def coordinates(b: F.Box, g: GridSpacing, lpd: int = 4) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """x(0:nx,0:lpd) colatitude [rad], y longitude [rad], z radius [m] of a
    box's GLL points (generate_models.f90:307-330; radius decreases with k)."""
    h = _half(lpd)
    i = np.arange(b.nx + 1, dtype=np.float32)[:, None]
    j = np.arange(b.ny + 1, dtype=np.float32)[:, None]
    k = np.arange(b.nz + 1, dtype=np.float32)[:, None]
    x = (f32(b.xmin) + i * g.dx) + h[None, :] * g.dx
    y = (f32(b.ymin) + j * g.dy) + h[None, :] * g.dy
    z = (f32(b.zmax) - k * g.dz) - h[None, :] * g.dz
    return x, y, z


# ---------------------------------------------------------------------------
# 1D models: rho [g/cm^3], cp, cs [km/s] and Q as functions of the radius z [m]
# ---------------------------------------------------------------------------

def _poly(r: np.ndarray, *terms) -> np.ndarray:
    """c0 (+|-) c1*r (+|-) c2*r*r ... left to right, each term coefficient
    first: terms are (sign, coefficient, power)."""
    _, c0, p0 = terms[0]
    acc = f32(c0) * np.ones_like(r)
    for _ in range(p0):
        acc = acc * r
    for sign, c, p in terms[1:]:
        t = f32(c) * np.ones_like(r)
        for _ in range(p):
            t = t * r
        acc = acc + t if sign > 0 else acc - t
    return acc


def powi(x: np.ndarray, n: int) -> np.ndarray:
    """x**n for a constant integer n >= 1 as the tools' build evaluates it: a
    left-to-right product x*x*...*x (for n = 4, ((x*x)*x)*x, which rounds
    differently from squaring twice)."""
    res = x
    for _ in range(n - 1):
        res = res * x
    return res


class _Bands:
    """An if / elseif chain over radius bands: the first band containing a
    point decides it; every point must be decided."""

    def __init__(self, z: np.ndarray):
        self.z = z
        self.done = np.zeros(z.shape, bool)

    def band(self, hi: Optional[float], lo: Optional[float]) -> np.ndarray:
        m = ~self.done
        if hi is not None:
            m &= self.z <= f32(hi)
        if lo is not None:
            m &= self.z >= f32(lo)
        self.done |= m
        return m

    def rest(self) -> np.ndarray:
        m = ~self.done
        self.done |= m
        return m

    def check(self, what: str) -> None:
        if not self.done.all():
            raise ValueError(f"{what}: radius outside the model's layers")


def _ql6_q(z: np.ndarray, is_diss: bool) -> np.ndarray:
    """Smooth QL6 shear Q (eumod_bg.f90:143-176, homogeneous_plus_Q.f90:51-81)."""
    if not is_diss:
        return np.full(z.shape, f32(100000.0), np.float32)
    r = (f32(6371.0) - z / f32(1000.0)) / f32(271.0)
    q = np.empty(z.shape, np.float32)
    B = _Bands(z)
    m = B.band(6371000.0, 6100000.0)
    p = (((f32(300.0) - f32(5370.82) * powi(r, 2)) + f32(14401.62) * powi(r, 3)) - f32(13365.78) * powi(r, 4)) \
        + f32(4199.98) * powi(r, 5)
    q[m] = p[m]
    q[B.band(6100000.0, 5701000.0)] = f32(165.0)
    q[B.band(5701000.0, 3480000.0)] = f32(355.0)
    q[B.band(3480000.0, 1221000.0)] = f32(0.0)
    q[B.rest()] = f32(104.0)
    return q


# isotropic PREM (prem_iso.f90:31-122): rho, cp, cs per band, polynomials in r = z/6371 km
_PREM_LOWER_MANTLE_RHO = ((+1, 7.9565, 0), (-1, 6.4761, 1), (+1, 5.5283, 2), (-1, 3.0807, 3))
_PREM = [
    (6371000.0, 6356000.0, ((+1, 2.60, 0),), ((+1, 5.80, 0),), ((+1, 3.20, 0),)),
    (6356000.0, 6346600.0, ((+1, 2.90, 0),), ((+1, 6.80, 0),), ((+1, 3.90, 0),)),
    (6346600.0, 6291000.0, ((+1, 2.6910, 0), (+1, 0.6924, 1)), ((+1, 4.1875, 0), (+1, 3.9382, 1)),
     ((+1, 2.1519, 0), (+1, 2.3481, 1))),
    (6291000.0, 6151000.0, ((+1, 2.6910, 0), (+1, 0.6924, 1)), ((+1, 4.1875, 0), (+1, 3.9382, 1)),
     ((+1, 2.1519, 0), (+1, 2.3481, 1))),
    (6151000.0, 5971000.0, ((+1, 7.1089, 0), (-1, 3.8045, 1)), ((+1, 20.3926, 0), (-1, 12.2569, 1)),
     ((+1, 8.9496, 0), (-1, 4.4597, 1))),
    (5971000.0, 5771000.0, ((+1, 11.2494, 0), (-1, 8.0298, 1)), ((+1, 39.7027, 0), (-1, 32.6166, 1)),
     ((+1, 22.3512, 0), (-1, 18.5856, 1))),
    (5771000.0, 5701000.0, ((+1, 5.3197, 0), (-1, 1.4836, 1)), ((+1, 19.0957, 0), (-1, 9.8672, 1)),
     ((+1, 9.9839, 0), (-1, 4.9324, 1))),
    (5701000.0, 5600000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 29.2766, 0), (-1, 23.6026, 1), (+1, 5.5242, 2), (-1, 2.5514, 3)),
     ((+1, 22.3459, 0), (-1, 17.2473, 1), (-1, 2.0834, 2), (+1, 0.9783, 3))),
    (5600000.0, 3630000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 24.9520, 0), (-1, 40.4673, 1), (+1, 51.4832, 2), (-1, 26.6419, 3)),
     ((+1, 11.1671, 0), (-1, 13.7818, 1), (+1, 17.4575, 2), (-1, 9.2777, 3))),
    (3630000.0, 3480000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 15.3891, 0), (-1, 5.3181, 1), (+1, 5.5242, 2), (-1, 2.5514, 3)),
     ((+1, 6.9254, 0), (+1, 1.4672, 1), (-1, 2.0834, 2), (+1, 0.9783, 3))),
    (3480000.0, 1221500.0, ((+1, 12.5815, 0), (-1, 1.2638, 1), (-1, 3.6426, 2), (-1, 5.5281, 3)),
     ((+1, 11.0487, 0), (-1, 4.0362, 1), (+1, 4.8023, 2), (-1, 13.5732, 3)), ((+1, 0.0, 0),)),
    (1221500.0, None, ((+1, 13.0885, 0), (-1, 8.8381, 2)), ((+1, 11.2622, 0), (-1, 6.3640, 2)),
     ((+1, 3.6678, 0), (-1, 4.4475, 2))),
]
_PREM_Q = (600.0, 600.0, 600.0, 80.0, 143.0, 143.0, 143.0, 312.0, 312.0, 312.0, 100000.0, 84.6)

# European background model (eumod_bg.f90:31-122): PREM-like bands with
# shifted velocities and a crust-mantle transition at 180 km
_EUMOD = [
    (6371000.0, 6356000.0, ((+1, 2.60, 0),), ((+1, 5.80, 0),), ((+1, 3.20, 0),)),
    (6356000.0, 6346600.0, ((+1, 2.90, 0),), ((+1, 6.80, 0),), ((+1, 3.90, 0),)),
    (6346600.0, 6291000.0, ((+1, 2.6910, 0), (+1, 0.6924, 1)), ((+1, 4.1875, 0), (+1, 3.9382, 1), (-1, 0.035, 0)),
     ((+1, 2.1519, 0), (+1, 2.3481, 1), (-1, 0.065, 0))),
    (6291000.0, 6191000.0, ((+1, 2.6910, 0), (+1, 0.6924, 1)), ((+1, 4.1875, 0), (+1, 3.9382, 1), (-1, 0.035, 0)),
     ((+1, 2.1519, 0), (+1, 2.3481, 1), (-1, 0.065, 0))),
    (6191000.0, 6051000.0, ((+1, 9.1790, 0), (-1, 5.9841, 1)), ((+1, 40.5988, 0), (-1, 33.5317, 1), (-1, 0.035, 0)),
     ((+1, 16.8261, 0), (-1, 12.7527, 1), (-1, 0.065, 0))),
    (6051000.0, 5971000.0, ((+1, 7.1089, 0), (-1, 3.8045, 1)), ((+1, 20.3926, 0), (-1, 12.2569, 1), (-1, 0.035, 0)),
     ((+1, 8.9496, 0), (-1, 4.4597, 1), (-1, 0.065, 0))),
    (5971000.0, 5771000.0, ((+1, 11.2494, 0), (-1, 8.0298, 1)), ((+1, 39.7027, 0), (-1, 32.6166, 1), (-1, 0.035, 0)),
     ((+1, 22.3512, 0), (-1, 18.5856, 1), (-1, 0.12, 0))),
    (5771000.0, 5701000.0, ((+1, 5.3197, 0), (-1, 1.4836, 1)), ((+1, 19.0957, 0), (-1, 9.8672, 1), (-1, 0.035, 0)),
     ((+1, 9.9839, 0), (-1, 4.9324, 1), (-1, 0.12, 0))),
    (5701000.0, 5600000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 29.2766, 0), (-1, 23.6026, 1), (+1, 5.5242, 2), (-1, 2.5514, 3)),
     ((+1, 22.3459, 0), (-1, 17.2473, 1), (-1, 2.0834, 2), (+1, 0.9783, 3), (-1, 0.12, 0))),
    (5600000.0, 3630000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 24.9520, 0), (-1, 40.4673, 1), (+1, 51.4832, 2), (-1, 26.6419, 3)),
     ((+1, 11.1671, 0), (-1, 13.7818, 1), (+1, 17.4575, 2), (-1, 9.2777, 3), (-1, 0.12, 0))),
    (3630000.0, 3480000.0, _PREM_LOWER_MANTLE_RHO,
     ((+1, 15.3891, 0), (-1, 5.3181, 1), (+1, 5.5242, 2), (-1, 2.5514, 3)),
     ((+1, 6.9254, 0), (+1, 1.4672, 1), (-1, 2.0834, 2), (+1, 0.9783, 3), (-1, 0.12, 0))),
    (3480000.0, 1221500.0, ((+1, 12.5815, 0), (-1, 1.2638, 1), (-1, 3.6426, 2), (-1, 5.5281, 3)),
     ((+1, 11.0487, 0), (-1, 4.0362, 1), (+1, 4.8023, 2), (-1, 13.5732, 3)), ((+1, 0.0, 0),)),
    (1221500.0, None, ((+1, 13.0885, 0), (-1, 8.8381, 2)), ((+1, 11.2622, 0), (-1, 6.3640, 2)),
     ((+1, 3.6678, 0), (-1, 4.4475, 2))),
]


def _banded(z: np.ndarray, table, what: str):
    r = z / f32(6371e3)
    rho, cp, cs = (np.empty(z.shape, np.float32) for _ in range(3))
    B = _Bands(z)
    for hi, lo, tr, tp, ts in table:
        m = B.band(hi, lo)
        rho[m] = _poly(r, *tr)[m]
        cp[m] = _poly(r, *tp)[m]
        cs[m] = _poly(r, *ts)[m]
    B.check(what)
    return rho, cp, cs


def _prem_iso(z, is_diss):
    rho, cp, cs = _banded(z, _PREM, "prem_iso")
    if is_diss:
        q = np.empty(z.shape, np.float32)
        B = _Bands(z)
        for (hi, lo, *_), qv in zip(_PREM, _PREM_Q):
            q[B.band(hi, lo)] = f32(qv)
    else:
        q = np.full(z.shape, f32(100000.0), np.float32)
    return rho, cp, cs, q


def _eumod_bg(z, is_diss):
    rho, cp, cs = _banded(z, _EUMOD, "eumod_bg")
    q = _ql6_q(z, is_diss)
    if z[0, 0] == f32(6371000.0):
        # the surface element layer gets upper-mantle values (eumod_bg.f90:182-191)
        rho[0, :], cp[0, :], cs[0, :] = f32(3.00), f32(7.20), f32(4.10)
    return rho, cp, cs, q


# ak135 (Kennett, Engdahl & Buland 1995, Geophys. J. Int. 122, 108-124), as
# sampled in ak135.f90:25-96: radius [km], vp [km/s], vs [km/s], rho [g/cm^3];
# repeated radii are the discontinuities
_AK135 = """
  6371.0   5.800000  3.460000   2.720000
  6351.0   5.800000  3.460000   2.720000
  6351.0   6.500000  3.850000   2.920000
  6336.0   6.500000  3.850000   2.920000
  6336.0   8.040000  4.480000   3.320000
  6293.5   8.045000  4.490000   3.345000
  6251.0   8.050000  4.500000   3.371000
  6251.0   8.050000  4.500000   3.371100
  6206.0   8.175000  4.509000   3.371100
  6161.0   8.300700  4.518400   3.324300
  6161.0   8.300700  4.518400   3.324300
  6111.0   8.482200  4.609400   3.366300
  6061.0   8.665000  4.696400   3.411000
  6011.0   8.847600  4.783200   3.457700
  5961.0   9.030200  4.870200   3.506800
  5961.0   9.360100  5.080600   3.931700
  5911.0   9.528000  5.186400   3.927300
  5861.0   9.696200  5.292200   3.923300
  5811.0   9.864000  5.398900   3.921800
  5761.0  10.032000  5.504700   3.920600
  5711.0  10.200000  5.610400   3.920100
  5711.0  10.790900  5.960700   4.238700
  5661.0  10.922200  6.089800   4.298600
  5611.0  11.055300  6.210000   4.356500
  5561.5  11.135500  6.242400   4.411800
  5512.0  11.222800  6.279900   4.465000
  5462.5  11.306800  6.316400   4.516200
  5413.0  11.389700  6.351900   4.565400
  5363.5  11.470400  6.386000   4.592600
  5314.0  11.549300  6.418200   4.619800
  5264.5  11.626500  6.451400   4.646700
  5215.0  11.702000  6.482200   4.673500
  5165.5  11.776800  6.513100   4.700100
  5116.0  11.849100  6.543100   4.726600
  5066.5  11.920800  6.572800   4.752800
  5017.0  11.989100  6.600900   4.779000
  4967.5  12.057100  6.628500   4.805000
  4918.0  12.124700  6.655400   4.830700
  4868.5  12.191200  6.681300   4.856200
  4819.0  12.255800  6.707000   4.881700
  4769.5  12.318100  6.732300   4.906900
  4720.0  12.381300  6.757900   4.932100
  4670.5  12.442700  6.782000   4.957000
  4621.0  12.503000  6.805600   4.981700
  4571.5  12.563800  6.828900   5.006200
  4522.0  12.622600  6.851700   5.030600
  4472.5  12.680700  6.874300   5.054800
  4423.0  12.738400  6.897200   5.078900
  4373.5  12.795600  6.919400   5.102700
  4324.0  12.852400  6.941600   5.126400
  4274.5  12.909300  6.962500   5.149900
  4225.0  12.966300  6.985200   5.173200
  4175.5  13.022600  7.006900   5.196300
  4126.0  13.078600  7.028600   5.219200
  4076.5  13.133700  7.050400   5.242000
  4027.0  13.189500  7.072200   5.264600
  3977.5  13.246500  7.093200   5.287000
  3928.0  13.301700  7.114400   5.309200
  3878.5  13.358400  7.136800   5.331300
  3829.0  13.415600  7.158400   5.353100
  3779.5  13.474100  7.180400   5.374800
  3731.0  13.531100  7.203100   5.396200
  3681.0  13.589900  7.225300   5.417600
  3631.0  13.649800  7.248500   5.438700
  3631.0  13.649800  7.248500   5.693400
  3581.3  13.653300  7.259300   5.719600
  3531.7  13.657000  7.270000   5.745800
  3479.5  13.660100  7.281700   5.772100
  3479.5   8.000000  0.000000   9.914500
  3431.6   8.038200  0.000000   9.994200
  3381.3   8.128300  0.000000  10.072200
  3331.0   8.221300  0.000000  10.148500
  3280.6   8.312200  0.000000  10.223300
  3230.3   8.400100  0.000000  10.296400
  3180.0   8.486100  0.000000  10.367900
  3129.7   8.569200  0.000000  10.437800
  3079.4   8.649600  0.000000  10.506200
  3029.0   8.728300  0.000000  10.573100
  2978.7   8.803600  0.000000  10.638500
  2928.4   8.876100  0.000000  10.702300
  2878.3   8.946100  0.000000  10.764700
  2827.7   9.013800  0.000000  10.825700
  2777.4   9.079200  0.000000  10.885200
  2727.0   9.142600  0.000000  10.943400
  2676.7   9.204200  0.000000  11.000100
  2626.4   9.263400  0.000000  11.055500
  2576.0   9.320500  0.000000  11.109500
  2525.7   9.376000  0.000000  11.162300
  2475.4   9.429700  0.000000  11.213700
  2425.0   9.481400  0.000000  11.263900
  2374.7   9.530600  0.000000  11.312700
  2324.4   9.577700  0.000000  11.360400
  2274.1   9.623200  0.000000  11.406900
  2223.7   9.667300  0.000000  11.452100
  2173.4   9.710000  0.000000  11.496200
  2123.1   9.751300  0.000000  11.539100
  2072.7   9.791400  0.000000  11.580900
  2022.4   9.830400  0.000000  11.621600
  1972.1   9.868200  0.000000  11.661200
  1921.7   9.905100  0.000000  11.699800
  1871.4   9.941000  0.000000  11.737300
  1821.1   9.976100  0.000000  11.773700
  1770.7  10.010300  0.000000  11.809200
  1720.4  10.043900  0.000000  11.843700
  1670.1  10.076800  0.000000  11.877200
  1619.8  10.109500  0.000000  11.909800
  1569.4  10.141500  0.000000  11.941400
  1519.1  10.173900  0.000000  11.972200
  1468.8  10.204900  0.000000  12.000100
  1418.4  10.232900  0.000000  12.031100
  1368.1  10.256500  0.000000  12.059300
  1317.8  10.274500  0.000000  12.086700
  1267.4  10.285400  0.000000  12.113300
  1217.5  10.289000  0.000000  12.139100
  1217.5  11.042700  3.504300  12.703700
  1166.4  11.058500  3.518700  12.728900
  1115.7  11.071800  3.531400  12.753000
  1064.9  11.085000  3.543500  12.776000
  1014.3  11.098300  3.555100  12.798000
   963.5  11.116600  3.566100  12.818800
  912.83  11.131600  3.576500  12.838700
  862.11  11.145700  3.586400  12.857400
  811.40  11.159000  3.595700  12.875100
  760.69  11.171500  3.604400  12.891700
  709.98  11.183200  3.612600  12.907200
  659.26  11.194100  3.620200  12.921700
  608.55  11.204100  3.627200  12.935100
  557.84  11.213400  3.633700  12.947400
  507.13  11.221900  3.639600  12.958600
  456.41  11.229500  3.645000  12.968800
  405.70  11.236400  3.649800  12.977900
  354.99  11.242400  3.654000  12.985900
  304.28  11.247700  3.657700  12.992900
  253.56  11.252100  3.660800  12.998800
  202.85  11.255700  3.663300  13.003600
  152.14  11.258600  3.665300  13.007400
  101.43  11.260600  3.666700  13.010000
   50.71  11.261800  3.667500  13.011700
     0.0  11.262200  3.667800  13.012200
"""


def _ak135_table():
    a = np.array([[float(t) for t in l.split()] for l in _AK135.strip().splitlines()])
    return tuple(f32(1000) * a[:, c].astype(np.float32) for c in range(4))   # m, m/s, m/s, kg/m^3


def _ak135(z):
    """Piecewise-linear in radius between the table's nodes (ak135.f90:101-126):
    every bracketing interval is visited in table order and the last one
    containing the point decides; a zero-width interval (a discontinuity)
    takes its upper node's values."""
    R, CP, CS, RHO = _ak135_table()
    cp, cs, rho = (np.full(z.shape, np.nan, np.float32) for _ in range(3))
    for n in range(len(R) - 1):
        m = (z >= R[n + 1]) & (z <= R[n])
        if not m.any():
            continue
        if R[n] - R[n + 1] == f32(0):
            cp[m], cs[m], rho[m] = CP[n], CS[n], RHO[n]
        else:
            ratio = (R[n] - z) / (R[n] - R[n + 1])
            one = f32(1) - ratio
            cp[m] = (CP[n] * one + CP[n + 1] * ratio)[m]
            cs[m] = (CS[n] * one + CS[n + 1] * ratio)[m]
            rho[m] = (RHO[n] * one + RHO[n + 1] * ratio)[m]
    if np.isnan(cp).any():
        raise ValueError("ak135: radius outside the table")
    return rho, cp, cs


MODEL_TYPES = {1: "homogeneous", 2: "prem_iso", 3: "homogeneous_plus_Q", 4: "eumod_bg", 7: "ak135"}


def model_1d(model_type: int, z: np.ndarray, is_diss: bool) -> Dict[str, np.ndarray]:
    """rhoinv, mu, lambda (A = B = C = 0) and, where the model defines it, Q
    on the radii z (nz+1, lpd+1), float32."""
    zero = np.zeros(z.shape, np.float32)
    if model_type in (1, 3):
        # homogeneous.f90 / homogeneous_plus_Q.f90: rho = 1, mu = lambda = 0
        out = {"rhoinv": np.full(z.shape, f32(1.0) / f32(1.0), np.float32), "mu": zero.copy(),
               "lambda": zero.copy()}
        if model_type == 3:
            out["Q"] = _ql6_q(z, is_diss)
        elif is_diss:
            raise ValueError("model_type 1 (homogeneous) defines no Q; the reference writes an unset array")
        return out
    if model_type == 2:
        rho, cp, cs, q = _prem_iso(z, is_diss)
    elif model_type == 4:
        rho, cp, cs, q = _eumod_bg(z, is_diss)
    elif model_type == 7:
        rho, cp, cs = _ak135(z)
        q = np.full(z.shape, f32(100000.0), np.float32)
    else:
        raise ValueError(f"model_type {model_type} is not one of {sorted(MODEL_TYPES)}")
    if model_type != 7:
        rho, cp, cs = f32(1000.0) * rho, f32(1000.0) * cp, f32(1000.0) * cs
    out = {"rhoinv": f32(1) / rho, "mu": (cs * cs) * rho, "lambda": ((cp * cp) - (f32(2) * cs) * cs) * rho}
    if is_diss or model_type != 7:
        out["Q"] = q
    return out


def _fill(b: F.Box, prof: np.ndarray, lpd: int) -> np.ndarray:
    """A (nz+1, lpd+1) radial profile on every point of the box (reference layout)."""
    n = lpd + 1
    shape = (b.nx + 1, b.ny + 1, b.nz + 1, n, n, n)
    return np.asfortranarray(np.broadcast_to(prof[None, None, :, None, None, :], shape).astype(np.float32))


def generate(setup: F.Setup) -> List[Tuple[F.Box, Dict[str, np.ndarray]]]:
    """(box, model fields) per rank, in rank order, as generate_models.exe
    writes them: rhoinv, mu, lambda, A, B, C and, with is_diss, Q."""
    bs, g = boxes(setup)
    out = []
    for b in bs:
        _, _, z = coordinates(b, g, setup.lpd)
        m = model_1d(setup.model_type, z, bool(setup.is_diss))
        fields = {k: _fill(b, v, setup.lpd) for k, v in m.items()}
        zero = _fill(b, np.zeros(z.shape, np.float32), setup.lpd)
        for k in ("A", "B", "C"):
            fields[k] = zero.copy()
        if not setup.is_diss:
            fields.pop("Q", None)
        out.append((b, fields))
    return out


# ---------------------------------------------------------------------------
# block perturbations
# ---------------------------------------------------------------------------

@dataclass
class Subvolume:
    """One subvolume of MODELS/MODELS_3D: block edges (colatitude and
    longitude in degrees, radius in km) and per block (x slowest, z fastest)
    the perturbations of vsh, vsv, vp [km/s] and rho [g/cm^3]."""
    bx: np.ndarray
    by: np.ndarray
    bz: np.ndarray
    dvsh: np.ndarray
    dvsv: np.ndarray
    dvp: np.ndarray
    drho: np.ndarray


def _range(edges: np.ndarray, lo: np.float32, hi: np.float32) -> Tuple[int, int]:
    """Blocks that may hold points of [lo, hi] (add_perturbation.f90:418-450):
    from the last edge below lo to the first edge above hi (0-based block
    indices, inclusive)."""
    nb = len(edges) - 1
    start, end = 0, nb - 1
    for i in range(nb):
        if edges[i] < lo:
            start = i
    for i in range(nb - 1, -1, -1):
        if edges[i] > hi:
            end = i
    return start, end


def perturb(b: F.Box, model: Dict[str, np.ndarray], subvolumes: Sequence[Subvolume],
            lpd: int = 4) -> Dict[str, np.ndarray]:
    """The box's model with the subvolumes' block perturbations added to vsh,
    vsv, vp and rho (add_perturbation.f90:259-533); A, C and Q are kept."""
    xmin, xmax, ymin, ymax = f32(b.xmin), f32(b.xmax), f32(b.ymin), f32(b.ymax)
    zmin, zmax = f32(b.zmin), f32(b.zmax)
    dx = (xmax - xmin) / f32(b.nx + 1)
    dy = (ymax - ymin) / f32(b.ny + 1)
    dz = (zmax - zmin) / f32(b.nz + 1)
    h = _half(lpd)
    cx = (xmin + np.arange(b.nx + 1, dtype=np.float32)[:, None] * dx) + h[None, :] * dx
    cy = (ymin + np.arange(b.ny + 1, dtype=np.float32)[:, None] * dy) + h[None, :] * dy
    cz = (zmax - np.arange(b.nz + 1, dtype=np.float32)[:, None] * dz) - h[None, :] * dz
    cox = cx[:, None, None, :, None, None]
    coy = cy[None, :, None, None, :, None]
    coz = cz[None, None, :, None, None, :]
    rhoinv, mu, lam, B = (np.asarray(model[k], np.float32) for k in ("rhoinv", "mu", "lambda", "B"))
    csh = np.sqrt(mu * rhoinv)
    csv = np.sqrt((mu + B) * rhoinv)
    cp = np.sqrt((lam + f32(2) * mu) * rhoinv)
    rho = f32(1) / rhoinv
    for s in subvolumes:
        bx = np.asarray(s.bx, np.float32) * PI / f32(180)
        by = np.asarray(s.by, np.float32) * PI / f32(180)
        bz = np.asarray(s.bz, np.float32) * f32(1000)
        shp = (len(bx) - 1, len(by) - 1, len(bz) - 1)
        d_sh, d_sv, d_p, d_rho = (f32(1000) * np.asarray(v, np.float32).reshape(shp)
                                  for v in (s.dvsh, s.dvsv, s.dvp, s.drho))
        (x0, x1), (y0, y1), (z0, z1) = _range(bx, xmin, xmax), _range(by, ymin, ymax), _range(bz, zmin, zmax)
        for i in range(x0, x1 + 1):
            mx = (cox >= bx[i]) & (cox < bx[i + 1])
            for j in range(y0, y1 + 1):
                mxy = mx & (coy >= by[j]) & (coy < by[j + 1])
                for k in range(z0, z1 + 1):
                    m = np.broadcast_to(mxy & (coz > bz[k]) & (coz <= bz[k + 1]), rho.shape)
                    csv = np.where(m, csv + d_sv[i, j, k], csv)
                    csh = np.where(m, csh + d_sh[i, j, k], csh)
                    cp = np.where(m, cp + d_p[i, j, k], cp)
                    rho = np.where(m, rho + d_rho[i, j, k], rho)
    mu = (rho * csh) * csh
    out = dict(model)
    out["mu"] = np.asfortranarray(mu)
    out["B"] = np.asfortranarray((rho * csv) * csv - mu)
    out["lambda"] = np.asfortranarray((rho * cp) * cp - f32(2) * mu)
    out["rhoinv"] = np.asfortranarray(f32(1) / rho)
    return out


# ---------------------------------------------------------------------------
# run trees
# ---------------------------------------------------------------------------

def write_tree(root: str, setup: Optional[F.Setup] = None) -> List[F.Box]:
    """Write ``MODELS/MODELS/boxfile`` and every rank's model files (and the
    GLL coordinate lists ``DATA/COORDINATES/{x,y,z}co_<rank>``) for the
    tree's ``INPUT/setup``, as generate_models.exe does from MODELS/MAIN."""
    setup = setup or F.Setup.read(os.path.join(root, "INPUT", "setup"))
    mdir = os.path.join(root, "MODELS", "MODELS")
    cdir = os.path.join(root, "DATA", "COORDINATES")
    os.makedirs(mdir, exist_ok=True)
    os.makedirs(cdir, exist_ok=True)
    bs, g = boxes(setup)
    F.write_boxes(os.path.join(mdir, "boxfile"), bs, setup.px, setup.py, setup.pz)
    for r, (b, fields) in enumerate(generate(setup)):
        F.write_model(mdir, fields, rank=r, is_diss=bool(setup.is_diss))
        for name, c in zip("xyz", coordinates(b, g, setup.lpd)):
            with open(os.path.join(cdir, f"{name}co_{r}"), "w") as fh:
                fh.write("".join(f"{float(v)!r}\n" for v in c.ravel()))
    return bs


def read_subvolumes(d3: str) -> List[Subvolume]:
    """MODELS/MODELS_3D/block_{x,y,z} and dvsh, dvsv, dvp, drho."""
    bx, by, bz = (F.read_block_file(os.path.join(d3, f"block_{a}")) for a in "xyz")
    vals = {n: F.read_projected(os.path.join(d3, n)) for n in ("dvsh", "dvsv", "dvp", "drho")}
    return [Subvolume(bx[k], by[k], bz[k], vals["dvsh"][k], vals["dvsv"][k], vals["dvp"][k], vals["drho"][k])
            for k in range(len(bx))]


def perturb_tree(root: str) -> None:
    """Add the perturbations of ``MODELS/MODELS_3D`` to every rank's model
    files in place, as add_perturbation.exe does from MODELS/MAIN."""
    mdir = os.path.join(root, "MODELS", "MODELS")
    subs = read_subvolumes(os.path.join(root, "MODELS", "MODELS_3D"))
    for r, b in enumerate(F.read_boxfile(os.path.join(mdir, "boxfile"))):
        shape = (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
        model = {n: F.read_field(os.path.join(mdir, f"{n}{r}"), shape) for n in F.MODEL_FIELDS}
        new = perturb(b, model, subs)
        for n in F.MODEL_FIELDS:
            F.write_field(os.path.join(mdir, f"{n}{r}"), new[n])


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="SES3D model generation on a run tree "
                                             "(generate_models.exe / add_perturbation.exe)")
    ap.add_argument("action", choices=("generate", "perturb"))
    ap.add_argument("root", help="run tree holding INPUT/ and MODELS/")
    a = ap.parse_args(argv)
    if a.action == "generate":
        write_tree(a.root)
    else:
        perturb_tree(a.root)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
