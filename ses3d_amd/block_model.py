"""Block models of MODELS/MODELS_3D as the reference's model tool reads,
combines and writes them (TOOLS/models.py:11-314).

A block model is a list of subvolumes, each a grid of blocks between edges in
colatitude (block_x, degrees), longitude (block_y, degrees) and radius
(block_z, km), with one value per block.  The edges live in the three block
files the kernel projection and the perturbation tool share; the values of one
quantity (dvsv, dvsh, dvp, drho, or a projected gradient such as
gradient_csv) live in a file of their own: the number of subvolumes, then per
subvolume the number of blocks and one value per line, colatitude slowest and
radius fastest.

``BlockModel`` offers the tool's ``read``, ``write``, scalar multiplication
and addition, which is what a model update needs (model + step * gradient,
then MODELS/SOURCE/add_perturbation.f90, ses3d_amd.generator.perturb_tree).
Rotated coordinates, smoothing, VTK export and plots are presentation and are
not carried over.
"""
from __future__ import annotations

import copy
import os
from typing import List, Sequence

import numpy as np

from . import formats as F
from .adjoint_source import py2_str


class SubModel:
    """One subvolume (TOOLS/models.py:15-31): edge lines and block values.

    ``colat`` keeps the block_x edges as read; ``lat`` = 90 - colat as the
    tool holds them."""

    def __init__(self, colat=(0.0,), lon=(0.0,), r=(0.0,), v=None):
        self.colat = np.asarray(colat, np.float64)
        self.lat = 90.0 - self.colat
        self.lon = np.asarray(lon, np.float64)
        self.r = np.asarray(r, np.float64)
        shape = (max(len(self.colat) - 1, 1), max(len(self.lon) - 1, 1), max(len(self.r) - 1, 1))
        self.v = np.zeros(shape) if v is None else np.asarray(v, np.float64).reshape(shape)

    @property
    def shape(self):
        return (len(self.colat) - 1, len(self.lon) - 1, len(self.r) - 1)


def _lines(path: str) -> np.ndarray:
    """The file's lines as floats (TOOLS/models.py:177-179, 236: one number
    per line, no blank lines inside)."""
    with open(path) as fh:
        return np.array(fh.read().strip().split("\n"), dtype=float)


class BlockModel:
    """A block model (TOOLS/models.py:38-314)."""

    def __init__(self):
        self.nsubvol = 0
        self.m: List[SubModel] = []

    @classmethod
    def from_blocks(cls, bx: Sequence[np.ndarray], by: Sequence[np.ndarray], bz: Sequence[np.ndarray],
                    values: Sequence[np.ndarray]) -> "BlockModel":
        """From per-subvolume edges (colatitude and longitude in degrees,
        radius in km) and values ((nx, ny, nz) or flat, radius fastest)."""
        if not (len(bx) == len(by) == len(bz) == len(values)):
            raise ValueError("edges and values must list the same subvolumes")
        out = cls()
        for x, y, z, v in zip(bx, by, bz, values):
            sub = SubModel(x, y, z)
            v = np.asarray(v, np.float64)
            if v.size != int(np.prod(sub.shape)):
                raise ValueError(f"subvolume of {sub.shape} blocks given {v.size} values")
            sub.v = v.reshape(sub.shape).copy()
            out.m.append(sub)
        out.nsubvol = len(out.m)
        return out

    def read(self, directory: str, filename: str) -> None:
        """Read ``directory`` + block_x, block_y, block_z and ``filename``
        (TOOLS/models.py:157-254); ``directory`` ends with '/'."""
        dx, dy, dz = (_lines(directory + "block_" + a) for a in "xyz")
        n = int(dx[0])
        if int(dy[0]) != n or int(dz[0]) != n:
            raise ValueError("block_x, block_y and block_z disagree on the number of subvolumes")
        v = _lines(directory + filename)
        if int(v[0]) != n:
            raise ValueError(f"{filename} holds {int(v[0])} subvolumes, the block files {n}")
        self.m, self.nsubvol = [], n
        ix = iy = iz = 1
        iv = 1
        for _ in range(n):
            cx, cy, cz = int(dx[ix]), int(dy[iy]), int(dz[iz])
            sub = SubModel(dx[ix + 1:ix + 1 + cx], dy[iy + 1:iy + 1 + cy], dz[iz + 1:iz + 1 + cz])
            ix, iy, iz = ix + cx + 1, iy + cy + 1, iz + cz + 1
            nb = int(v[iv])
            if nb != int(np.prod(sub.shape)):
                raise ValueError(f"{filename}: subvolume of {sub.shape} blocks holds {nb} values")
            sub.v = v[iv + 1:iv + 1 + nb].reshape(sub.shape)
            iv += nb + 1
            self.m.append(sub)

    def write(self, directory: str, filename: str) -> None:
        """Write the values as the tool does (TOOLS/models.py:283-312): counts,
        then one value per line in Python 2's ``str`` (12 significant digits)."""
        out = [str(self.nsubvol)]
        for sub in self.m:
            out.append(str(int(np.prod(sub.shape))))
            out += [py2_str(x) for x in sub.v.reshape(-1)]
        with open(directory + filename, "w") as fh:
            fh.write("\n".join(out) + "\n")

    def write_blocks(self, directory: str) -> None:
        """block_x, block_y, block_z of this model's edges (colatitude,
        longitude, radius), as the projection and perturbation tools read them."""
        os.makedirs(directory, exist_ok=True)
        for a, edges in (("x", [s.colat for s in self.m]), ("y", [s.lon for s in self.m]),
                         ("z", [s.r for s in self.m])):
            F.write_block_file(os.path.join(directory, "block_" + a), edges)

    def _check(self, other: "BlockModel") -> None:
        if other.nsubvol != self.nsubvol or any(a.v.shape != b.v.shape for a, b in zip(self.m, other.m)):
            raise ValueError("models on different block grids")

    def __rmul__(self, factor: float) -> "BlockModel":
        res = copy.deepcopy(self)
        for sub in res.m:
            sub.v = factor * sub.v
        return res

    def __add__(self, other: "BlockModel") -> "BlockModel":
        self._check(other)
        res = copy.deepcopy(self)
        for sub, o in zip(res.m, other.m):
            sub.v = sub.v + o.v
        return res


# the reference tool's class names (TOOLS/models.py:15, :38)
ses3d_submodel = SubModel
ses3d_model = BlockModel
