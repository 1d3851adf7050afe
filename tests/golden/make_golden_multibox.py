"""Golden fixture of a run tree split into rank boxes, from the reference
solver run on as many MPI ranks (oracle/_ref/ses3d.exe under the image's
MPICH mpiexec).

Case fixture_case.MULTIBOX: the section of `multibox_adjoint` (6x5x4
elements, laterally varying elastic model, moment tensor) split 2 x 1 x 2,
run by the reference built for rank boxes of 3 x 5 x 2 elements
(oracle/pyoracle.build_ref_multirank: its static extents nx_max, ny_max,
nz_max = 2, 4, 1, as the reference's workflow sets them per run).  The
extents must match: SOURCE/ses3d_miscellaneous.f90:156-166 declares the
mass-matrix exchange's argument with the rank's own extents
((nx+1)*lpd+1, ...) while the array passed is sized by the static ones
(SOURCE/ses3d_modules.f90:105), so with larger static extents the exchange
adds the neighbour's face masses at the wrong addresses and every rank-face
node keeps half its mass (face velocities double: measured here).
The fixture holds the single-rank outputs make_golden.py stores for any case
(the drop-in driver run on the split tree must reproduce them bit for bit: the
boxes are one section), and, from the 4-rank reference run of the split tree,
  mr_seismograms (3, nr, nt)      the receivers' files
  mr_shape_<r>                    the extent of rank r's field files
  mr_<vx|vy|vz>_<r>               rank r's velocity snapshot at the last step
  mr_rel_seismograms              relative L2 distance of mr_seismograms from
                                  the reference's own single-rank run
The reference's multi-rank run differs from its single-rank run by the
rounding of per-rank geometry and of the cross-rank face sums
(SOURCE/ses3d_miscellaneous.f90:257-368); mr_rel_seismograms records that
distance.  (Ranks > 0 of an anelastic run also use sum(D_p) = 0, formed on
rank 0 only at SOURCE/ses3d_input.f90:353; this case is elastic.)

    python tests/golden/make_golden_multibox.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import formats as F, inputs as I  # noqa: E402
from tests.golden import fixture_case as FC, make_golden as MG  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make(name: str) -> str:
    path = MG.make(name)                       # single-rank outputs under the usual keys
    out = dict(np.load(path))
    decomp = FC.MULTIBOX[name]
    nproc = int(np.prod(decomp))
    root = tempfile.mkdtemp(prefix=f"golden_{name}_mr_")
    try:
        FC.write_tree(out, root, decomp=decomp)
        boxes = F.read_boxfile(os.path.join(root, "MODELS", "MODELS", "boxfile"))
        nel = FC.CASES[name][0]
        if any(n % d for n, d in zip(nel, decomp)):
            raise ValueError(f"{nel} elements do not split evenly into {decomp} rank boxes")
        per_rank = tuple(n // d for n, d in zip(nel, decomp))
        exe = PO.build_ref_multirank(per_rank)
        r = PO.run_ref(root, nprocs=nproc, exe=exe)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        run = I.load_run(root)
        e = run.events[0]
        nt = e.nt
        ofd = os.path.join(root, "MAIN", e.event.ofd)
        seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
        out["mr_seismograms"] = seis
        out["mr_rel_seismograms"] = _rel(seis, out["seismograms"])

        def rshape(b):
            return (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)

        def full(b):   # the build's static extents (SOURCE/ses3d_modules.f90:13-15) = the rank's
            return tuple(per_rank) + (5, 5, 5)

        for b in boxes:
            rk = b.index - 1
            out[f"mr_shape_{rk}"] = np.array(rshape(b))
            # the rank's snapshot file holds the reference's static extents
            assert os.path.getsize(os.path.join(ofd, f"vx_{rk}_{nt}")) == 8 + 4 * int(np.prod(full(b)))
            for c in ("vx", "vy", "vz"):
                out[f"mr_{c}_{rk}"] = F.read_field(os.path.join(ofd, f"{c}_{rk}_{nt}"), rshape(b))
        out["mr_decomp"] = np.array(decomp)
        print(f"{name}: {nproc} ranks, seismograms rel L2 to the single-rank run {out['mr_rel_seismograms']:.2e}")
        np.savez_compressed(path, **out)
        return path
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    if not PO.build_ref():
        sys.exit("reference binary unavailable (needs /root/reference and amdflang)")
    for n in (sys.argv[1:] or FC.MULTIBOX):
        print(make(n))
