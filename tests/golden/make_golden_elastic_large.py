"""Reference outputs for the elastic workload at the largest mesh the
reference build holds on one rank: 23x28x8 elements (its compile-time
maximum, SOURCE/ses3d_modules.f90:13-15), 1D PREM-like model, moment tensor
at 10 km depth below colatitude 50 / longitude 20, 20 surface receivers,
dt=0.2 s, nt=2000 (the 24x24x12 elastic workload needs 12 radial elements,
more than one rank of this build holds).

Stored: the seismograms (3, 20, 2000) read from the reference's own output
files, and sha256 digests of its final velocity snapshots (vx, vy, vz at
step 2000, cropped to the mesh, float32 in the reference's Fortran layout).
The inputs are rebuilt by ses3d_amd.configs.build with seed 0; their digest
is stored to detect drift.

    python tests/golden/make_golden_elastic_large.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import configs, formats as F, models as M  # noqa: E402
from tests.golden.fixture_case import (ELASTIC_LARGE as WORKLOAD, ELASTIC_LARGE_FIXTURE as FIXTURE,  # noqa: E402
                                       field_digest, input_digest)

FULL = (23, 28, 8, 5, 5, 5)      # the reference's static array extents


def main() -> str:
    run = configs.build(WORKLOAD, seed=0)
    e = run.events[0]
    nt = e.nt
    # output_displacement=1 makes the reference write its velocity snapshot at nt
    ev = replace(e.event, output_displacement=1)
    case = M.RunCase(run.setup, [ev], [e.receivers], e.stf, run.model)
    root = tempfile.mkdtemp(prefix="golden_elastic_large_")
    try:
        M.write_run_tree(root, case)
        r = PO.run_ref(root, timeout=1800)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        ofd = os.path.join(root, "MAIN", e.event.ofd)
        seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
        dig = {c: field_digest(F.read_field(os.path.join(ofd, f"{c}_0_{nt}"), FULL)[: run.nx + 1, : run.ny + 1,
                                                                                   : run.nz + 1])
               for c in ("vx", "vy", "vz")}
    finally:
        shutil.rmtree(root, ignore_errors=True)
    np.savez_compressed(FIXTURE, seismograms=seis, nt=nt, digest=input_digest(run),
                        vx_digest=dig["vx"], vy_digest=dig["vy"], vz_digest=dig["vz"])
    print(f"wrote {FIXTURE}: seismograms {seis.shape}, max |v| {np.abs(seis).max():.3e}")
    return FIXTURE


if __name__ == "__main__":
    main()
