"""Reference outputs of the production-size workload's first steps:
anelastic_96x96x48 (55.3 M GLL points, 3 relaxation mechanisms, the seeded
3D model, 3-element absorbing layers, 200 receivers), nt=20, from the
reference solver itself run on one rank of a build sized for the section
(oracle/Makefile ref-mr MR=95_95_47_3: static extents 95, 95, 47 and the
absorbing width pml=3, the reference's compile-time parameters).

Stored: the seismograms (3, 200, 20) from the reference's files, the sha256
of its velocity snapshots at step 20 (vx, vy, vz, float32, Fortran layout)
and their L2 norms, and the digest of the inputs (rebuilt by
ses3d_amd.configs.build with seed 0) to detect drift.

    python tests/golden/make_golden_fullsize_ref.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile
import time
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import configs, formats as F, models as M  # noqa: E402
from tests.golden.fixture_case import field_digest, input_digest  # noqa: E402

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "anelastic_96x96x48_ref.npz")
WORKLOAD = "anelastic_96x96x48"
NT = 20


def build_run():
    return configs.build(configs.WORKLOADS[WORKLOAD], seed=0, nt=NT)


def main() -> str:
    w = configs.WORKLOADS[WORKLOAD]
    run = build_run()
    e = run.events[0]
    exe = PO.build_ref_multirank((run.nx + 1, run.ny + 1, run.nz + 1), pml=w.pml)
    ev = replace(e.event, output_displacement=1, ssamp=NT)     # the snapshot at step NT is written
    case = M.RunCase(run.setup, [ev], [e.receivers], e.stf, run.model)
    root = tempfile.mkdtemp(prefix="golden_fullsize_ref_")
    try:
        M.write_run_tree(root, case)
        t0 = time.time()
        r = PO.run_ref(root, timeout=7200, exe=exe)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        print(f"reference: {NT} steps in {time.time() - t0:.0f} s", flush=True)
        ofd = os.path.join(root, "MAIN", e.event.ofd)
        seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
        shape = (run.nx + 1, run.ny + 1, run.nz + 1, 5, 5, 5)
        out = {"nt": NT, "digest": input_digest(run), "seismograms": seis}
        for c in ("vx", "vy", "vz"):
            v = F.read_field(os.path.join(ofd, f"{c}_0_{NT}"), shape)
            out[f"{c}_digest"] = field_digest(v)
            out[f"{c}_norm"] = np.linalg.norm(v.astype(np.float64))
    finally:
        shutil.rmtree(root, ignore_errors=True)
    np.savez_compressed(FIXTURE, **out)
    return FIXTURE


if __name__ == "__main__":
    print(main())
