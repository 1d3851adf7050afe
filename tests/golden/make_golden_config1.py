"""Digests of the elastic 24x24x12 workload (1D PREM-like model, moment
tensor at 10 km depth below colatitude 50 / longitude 20, 20 surface
receivers, dt=0.2 s, nt=2000) for its GPU parity test, from the reference
solver itself and from the float32 oracle, which must agree bit for bit.

The reference's default build holds at most 23x28x8 elements per rank
(SOURCE/ses3d_modules.f90:13-15, static extents); this workload runs on one
rank of the reference built with nx_max, ny_max, nz_max = 23, 23, 11
(oracle/Makefile ref-mr, the three parameters the reference's workflow sets
per run).

Stored: sha256 of the seismograms (3, 20, 2000) float32 and of the final
velocity snapshots (vx, vy, vz gathered at the start of step 2000, reference
layout), their norms, the digest of the inputs (rebuilt by
ses3d_amd.configs.build with seed 0) to detect drift, and the reference
run's seismograms (ref_seismograms).

    python tests/golden/make_golden_config1.py
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import shutil  # noqa: E402
import tempfile  # noqa: E402
from dataclasses import replace  # noqa: E402

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import configs, formats as F, models as M  # noqa: E402
from tests.cases import oracle_for  # noqa: E402
from tests.golden.fixture_case import field_digest, input_digest  # noqa: E402

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elastic_24x24x12_oracle.npz")
WORKLOAD = "elastic_24x24x12"


def seis_digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def reference_run(run):
    """The reference solver on one rank sized for the section: seismograms
    (3, nr, nt) from its output files and its velocity snapshots at nt."""
    e = run.events[0]
    nt = e.nt
    exe = PO.build_ref_multirank((run.nx + 1, run.ny + 1, run.nz + 1))
    ev = replace(e.event, output_displacement=1)     # the snapshot at nt is written
    case = M.RunCase(run.setup, [ev], [e.receivers], e.stf, run.model)
    root = tempfile.mkdtemp(prefix="golden_config1_")
    try:
        M.write_run_tree(root, case)
        t0 = time.time()
        r = PO.run_ref(root, timeout=3600, exe=exe)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        print(f"reference: {nt} steps in {time.time() - t0:.0f} s")
        ofd = os.path.join(root, "MAIN", e.event.ofd)
        seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
        shape = (run.nx + 1, run.ny + 1, run.nz + 1, 5, 5, 5)
        snap = {c: F.read_field(os.path.join(ofd, f"{c}_0_{nt}"), shape) for c in ("vx", "vy", "vz")}
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return seis, snap


def main() -> str:
    w = configs.WORKLOADS[WORKLOAD]
    run = configs.build(w, seed=0)
    e = run.events[0]
    nt = e.nt
    ref_seis, ref_snap = reference_run(run)
    o = oracle_for(run, e, adjoint_flag=0, pml=w.pml)
    t0 = time.time()
    o.run(1, nt - 1)
    snap = {c: o.get(c + "g").astype(np.float32) for c in ("vx", "vy", "vz")}
    o.run(nt, 1)
    # the snapshot of step nt holds the velocity gathered at its start
    loc = {c: o.get(c).astype(np.float32) for c in ("vx", "vy", "vz")}
    seis = o.get("seismograms").astype(np.float32)
    print(f"oracle: {nt} steps in {time.time() - t0:.0f} s, max |v| {np.abs(seis).max():.3e}")
    if not np.array_equal(ref_seis, seis):
        raise RuntimeError("oracle and reference seismograms differ")
    for c in loc:
        if field_digest(ref_snap[c]) != field_digest(loc[c]):
            raise RuntimeError(f"oracle and reference {c} snapshots differ")
    np.savez_compressed(FIXTURE, nt=nt, digest=input_digest(run), seismograms_digest=seis_digest(seis),
                        ref_seismograms=ref_seis,
                        seismograms_norm=np.linalg.norm(seis.astype(np.float64)),
                        **{f"{c}_digest": field_digest(loc[c]) for c in loc},
                        **{f"{c}_global_norm": np.linalg.norm(snap[c].astype(np.float64)) for c in snap})
    return FIXTURE


if __name__ == "__main__":
    print(main())
