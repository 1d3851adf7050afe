"""The elastic workload's physics (1D PREM-like model, moment tensor at 10 km
depth below colatitude 50 / longitude 20, 20 surface receivers, dt=0.2 s)
at the largest mesh one rank of the reference build holds, 23x28x8
elements, for the full 2000 steps on the GPU: seismograms bit-identical to
the reference binary's output files and final velocity snapshots with the
reference's sha256 digests (fixture: tests/golden/make_golden_elastic_large.py)."""
import numpy as np
import pytest

from tests.golden.fixture_case import ELASTIC_LARGE, ELASTIC_LARGE_FIXTURE, field_digest, input_digest

pytestmark = pytest.mark.gpu


def test_elastic_23x28x8_full_run_matches_reference():
    from ses3d_amd import configs
    from ses3d_amd.solver import Solver
    fx = np.load(ELASTIC_LARGE_FIXTURE)
    run = configs.build(ELASTIC_LARGE, seed=0)
    assert input_digest(run) == str(fx["digest"]), "inputs drifted from the fixture's"
    nt = int(fx["nt"])
    s = Solver.from_run(run, adjoint_flag=0, pml=ELASTIC_LARGE.pml)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    print(f"{ELASTIC_LARGE.name}: {nt} steps", flush=True)
    s.run(1, nt - 1)
    snap = {c: s.field(c) for c in ("vx", "vy", "vz")}   # the local copy the last step gathers
    s.run(nt, 1)
    np.testing.assert_array_equal(s.seismograms(), fx["seismograms"])
    for c in ("vx", "vy", "vz"):
        assert field_digest(snap[c]) == str(fx[c + "_digest"]), c
    s.close()
