"""Error of the GPU's float32 results against the float64 restatement of the
reference algorithm (the float32 results equal the reference's float32
build bit for bit; this states how far that arithmetic is from float64).

Anelastic 16x16x8 elements, 300 steps, moment tensor source, 4 surface
receivers; relative L2 per result <= 1e-4 (measured: seismograms 1.7e-5,
final velocity 4-7e-6, split stresses 4-6e-6, memory variables 5-7e-6).  The
float64 oracle's kernels are compared in tests/test_gpu_adjoint.py (<= 1e-4)."""
import numpy as np
import pytest

from tests.cases import make_case, oracle_for

pytestmark = pytest.mark.gpu


def _rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_float32_gpu_vs_float64_oracle(tmp_path):
    from ses3d_amd.solver import Solver
    nt = 300
    run = make_case(str(tmp_path), nel=(16, 16, 8), nt=nt, is_diss=True, seed=21, period=20.0)
    e = run.events[0]
    s = Solver.from_run(run, adjoint_flag=0)
    s.set_model(run.model)
    s.load_event(0, e)
    s.init()
    s.run()
    o = oracle_for(run, e, precision=64)
    o.run(1, nt)
    res = {"seismograms": _rel(s.seismograms(), o.get("seismograms"))}
    for c in ("vx", "vy", "vz"):
        res[c + "_global"] = _rel(s.field(c + "_global"), o.get(c + "g"))
    for f in ("sxx", "syy", "szz", "sxy", "sxz", "syz", "Mxx", "Mxy", "Myz"):
        res[f] = _rel(s.field(f), o.get(f))
    for k, v in res.items():
        print(f"relative L2 vs float64 oracle, {k}: {v:.2e}", flush=True)
    assert max(res.values()) <= 1e-4, res
    s.close()
