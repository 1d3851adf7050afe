"""GPU path against the reference solver's own outputs (golden fixtures written
by tests/golden/make_golden.py from the reference Fortran build: 1D and
laterally varying models, moment tensor and point forces along theta, phi and
r, elastic, anelastic and anisotropic).  Forward fields and seismograms:
bit-exact.  Kernels (the four elastic ones and, anelastic, grad_Q_mu of the
reference's grad_Q build): with float32 accumulators (the reference's type)
bit-exact; with the default float64 accumulators the same float32 increments
are rounded less, so the bar is relative L2 <= 1e-5."""
import os

import numpy as np
import pytest

from ses3d_amd import inputs as I
from tests.golden import fixture_case as FC

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _solver(run, af=None, kernel_bits=64):
    from ses3d_amd.solver import Solver
    s = Solver.from_run(run, adjoint_flag=af)
    if af == 2:
        s.set_kernel_precision(kernel_bits)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    return s


@pytest.mark.parametrize("name", list(FC.CASES))
def test_gpu_matches_reference_fixture(name, tmp_path):
    fx = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    FC.write_tree(fx, str(tmp_path))
    run = I.load_run(str(tmp_path))
    nt = run.events[0].nt
    print(f"fixture {name}", flush=True)
    s = _solver(run)
    s.run(1, nt - 1)
    snap = {c: s.field(c) for c in ("vx", "vy", "vz")}
    s.run(nt, 1)
    np.testing.assert_array_equal(s.seismograms(), fx["seismograms"])
    for c in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(snap[c], fx[c])
    if not int(fx["adjoint"]):
        return
    np.testing.assert_array_equal(s.saving_vector(), fx["saving_vector"])
    for tag in ("first", "last"):
        if f"history_{tag}" in fx:
            np.testing.assert_array_equal(s.history(int(fx[f"history_{tag}_step"])), fx[f"history_{tag}"])
    FC.write_adjoint(fx, str(tmp_path), fx["seismograms"])
    run2 = I.load_run(str(tmp_path))
    for bits in (64, 32):
        a = _solver(run2, af=2, kernel_bits=bits)
        for st in np.flatnonzero(s.saving_vector()) + 1:
            a.put_history(int(st), s.history(int(st)))
        a.run(1, nt - 1)
        avx = a.field("vx")
        a.run(nt, 1)
        np.testing.assert_array_equal(avx, fx["adjoint_vx"])
        diss = int(fx["is_diss"])
        k = a.kernels(with_q=bool(diss))
        a.close()
        # grad_Q_mu: the reference's grad_Q build (SOURCE/ses3d_grad_Q.f90:111-128,151-153)
        for g in ("grad_rho", "grad_cp", "grad_csh", "grad_csv") + (("grad_Q_mu",) if diss else ()):
            ref = fx[g].astype(np.float64)
            assert np.linalg.norm(ref) > 0, g
            if bits == 32:
                np.testing.assert_array_equal(k[g], ref, err_msg=g)
            else:
                rel = np.linalg.norm(k[g] - ref) / np.linalg.norm(ref)
                assert rel <= 1e-5, (g, rel)
