"""Block-grid projection of kernels (TOOLS/SOURCE/project_kernel.f90) on the GPU.

* Against the reference tool itself: the reference's float32 kernels of the
  anelastic_adjoint fixture projected by oracle/_ref/project_kernel.exe
  (tests/golden/make_golden_projection.py) on a grid that cuts elements at
  arbitrary positions; the tool prints 7 significant digits, so the bar is
  relative L2 <= 1e-6 and every block within 1e-5 of the largest (measured:
  identical to the printed digits for all four kernels).
* Closed form: a constant field over one block equal to the whole section
  integrates each element's 125 Lagrange polynomials to 2 x 2 x 2, so the
  block value is 8 x elements / (block width product), to the rounding of a
  float32 sum of 12,500 terms in the tool's order (measured 1.5e-5; bound 1e-4).
"""
import os

import numpy as np
import pytest

from tests.golden import fixture_case as FC

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PAIRS = (("grad_csh", "gradient_csh"), ("grad_csv", "gradient_csv"), ("grad_cp", "gradient_cp"),
         ("grad_rho", "gradient_rho"))


def _solver(tmp_path):
    from ses3d_amd import inputs as I
    from ses3d_amd.solver import Solver
    fx = dict(np.load(os.path.join(HERE, "anelastic_adjoint.npz")))
    FC.write_tree(fx, str(tmp_path))
    run = I.load_run(str(tmp_path))
    return fx, run, Solver.from_run(run, adjoint_flag=0)


def test_projection_matches_reference_tool(tmp_path):
    fx, run, s = _solver(tmp_path)
    pj = np.load(os.path.join(HERE, "projection_anelastic_adjoint.npz"))
    for g, o in PAIRS:
        got = s.project(fx[g], pj["bx"], pj["by"], pj["bz"]).astype(np.float64)
        ref = pj[o].astype(np.float64)
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        worst = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"{o}: relative L2 {rel:.2e}, worst block {worst:.2e} of the largest", flush=True)
        assert np.abs(ref).max() > 0 and rel <= 1e-6 and worst <= 1e-5, (o, rel, worst)
    s.close()


def test_projection_of_constant_over_whole_section(tmp_path):
    fx, run, s = _solver(tmp_path)
    st = run.setup
    ones = np.ones((run.nx + 1, run.ny + 1, run.nz + 1, 5, 5, 5), np.float32)
    bx = np.array([st.theta_min, st.theta_max])
    by = np.array([st.phi_min, st.phi_max])
    bz = np.array([st.z_min, st.z_max]) / 1000.0
    got = float(s.project(ones, bx, by, bz)[0, 0, 0])
    n_el = (run.nx + 1) * (run.ny + 1) * (run.nz + 1)
    expect = 8.0 * n_el / ((bx[1] - bx[0]) * (by[1] - by[0]) * (bz[1] - bz[0]))
    assert abs(got / expect - 1.0) < 1e-4, (got, expect)
    s.close()
