"""ses3d_hip_project_noalloc (the projection staged in the library's fixed
per-GPU scratch array, no device allocation) against ses3d_hip_project
(per-call device buffer), both through the C-ABI: bit-identical block values on
two calls in a row, the second on a smaller mesh and block grid (the scratch
keeps the first call's contents, so stale data would show), and a clear error
for a projection larger than the scratch array."""
import ctypes as C

import numpy as np
import pytest

from tests.cases import make_case

pytestmark = pytest.mark.gpu


def _edges(run, nb, rng):
    st = run.setup
    bx = np.sort(rng.uniform(st.theta_min, st.theta_max, nb[0]))
    by = np.sort(rng.uniform(st.phi_min, st.phi_max, nb[1]))
    bz = np.sort(rng.uniform(st.z_min, st.z_max, nb[2])) / 1000.0
    return [np.ascontiguousarray(a, np.float32) for a in (bx, by, bz)]


def _call(fn, s, field, e):
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    out = np.full((len(e[0]) - 1) * (len(e[1]) - 1) * (len(e[2]) - 1), np.nan, np.float32)
    rc = fn(s.h, fp(field), len(e[0]), fp(e[0]), len(e[1]), fp(e[1]), len(e[2]), fp(e[2]), fp(out))
    return rc, out


@pytest.mark.parametrize("order", ["noalloc_first", "alloc_first"])
def test_noalloc_projection_equals_allocating_one(tmp_path, order):
    from ses3d_amd.solver import Solver, load_library
    lib = load_library()
    rng = np.random.default_rng(5)
    cases = (((16, 16, 8), (13, 11, 9)), ((5, 5, 4), (6, 5, 4)))   # second call: smaller mesh and grid
    for k, (nel, nb) in enumerate(cases):
        print(f"call {k}: mesh {nel}, block edges {nb}, {order}", flush=True)
        run = make_case(str(tmp_path / f"c{k}"), nel=nel, nt=2, perturb=False)
        s = Solver.from_run(run, adjoint_flag=0)
        field = np.ascontiguousarray(rng.standard_normal(run.shape_local).astype(np.float32).ravel(order="F"))
        e = _edges(run, nb, rng)
        fns = [lib.ses3d_hip_project_noalloc, lib.ses3d_hip_project]
        if order == "alloc_first":
            fns.reverse()
        r0, o0 = _call(fns[0], s, field, e)
        r1, o1 = _call(fns[1], s, field, e)
        assert r0 == 0 and r1 == 0, lib.ses3d_hip_last_error()
        assert np.isfinite(o0).all() and np.abs(o0).max() > 0
        np.testing.assert_array_equal(o0, o1)
        s.close()


def test_noalloc_projection_rejects_oversize(tmp_path):
    """A 64x64x33 mesh has 16.9 million GLL points: its field alone exceeds the
    64 MiB scratch, so the entry point refuses before touching the device."""
    from ses3d_amd.solver import Solver, load_library
    lib = load_library()
    d2r = np.float32(np.pi / 180)
    s = Solver(nx=63, ny=63, nz=32, xmin=40 * d2r, xmax=60 * d2r, ymin=10 * d2r, ymax=30 * d2r,
               zmin=5371000.0, zmax=6371000.0, zmax_global=6371000.0, nt=2, dt=0.1)
    field = np.zeros(s.n_points, np.float32)
    e = [np.array(v, np.float32) for v in ([42.0, 50.0, 58.0], [12.0, 20.0, 28.0], [5400.0, 5900.0, 6300.0])]
    rc, _ = _call(lib.ses3d_hip_project_noalloc, s, field, e)
    assert rc == -1 and b"at most" in lib.ses3d_hip_last_error()
    rc, out = _call(lib.ses3d_hip_project, s, field, e)
    assert rc == 0 and np.array_equal(out, np.zeros_like(out))
    s.close()
