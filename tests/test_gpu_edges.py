"""Edge cases of the reference's set-up and stepping on the GPU, each against
the float32 oracle bit for bit (the oracle restates the reference's rules and
is pinned to the reference binary by the golden fixtures):

* meshes of 1 to 4 elements per axis with absorbing layers 0 to 3 elements
  wide;
* receivers outside the section (dropped), exactly on an element face (the
  last matching element wins, SOURCE/ses3d_init.f90:292-347) and an event with
  no receivers at all;
* a moment-tensor source exactly on a GLL node shared by 8 elements (the
  first nearest node wins, SOURCE/ses3d_init.f90:373-417);
* adjoint sources: two in one element (applied in list order), one outside
  the section.
"""
import numpy as np
import pytest

from tests.cases import make_case, oracle_for
from tests.test_gpu_forward import _assert_same

pytestmark = pytest.mark.gpu

FIELDS = ("sxx", "sxy", "syx", "szz", "Mxx", "Myz")


def _x(run, i, n=0):
    """float32 GLL coordinate x(i, n) as the reference computes it."""
    f = np.float32
    dx = f((f(run.xmax) - f(run.xmin)) / f(run.nx + 1))
    knots = np.array([-1.0, -0.6546536707079772, 0.0, 0.6546536707079772, 1.0], np.float32)
    return f(f(f(run.xmin) + f(i) * dx) + f(f(0.5) * f(f(1.0) + knots[n])) * dx)


def _y(run, j, n=0):
    f = np.float32
    dy = f((f(run.ymax) - f(run.ymin)) / f(run.ny + 1))
    knots = np.array([-1.0, -0.6546536707079772, 0.0, 0.6546536707079772, 1.0], np.float32)
    return f(f(f(run.ymin) + f(j) * dy) + f(f(0.5) * f(f(1.0) + knots[n])) * dy)


def _z(run, k, n=0):
    f = np.float32
    dz = f((f(run.zmax) - f(run.zmin)) / f(run.nz + 1))
    knots = np.array([-1.0, -0.6546536707079772, 0.0, 0.6546536707079772, 1.0], np.float32)
    return f(f(f(run.zmax) - f(k) * dz) - f(f(0.5) * f(f(1.0) + knots[n])) * dz)


def _compare(s, o, what, with_seis=True):
    for c in ("vx", "vy", "vz"):
        _assert_same(s.field(c + "_global"), o.get(c + "g"), f"{c} {what}")
    for f in FIELDS:
        if f.startswith("M") and not s.cfg.is_diss:
            continue
        _assert_same(s.field(f), o.get(f), f"{f} {what}")
    if with_seis:
        _assert_same(s.seismograms(), o.get("seismograms"), f"seismograms {what}")


@pytest.mark.parametrize("nel,pml", [((1, 1, 1), 0), ((2, 1, 3), 1), ((3, 4, 2), 2), ((4, 4, 4), 3)])
def test_small_meshes_and_layer_widths(tmp_path, nel, pml):
    from ses3d_amd.solver import Solver
    run = make_case(str(tmp_path), nel=nel, nt=30, is_diss=True, seed=31, n_rec=(1, 2))
    e = run.events[0]
    print(f"case nel={nel} pml={pml}", flush=True)
    s = Solver.from_run(run, adjoint_flag=0, pml=pml)
    s.set_model(run.model)
    s.load_event(0, e)
    s.init()
    s.run()
    o = oracle_for(run, e, pml=pml)
    o.run(1, e.nt)
    _compare(s, o, f"nel={nel} pml={pml}")
    s.close()


def test_receivers_on_faces_outside_and_none(tmp_path):
    from ses3d_amd.solver import Solver
    run = make_case(str(tmp_path), nel=(5, 5, 4), nt=30, is_diss=False, seed=32, n_events=2)
    e = run.events[0]
    f = np.float32
    zmax_g = f(run.zmax_global)
    rec = np.array([
        [_x(run, 2), _y(run, 3, 2), f(0.0)],                         # on the face between elements 1 and 2 in x
        [_x(run, 1, 2), _y(run, 2), f(zmax_g - _z(run, 1))],        # on a y face and a z face
        [f(np.deg2rad(75.0)), _y(run, 1, 1), f(0.0)],                 # outside the section
        [_x(run, 3, 3), _y(run, 4, 1), f(zmax_g - _z(run, 2, 3))],  # inside an element
    ], np.float32)
    s = Solver.from_run(run, adjoint_flag=0, n_events=2)
    s.set_model(run.model)
    s.set_source(0, e.source_type, e.xxs, e.yys, e.zzs, e.moment, e.stf)
    s.set_receivers(0, rec)
    s.set_source(1, e.source_type, e.xxs, e.yys, e.zzs, e.moment, e.stf)
    s.set_receivers(1, np.zeros((0, 3), np.float32))
    s.init()
    s.run()
    assert s.nr(0) == 3 and list(s.receiver_index(0)) == [0, 1, 3]
    assert s.nr(1) == 0 and s.seismograms(1).shape == (3, 0, e.nt)
    o = oracle_for(run, e)
    o.set_receivers(rec)
    o.init()
    o.run(1, e.nt)
    _compare(s, o, "receivers on faces")
    np.testing.assert_array_equal(s.field("vx_global", 1), s.field("vx_global", 0))
    s.close()


def test_source_on_a_node_shared_by_eight_elements(tmp_path):
    from ses3d_amd.solver import Solver
    run = make_case(str(tmp_path), nel=(4, 4, 4), nt=30, is_diss=True, seed=33)
    e = run.events[0]
    xs, ys = _x(run, 2), _y(run, 2)
    zs = np.float32(np.float32(run.zmax_global) - _z(run, 2))
    s = Solver.from_run(run, adjoint_flag=0)
    s.set_model(run.model)
    s.set_source(0, 10, xs, ys, zs, e.moment, e.stf)
    s.set_receivers(0, e.rec32)
    s.init()
    s.run()
    o = oracle_for(run, e)
    o.set_source(10, xs, ys, zs, e.moment, e.stf)
    o.init()
    o.run(1, e.nt)
    assert np.abs(o.get("seismograms")).max() > 0
    _compare(s, o, "source on a shared node")
    s.close()


def test_adjoint_sources_sharing_an_element_and_outside(tmp_path):
    from ses3d_amd.solver import Solver
    from oracle import pyoracle as PO
    run = make_case(str(tmp_path), nel=(5, 4, 4), nt=30, is_diss=True, seed=34)
    f = np.float32
    zmax_g = f(run.zmax_global)
    loc = np.array([
        [_x(run, 2, 1), _y(run, 1, 2), f(zmax_g - _z(run, 1, 1))],
        [_x(run, 2, 3), _y(run, 1, 3), f(zmax_g - _z(run, 1, 2))],   # same element as the first
        [f(np.deg2rad(20.0)), _y(run, 1, 1), f(1000.0)],              # outside the section
        [_x(run, 3, 2), _y(run, 2, 2), f(zmax_g - _z(run, 2, 2))],
    ], np.float32)
    rng = np.random.default_rng(5)
    nt = run.events[0].nt
    rows = (rng.standard_normal((len(loc), nt, 3)) * 1e-3).astype(np.float32)
    s = Solver.from_run(run, adjoint_flag=2)
    s.set_model(run.model)
    s.set_adjoint_sources(0, loc, rows)
    s.init()
    o = PO.Oracle(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin, ymax=run.ymax,
                  zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global, nt=nt, dt=run.events[0].dt,
                  is_diss=run.is_diss, adjoint_flag=2, samp_ad=run.samp_ad, pml=2, tau_p=run.tau_p, D_p=run.D_p)
    o.set_model(run.model)
    o.set_adjoint_sources(loc, rows)
    o.init()
    # the adjoint wavefield does not depend on the stored forward fields; run
    # both without them (zero history) and compare the wavefields
    zeros = np.zeros((9,) + s.shape_local, np.float32)
    for st in np.flatnonzero(s.saving_vector()) + 1:
        s.put_history(int(st), zeros)
    s.run()
    o.run(1, nt)
    _compare(s, o, "adjoint sources", with_seis=False)
    s.close()
