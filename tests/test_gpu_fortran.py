"""The Fortran drop-in driver (ses3d_amd/fortran, iso_c_binding over the
C-ABI) run from MAIN/ on a reference run tree: its output files must hold the
reference solver's results (golden fixtures from the reference build), also
when the tree is split into rank boxes (per-rank model records in, per-rank
field and kernel files out) and with the absorbing width set at run time."""
import os
import subprocess

import numpy as np
import pytest

from ses3d_amd import formats as F, inputs as I
from tests.golden import fixture_case as FC

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _driver():
    from ses3d_amd import build
    if not os.path.exists(build.DRIVER):
        pytest.skip("Fortran driver not built")
    return build.DRIVER


def _run(root, *args):
    r = subprocess.run([_driver(), *args], cwd=os.path.join(root, "MAIN"), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("name", ["elastic_mt", "aniso_mt", "anelastic_adjoint"])
def test_driver_outputs_match_reference(name, tmp_path):
    fx = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    root = str(tmp_path)
    FC.write_tree(fx, root)
    print(f"driver fixture {name}", flush=True)
    _run(root)
    run = I.load_run(root)
    e = run.events[0]
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    shape = run.shape_local
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    np.testing.assert_array_equal(seis, fx["seismograms"])
    for c in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(F.read_field(os.path.join(ofd, f"{c}_0_{e.nt}"), shape), fx[c])
    if not int(fx["adjoint"]):
        return
    ffd = run.ffd("1")
    np.testing.assert_array_equal(F.read_saving_vector(os.path.join(ffd, "saving_vector_0"), e.nt), fx["saving_vector"])
    stored = [int(s) for s in fx["saving_vector"] if s > 0]
    for tag in ("first", "last"):
        step = int(fx[f"history_{tag}_step"])
        for k, fname in enumerate(("vx", "vy", "vz", "exx", "eyy", "ezz", "exy", "exz", "eyz")):
            rec = F.read_records(os.path.join(ffd, f"{fname}_0"))[stored.index(step)]
            np.testing.assert_array_equal(np.frombuffer(rec, np.float32).reshape(shape, order="F"),
                                          fx[f"history_{tag}"][k])
    FC.write_adjoint(fx, root, fx["seismograms"])
    _run(root)
    # float32 kernel sums as the reference's: its kernel files bit for bit
    for g in ("grad_rho", "grad_cp", "grad_csh", "grad_csv"):
        k = F.read_field(os.path.join(ofd, f"{g}_0"), shape)
        assert np.linalg.norm(fx[g].astype(np.float64)) > 0, g
        np.testing.assert_array_equal(k, fx[g], err_msg=g)


def test_driver_on_rank_boxes(tmp_path):
    """A 2 x 1 x 2 split of the section: the driver stitches the four ranks'
    model records into the section, and each rank's snapshot, forward-field,
    saving-vector and kernel files hold exactly that rank's part of the
    reference's single-rank results (the files' extents are the ranks' own,
    as the reference's multi-rank run writes them); its seismograms equal the
    reference's single-rank files bit for bit.  Against the reference's own
    4-rank run of this tree (built for these rank boxes,
    tests/golden/make_golden_multibox.py), which differs from its single-rank
    run by the rounding of per-rank geometry and cross-rank face sums
    (relative 7.8e-9 in the seismograms): seismograms within 1e-7 relative,
    every rank's velocity snapshot within 1e-6 of its largest value.  A face
    node with a wrong mass (the doubled face velocities of a mismatched build)
    misses by 1e-3 or more."""
    fx = dict(np.load(os.path.join(GOLDEN, "multibox_adjoint.npz")))
    decomp = tuple(int(v) for v in fx["mr_decomp"])
    root = str(tmp_path)
    FC.write_tree(fx, root, decomp=decomp)
    print(f"driver on {decomp} rank boxes", flush=True)
    _run(root)
    run = I.load_run(root)
    assert len(run.boxes) == int(np.prod(decomp))
    e = run.events[0]
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    np.testing.assert_array_equal(seis, fx["seismograms"])
    d = np.linalg.norm((seis - fx["mr_seismograms"]).astype(np.float64)) / np.linalg.norm(fx["mr_seismograms"])
    assert d < 1e-7, d
    ffd = run.ffd("1")
    stored = [int(v) for v in fx["saving_vector"] if v > 0]
    for r, b in enumerate(run.boxes):
        shape = tuple(int(v) for v in fx[f"mr_shape_{r}"])
        assert shape == (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
        for c in ("vx", "vy", "vz"):
            v = F.read_field(os.path.join(ofd, f"{c}_{r}_{e.nt}"), shape)
            np.testing.assert_array_equal(v, run.box_part(fx[c], r), err_msg=f"{c} rank {r}")
            mr = fx[f"mr_{c}_{r}"].astype(np.float64)
            assert np.abs(v - mr).max() <= 1e-6 * np.abs(mr).max(), (c, r)
        np.testing.assert_array_equal(F.read_saving_vector(os.path.join(ffd, f"saving_vector_{r}"), e.nt),
                                      fx["saving_vector"])
        assert len(F.read_records(os.path.join(ffd, f"eyz_{r}"))) == len(stored)
    FC.write_adjoint(fx, root, fx["seismograms"])
    _run(root)
    for r in range(len(run.boxes)):
        shape = tuple(int(v) for v in fx[f"mr_shape_{r}"])
        for g in ("grad_rho", "grad_cp", "grad_csh", "grad_csv"):
            k = F.read_field(os.path.join(ofd, f"{g}_{r}"), shape)
            np.testing.assert_array_equal(k, run.box_part(fx[g], r), err_msg=f"{g} rank {r}")


def test_driver_on_the_elastic_benchmark(tmp_path):
    """The elastic 24x24x12 benchmark run tree (all 2000 steps): seismogram
    files bit-identical to those of the reference solver run on one rank of a
    build sized for the section (tests/golden/make_golden_config1.py)."""
    from ses3d_amd import configs, models as M
    from tests.golden.fixture_case import input_digest
    fx = np.load(os.path.join(GOLDEN, "elastic_24x24x12_oracle.npz"))
    run = configs.build(configs.WORKLOADS["elastic_24x24x12"], seed=0)
    assert input_digest(run) == str(fx["digest"]), "inputs drifted from the fixture's"
    e = run.events[0]
    root = str(tmp_path)
    M.write_run_tree(root, M.RunCase(run.setup, [e.event], [e.receivers], e.stf, run.model))
    print("driver on the elastic 24x24x12 benchmark, 2000 steps", flush=True)
    _run(root)
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    np.testing.assert_array_equal(seis, fx["ref_seismograms"])


def test_driver_on_the_production_mesh(tmp_path):
    """The 96x96x48 anelastic run tree (55.3 M GLL points, 200 receivers,
    --pml 3), first 20 steps: seismogram and velocity snapshot files
    bit-identical to the reference solver's (tests/golden/make_golden_fullsize_ref.py)."""
    from dataclasses import replace
    from ses3d_amd import models as M
    from tests.golden import make_golden_fullsize_ref as FR
    from tests.golden.fixture_case import field_digest, input_digest
    fx = np.load(os.path.join(GOLDEN, "anelastic_96x96x48_ref.npz"))
    run = FR.build_run()
    assert input_digest(run) == str(fx["digest"]), "inputs drifted from the fixture's"
    e = run.events[0]
    nt = int(fx["nt"])
    root = str(tmp_path)
    ev = replace(e.event, output_displacement=1, ssamp=nt)
    M.write_run_tree(root, M.RunCase(run.setup, [ev], [e.receivers], e.stf, run.model))
    print(f"driver on the 96x96x48 section, {nt} steps", flush=True)
    _run(root, "--pml", "3")
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    np.testing.assert_array_equal(seis, fx["seismograms"])
    shape = (run.nx + 1, run.ny + 1, run.nz + 1, 5, 5, 5)
    for c in ("vx", "vy", "vz"):
        assert field_digest(F.read_field(os.path.join(ofd, f"{c}_0_{nt}"), shape)) == str(fx[f"{c}_digest"]), c


def test_driver_on_the_gradient_mesh(tmp_path):
    """The gradient benchmark's mesh (48x48x24, anelastic, 3D model, 100
    receivers, 300 steps, forward fields stored every 15 steps: 5 GB of
    history files): seismogram files and, after the adjoint run, the four
    kernel files bit-identical to the reference solver's
    (tests/golden/make_golden_gradient_ref.py)."""
    from dataclasses import replace
    from ses3d_amd import models as M
    from tests.golden import make_golden_gradient_ref as GR
    from tests.golden.fixture_case import field_digest, input_digest
    fx = np.load(os.path.join(GOLDEN, "gradient_48x48x24_ref.npz"))
    run = GR.build_run()
    assert input_digest(run) == str(fx["digest"]), "inputs drifted from the fixture's"
    e = run.events[0]
    root = str(tmp_path)
    M.write_run_tree(root, M.RunCase(replace(run.setup, adjoint_flag=1), [e.event], [e.receivers], e.stf, run.model))
    print("driver on the 48x48x24 gradient mesh: forward", flush=True)
    _run(root)
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    np.testing.assert_array_equal(seis, fx["seismograms"])
    GR.write_adjoint_files(root, run, fx["seismograms"])
    print("adjoint", flush=True)
    _run(root)
    shape = (run.nx + 1, run.ny + 1, run.nz + 1, 5, 5, 5)
    for g in ("grad_rho", "grad_cp", "grad_csh", "grad_csv"):
        assert field_digest(F.read_field(os.path.join(ofd, f"{g}_0"), shape)) == str(fx[f"{g}_digest"]), g


def test_driver_absorbing_width_at_run_time(tmp_path):
    """--pml 3 (the reference's compile-time pml, SOURCE/ses3d_modules.f90:21)
    gives the seismograms of the library run with pml=3, which differ from the
    default width's."""
    from ses3d_amd.solver import Solver
    fx = dict(np.load(os.path.join(GOLDEN, "elastic_mt.npz")))
    root = str(tmp_path)
    FC.write_tree(fx, root)
    _run(root, "--pml", "3")
    run = I.load_run(root)
    e = run.events[0]
    ofd = os.path.join(root, "MAIN", e.event.ofd)
    seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
    s = Solver.from_run(run, adjoint_flag=0, pml=3)
    s.set_model(run.model)
    s.load_event(0, e)
    s.init()
    s.run()
    np.testing.assert_array_equal(seis, s.seismograms())
    s.close()
    assert not np.array_equal(seis, fx["seismograms"])
