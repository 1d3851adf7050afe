"""The Fortran drop-in driver (ses3d_amd/fortran, iso_c_binding over the
C-ABI) run from MAIN/ on a reference run tree: its output files must hold the
reference solver's results (golden fixtures from the reference build)."""
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


def _run(root):
    r = subprocess.run([_driver()], cwd=os.path.join(root, "MAIN"), capture_output=True, text=True, timeout=120)
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
