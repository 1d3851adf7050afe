"""The reference's own example inputs (INPUT/, ADJOINT/1/, MODELS/MODELS_3D/ of
the reference tree) parsed by ses3d_amd.formats, read where they lie; skipped
on a machine that holds only this repository.  Expected values are the files'
contents as the reference's list-directed reads take them
(SOURCE/ses3d_input.f90, SOURCE/ses3d_init.f90:479-490,
TOOLS/SOURCE/project_kernel.f90:306-372)."""
import os

import numpy as np
import pytest

from ses3d_amd import formats as F

REF = "/root/reference"
pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "INPUT")),
                                reason="reference tree absent")


def _p(*a):
    return os.path.join(REF, *a)


def test_setup():
    s = F.Setup.read(_p("INPUT", "setup"))
    assert (s.theta_min, s.theta_max, s.phi_min, s.phi_max) == (47.1, 55.9, 23.1, 42.9)
    assert (s.z_min, s.z_max, s.is_diss, s.model_type) == (5900000.0, 6371000.0, 0, 1)
    assert (s.nx_global, s.ny_global, s.nz_global, s.lpd) == (66, 108, 28, 4)
    assert (s.px, s.py, s.pz) == (3, 4, 4)
    assert (s.adjoint_flag, s.samp_ad) == (0, 15)
    assert s.ffd == "../DATA/OUTPUT/NEEDS_A_LOT_OF_SPACE/ADJOINT"
    # 69 x 112 x 32 elements split over 48 rank boxes
    boxes = F.make_boxes(s)
    assert len(boxes) == 48
    for axis, total in ((0, 69), (1, 112), (2, 32)):
        n = {}
        for b in boxes:
            lo, hi = [(b.nx_min, b.nx_max), (b.ny_min, b.ny_max), (b.nz_min, b.nz_max)][axis]
            n[b.multi[axis]] = hi - lo + 1
        assert sum(n.values()) == total, (axis, n)


def test_event():
    e = F.Event.read(_p("INPUT", "event_1"))
    assert (e.nt, e.dt) == (4000, 0.13)
    assert (e.xxs, e.yys, e.zzs, e.source_type) == (50.74, 41.04, 5000.0, 10)
    assert tuple(e.moment) == (1e16, 1e16, 1e16, 0.0, 0.0, 0.0)
    assert e.ofd == "../DATA/OUTPUT/1.8s/"
    assert (e.ssamp, e.output_displacement) == (15000, 0)
    assert F.read_event_list(_p("INPUT", "event_list")) == ["1"]


def test_relax():
    tau, D = F.read_relax(_p("INPUT", "relax"))
    np.testing.assert_array_equal(tau, np.array([1.7308, 14.3961, 22.9973], np.float32))
    np.testing.assert_array_equal(D, np.array([2.5100, 2.4354, 0.0879], np.float32))


def test_recfile():
    r = F.Receivers.read(_p("INPUT", "recfile_1"))
    assert len(r.names) == 320
    assert r.names[0] == "ADVT_.KO.___" and r.names[-1] == "ZEY__.TU.___"
    np.testing.assert_array_equal([r.colat[0], r.lon[0], r.depth[0]], [49.5668, 29.7383, 0.0])
    np.testing.assert_array_equal([r.colat[-1], r.lon[-1], r.depth[-1]], [51.7611, 26.5033, 0.0])


def test_headerless_stf_loses_its_first_four_samples():
    """INPUT/stf has no header, but the reader skips 4 records
    (SOURCE/ses3d_input.f90:419-422): the first sample used is the file's
    fifth line."""
    raw = [np.float32(l) for l in open(_p("INPUT", "stf")).read().split()]
    assert raw[0] == np.float32(-1.07622e-17) and len(raw) == 10001
    so = F.read_stf(_p("INPUT", "stf"), 4000)
    np.testing.assert_array_equal(so, np.array(raw[4:4004], np.float32))
    assert so[0] == np.float32(0.000126585)


def test_adjoint_source_files():
    locs = F.read_ad_srcfile(_p("ADJOINT", "1", "ad_srcfile"))
    # colatitude and longitude converted to radians in float32 (SOURCE/ses3d_input.f90:594-601)
    np.testing.assert_array_equal(locs, [[F.deg2rad32(np.float32(50.36)), F.deg2rad32(np.float32(27.88)), 1000.0]])
    rows = F.read_ad_src(_p("ADJOINT", "1", "ad_src_1"), 4000)
    assert rows.shape == (4000, 3)
    lines = open(_p("ADJOINT", "1", "ad_src_1")).read().splitlines()[4:4004]
    np.testing.assert_array_equal(rows, np.array([[np.float32(t) for t in l.split()] for l in lines]))
    assert np.abs(rows[:, 2]).max() == np.float32(1724967.17294) and not rows[:, :2].any()


def test_header_reads_skip_blank_records(tmp_path):
    """A list-directed header read of a character item passes over blank
    records (as amdflang's runtime does for the reference's reads): blank
    lines before or between header lines are not headers."""
    p = tmp_path / "ad_src_1"
    p.write_text("\n-- adjoint source --\n\n-- coords --\n50.0 20.0 0.0\n   \n-- stf --\n1.5 2.5 3.5\n\n4.0 5.0 6.0\n")
    np.testing.assert_array_equal(F.read_ad_src(str(p), 2), np.array([[1.5, 2.5, 3.5], [4.0, 5.0, 6.0]], np.float32))


def test_block_files():
    x = F.read_block_file(_p("MODELS", "MODELS_3D", "block_x"))
    y = F.read_block_file(_p("MODELS", "MODELS_3D", "block_y"))
    z = F.read_block_file(_p("MODELS", "MODELS_3D", "block_z"))
    assert [len(v) for v in x] == [37, 91] and [len(v) for v in y] == [81, 201] and [len(v) for v in z] == [80, 22]
    assert x[0][0] == 47.0 and x[0][1] == 47.25 and x[0][-1] == 56.0
    for e in x + y + z:
        assert np.all(np.diff(e) > 0)
