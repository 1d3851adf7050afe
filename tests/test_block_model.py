"""Block models (ses3d_amd.block_model) as the reference's model tool
(TOOLS/models.py:157-314) reads, combines and writes them, on the
reference's own block files and through the readers of the projection and
perturbation tools."""
import os

import numpy as np
import pytest

from ses3d_amd import block_model as B, formats as F, generator as G

REF3D = "/root/reference/MODELS/MODELS_3D/"


def _small():
    bx = [np.array([40.0, 45.0, 50.5]), np.array([48.0, 52.0])]
    by = [np.array([10.0, 20.0]), np.array([12.0, 14.0, 16.0])]
    bz = [np.array([6000.0, 6200.0, 6371.0]), np.array([6300.0, 6371.0])]
    vals = [np.array([[[0.1, -0.25]], [[1.0 / 3.0, 2.0]]]), np.array([[[1e-20], [-7.5]]])]
    return bx, by, bz, vals


def test_write_is_the_tools_text(tmp_path):
    m = B.BlockModel.from_blocks(*_small())
    m.write(str(tmp_path) + "/", "dvsv")
    assert (tmp_path / "dvsv").read_text() == \
        "2\n4\n0.1\n-0.25\n0.333333333333\n2.0\n2\n1e-20\n-7.5\n"


def test_read_write_round_trip_and_arithmetic(tmp_path):
    d = str(tmp_path) + "/"
    m = B.BlockModel.from_blocks(*_small())
    m.write_blocks(d)
    m.write(d, "dvsv")
    r = B.ses3d_model()
    r.read(d, "dvsv")
    assert r.nsubvol == 2
    for a, b in zip(m.m, r.m):
        np.testing.assert_array_equal(a.colat, b.colat)
        np.testing.assert_array_equal(b.lat, 90.0 - a.colat)
        np.testing.assert_array_equal(a.lon, b.lon)
        np.testing.assert_array_equal(a.r, b.r)
        np.testing.assert_array_equal(b.v, np.vectorize(lambda x: float(B.py2_str(x)))(a.v))
    u = m + (-0.5) * r
    for a, b, c in zip(m.m, r.m, u.m):
        np.testing.assert_array_equal(c.v, a.v + (-0.5) * b.v)
        np.testing.assert_array_equal(c.colat, a.colat)
    # the operands are left as they were
    np.testing.assert_array_equal(m.m[0].v, _small()[3][0])
    with pytest.raises(ValueError):
        m + B.BlockModel.from_blocks([np.array([1.0, 2.0])], [np.array([1.0, 2.0])], [np.array([1.0, 2.0])],
                                     [np.zeros(1)])


@pytest.mark.skipif(not os.path.isdir(REF3D), reason="reference tree absent")
def test_reference_block_files(tmp_path):
    """The reference's MODELS_3D/block_{x,y,z} (two subvolumes of 36x80x79 and
    90x200x21 blocks) with a value file in the tool's layout."""
    bx, by, bz = (F.read_block_file(REF3D + "block_" + a) for a in "xyz")
    rng = np.random.default_rng(4)
    vals = [np.round(rng.normal(0, 0.05, (len(x) - 1) * (len(y) - 1) * (len(z) - 1)), 6)
            for x, y, z in zip(bx, by, bz)]
    assert [v.size for v in vals] == [36 * 80 * 79, 90 * 200 * 21]
    d = str(tmp_path) + "/"
    for a in "xyz":
        with open(REF3D + "block_" + a) as src, open(d + "block_" + a, "w") as dst:
            dst.write(src.read())
    with open(d + "dvsv", "w") as fh:
        fh.write("2\n" + "".join(f"{v.size}\n" + "".join(f"{float(x)!r}\n" for x in v) for v in vals))
    m = B.BlockModel()
    m.read(d, "dvsv")
    assert [s.v.shape for s in m.m] == [(36, 80, 79), (90, 200, 21)]
    for s, x, y, z, v in zip(m.m, bx, by, bz, vals):
        np.testing.assert_array_equal(s.colat, x)
        np.testing.assert_array_equal(s.lon, y)
        np.testing.assert_array_equal(s.r, z)
        np.testing.assert_array_equal(s.v.ravel(), v)
        assert s.v[0, 0, 1] == v[1] and s.v[0, 1, 0] == v[z.size - 1]     # radius fastest


def test_update_feeds_the_perturbation_tool(tmp_path):
    """model + step * projected gradient, written by the tool's writer, is
    what the perturbation tool reads (float32 of the printed values) and
    perturbs the run tree with."""
    bx, by, bz, _ = _small()
    grad = [np.array([1.5e-3, -2.0e-3, 0.5e-3, 0.0]), np.array([3.0e-3, -1.0e-3])]
    d = str(tmp_path / "MODELS" / "MODELS_3D") + "/"
    os.makedirs(d)
    F.write_projected(d + "gradient_csv", grad)            # as project_kernel writes it
    g = B.BlockModel.from_blocks(bx, by, bz, grad)
    g.write_blocks(d)
    gr = B.BlockModel()
    gr.read(d, "gradient_csv")
    zero = 0.0 * gr
    step = zero + (-20.0) * gr
    for name in ("dvsv", "dvsh", "dvp", "drho"):
        (step if name == "dvsv" else zero).write(d, name)
    got = F.read_projected(d + "dvsv")
    for k, sub in enumerate(step.m):
        np.testing.assert_array_equal(got[k], np.float32(sub.v.ravel()))
    subs = G.read_subvolumes(d)
    assert len(subs) == 2
    np.testing.assert_array_equal(subs[0].dvsv, got[0])
    np.testing.assert_array_equal(subs[1].dvsh, np.zeros(2, np.float32))
