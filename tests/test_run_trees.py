"""Run trees split into rank boxes load as the one section they tile
(ses3d_amd.inputs.load_run): the stitched model equals the single-box tree's
model bit for bit, the bounds are the section's, and each box's part of a
section field is the record that box's rank reads (SOURCE/ses3d_input.f90:
49-170, 472-501; boxes as MODELS/SOURCE/generate_models.f90:200-282 lays
them out)."""
import os

import numpy as np
import pytest

from ses3d_amd import formats as F, inputs as I
from tests.golden import fixture_case as FC

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("decomp", [(2, 1, 2), (3, 2, 1), (1, 5, 4), (6, 1, 1)])
def test_split_tree_loads_as_the_section(tmp_path, decomp):
    fx = dict(np.load(os.path.join(GOLDEN, "multibox_adjoint.npz")))
    one, split = str(tmp_path / "one"), str(tmp_path / "split")
    FC.write_tree(fx, one)
    FC.write_tree(fx, split, decomp=decomp)
    a, b = I.load_run(one), I.load_run(split)
    assert len(b.boxes) == int(np.prod(decomp))
    assert (a.nx, a.ny, a.nz) == (b.nx, b.ny, b.nz)
    for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "zmax_global"):
        assert getattr(a, k) == getattr(b, k), k
    assert sorted(a.model) == sorted(b.model)
    for k in a.model:
        np.testing.assert_array_equal(a.model[k], b.model[k], err_msg=k)
    # every element belongs to exactly one box
    cover = np.zeros(a.shape_local[:3], int)
    for sl in b.box_slices:
        cover[sl] += 1
    assert (cover == 1).all()
    # each box's part is its rank's model record
    mdir = os.path.join(split, "MODELS", "MODELS")
    for r, box in enumerate(b.boxes):
        rec = F.read_model(mdir, (box.nx + 1, box.ny + 1, box.nz + 1, 5, 5, 5), box.index - 1, False)
        np.testing.assert_array_equal(b.box_part(b.model["mu"], r), rec["mu"])


def test_boxes_that_do_not_tile_are_refused(tmp_path):
    fx = dict(np.load(os.path.join(GOLDEN, "multibox_adjoint.npz")))
    FC.write_tree(fx, str(tmp_path), decomp=(2, 1, 2))
    boxes = F.read_boxfile(os.path.join(str(tmp_path), "MODELS", "MODELS", "boxfile"))
    with pytest.raises(ValueError):
        I.section_box(boxes[:-1])


def test_multirank_fixture_agrees_with_the_single_rank_run():
    """The fixture's 4-rank reference run (a build sized for its 3x5x2-element
    rank boxes) differs from the reference's single-rank run only by the
    rounding of per-rank geometry and face sums; a build with larger static
    extents would leave every rank-face node with half its mass (face
    velocities doubled, a distance of order 1e-3 here)."""
    fx = dict(np.load(os.path.join(GOLDEN, "multibox_adjoint.npz")))
    assert tuple(int(v) for v in fx["mr_decomp"]) == (2, 1, 2)
    assert float(fx["mr_rel_seismograms"]) < 1e-7
    rel = np.linalg.norm((fx["mr_seismograms"] - fx["seismograms"]).astype(np.float64)) / \
        np.linalg.norm(fx["seismograms"].astype(np.float64))
    assert rel < 1e-7
