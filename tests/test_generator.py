"""Model generation against the reference's own model tools.

The fixture (tests/golden/make_golden_models.py) holds what
MODELS/SOURCE/generate_models.f90 with the 1D models of MODELS/MODELS_1D and
MODELS/SOURCE/add_perturbation.f90, compiled from their sources, wrote for
five setups (model types 1, 2, 3, 4, 7; one to four rank boxes; sections
reaching from the surface into the inner core) and for a two-subvolume block
perturbation of the PREM case.  ses3d_amd.generator must reproduce the rank
boxes and every model file bit for bit, through its functions and through
the run-tree files it writes."""
import os

import numpy as np
import pytest

from ses3d_amd import formats as F, generator as G

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model_tools.npz"))
CASES = ("homogeneous", "prem", "homogeneous_q", "eumod", "ak135")


def _setup(name: str) -> F.Setup:
    a = FIX[f"setup_{name}"]
    return F.Setup(theta_min=float(a[0]), theta_max=float(a[1]), phi_min=float(a[2]), phi_max=float(a[3]),
                   z_min=float(a[4]), z_max=float(a[5]), is_diss=int(a[6]), model_type=int(a[7]),
                   nx_global=int(a[8]), ny_global=int(a[9]), nz_global=int(a[10]),
                   px=int(a[11]), py=int(a[12]), pz=int(a[13]))


def _fields(setup: F.Setup):
    return F.MODEL_FIELDS + (("Q",) if setup.is_diss else ())


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _ref_boxes(name: str):
    idx, lim = FIX[f"boxes_{name}"], FIX[f"bounds_{name}"]
    return [F.Box(r + 1, (0, 0, 0), *[int(v) for v in idx[r]], *[float(v) for v in lim[r]])
            for r in range(len(idx))]


@pytest.mark.parametrize("name", CASES)
def test_generate_equals_reference_tool(name):
    s = _setup(name)
    gen = G.generate(s)
    ref_boxes = _ref_boxes(name)
    assert len(gen) == len(ref_boxes) == s.px * s.py * s.pz
    for r, ((b, fields), rb) in enumerate(zip(gen, ref_boxes)):
        assert (b.nx_min, b.nx_max, b.ny_min, b.ny_max, b.nz_min, b.nz_max) == \
            (rb.nx_min, rb.nx_max, rb.ny_min, rb.ny_max, rb.nz_min, rb.nz_max)
        # the boxfile's bounds as the solver reads them back
        assert np.array_equal(np.float32([b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax]),
                              FIX[f"bounds_{name}"][r]), f"box {r}"
        assert sorted(fields) == sorted(_fields(s))
        for f in _fields(s):
            ref = FIX[f"{name}_{f}_{r}"]
            assert fields[f].shape == ref.shape
            assert np.array_equal(_bits(fields[f]), _bits(ref)), f"{name} rank {r} {f}: " \
                f"max |diff| {np.abs(fields[f].astype(np.float64) - ref).max()}"


def test_reference_example_setup():
    """The reference's own INPUT/setup (model type 1, 69 x 112 x 32 elements on
    3 x 4 x 4 rank boxes): every box and bound of its boxfile, rank 0's fields."""
    s = _setup("example")
    gen = G.generate(s)
    assert len(gen) == 48 == len(FIX["boxes_example"])
    for r, (b, fields) in enumerate(gen):
        assert [b.nx_min, b.nx_max, b.ny_min, b.ny_max, b.nz_min, b.nz_max] == list(FIX["boxes_example"][r])
        assert np.array_equal(np.float32([b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax]),
                              FIX["bounds_example"][r]), f"box {r}"
    for f in F.MODEL_FIELDS:
        assert np.array_equal(_bits(gen[0][1][f]), _bits(FIX[f"example_{f}_0"])), f


def _subvolumes():
    subs, k = [], 0
    while f"pert_bx_{k}" in FIX.files:
        subs.append(G.Subvolume(*(FIX[f"pert_{n}_{k}"] for n in ("bx", "by", "bz", "dvsh", "dvsv", "dvp", "drho"))))
        k += 1
    return subs


def test_perturb_equals_reference_tool():
    s = _setup("perturbed")
    subs = _subvolumes()
    assert len(subs) == 2
    for r, b in enumerate(_ref_boxes("perturbed")):
        model = {f: FIX[f"prem_{f}_{r}"] for f in _fields(s)}
        new = G.perturb(b, model, subs)
        for f in _fields(s):
            assert np.array_equal(_bits(new[f]), _bits(FIX[f"perturbed_{f}_{r}"])), f"rank {r} {f}"


def test_run_tree_files_equal_reference_tool(tmp_path):
    """write_tree then perturb_tree on a run tree: the files it leaves are the
    ones the two reference tools leave (through the boxfile and records)."""
    s = _setup("prem")
    root = str(tmp_path)
    os.makedirs(os.path.join(root, "INPUT"))
    s.write(os.path.join(root, "INPUT", "setup"))
    G.write_tree(root)
    mdir = os.path.join(root, "MODELS", "MODELS")
    boxes = F.read_boxfile(os.path.join(mdir, "boxfile"))
    assert np.array_equal(np.float32([[b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax] for b in boxes]),
                          FIX["bounds_prem"])
    for r, b in enumerate(boxes):
        shape = (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
        for f in _fields(s):
            assert np.array_equal(_bits(F.read_field(os.path.join(mdir, f"{f}{r}"), shape)),
                                  _bits(FIX[f"prem_{f}_{r}"])), f"written rank {r} {f}"
    # the block files as the perturbation tool reads them
    d3 = os.path.join(root, "MODELS", "MODELS_3D")
    os.makedirs(d3)
    subs = _subvolumes()
    for a in "xyz":
        F.write_block_file(os.path.join(d3, f"block_{a}"), [getattr(v, f"b{a}") for v in subs])
    for n in ("dvsh", "dvsv", "dvp", "drho"):
        F.write_projected(os.path.join(d3, n), [getattr(v, n) for v in subs])
    G.perturb_tree(root)
    for r, b in enumerate(boxes):
        shape = (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
        for f in F.MODEL_FIELDS:
            assert np.array_equal(_bits(F.read_field(os.path.join(mdir, f"{f}{r}"), shape)),
                                  _bits(FIX[f"perturbed_{f}_{r}"])), f"perturbed rank {r} {f}"


def test_integer_power_is_left_to_right_product():
    """powi(x, n) multiplies as the reference's build evaluates x**n: a
    left-to-right product, which for n = 4 rounds differently from squaring
    twice (the eumod_bg fixture's Q at r = 6164168.5 m tells them apart)."""
    x = np.float32(1.0000001) + np.arange(1000, dtype=np.float32) * np.float32(0.0137)
    assert np.array_equal(G.powi(x, 2), x * x)
    assert np.array_equal(G.powi(x, 4), ((x * x) * x) * x)
    assert np.array_equal(G.powi(x, 5), (((x * x) * x) * x) * x)
    assert not np.array_equal(((x * x) * x) * x, (x * x) * (x * x))
