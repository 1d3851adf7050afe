"""Golden fixtures of the reference's model tools: MODELS/SOURCE/generate_models.f90
(with the 1D models of MODELS/MODELS_1D) and MODELS/SOURCE/add_perturbation.f90,
built from their own sources under oracle/_ref (`make -C oracle ref`) and run
under the image's MPICH on as many ranks as each setup has boxes.

For every case: the setup it ran on (setup_<case>: theta/phi/z ranges,
is_diss, model_type, nx/ny/nz_global, px, py, pz), the boxfile as the solver
parses it (boxes_<case>: per box nx_min nx_max ny_min ny_max nz_min nz_max;
bounds_<case>: per box xmin xmax ymin ymax zmin zmax as float32) and every
rank's model fields (<case>_<field>_<rank>, float32, reference layout).
The perturbation case also stores its block edges and perturbations, and the
reference's own example setup (INPUT/setup, 48 rank boxes) its boxfile and
rank 0's fields.

    python tests/golden/make_golden_models.py
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ses3d_amd import formats as F  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref")
GM = os.path.join(REF, "generate_models.exe")
AP = os.path.join(REF, "add_perturbation.exe")
OUT = os.path.join(HERE, "model_tools.npz")

# (name, setup fields); the sections reach from the surface into the core so
# that the profiles' layers and discontinuities are crossed by GLL points
CASES = [
    ("homogeneous", dict(model_type=1, is_diss=0, nx_global=2, ny_global=2, nz_global=1, px=1, py=1, pz=1)),
    ("prem", dict(model_type=2, is_diss=1, z_min=3000000.0, nx_global=4, ny_global=2, nz_global=5,
                  px=2, py=2, pz=1)),
    ("homogeneous_q", dict(model_type=3, is_diss=1, z_min=5000000.0, nx_global=2, ny_global=4, nz_global=3,
                           px=1, py=2, pz=1)),
    ("eumod", dict(model_type=4, is_diss=1, nx_global=3, ny_global=2, nz_global=2, px=2, py=1, pz=2)),
    ("ak135", dict(model_type=7, is_diss=1, z_min=1000000.0, nx_global=2, ny_global=2, nz_global=6,
                   px=1, py=1, pz=2)),
]
FIELDS = ("rhoinv", "mu", "lambda", "A", "B", "C", "Q")

# 3D perturbation of the PREM case: two subvolumes of blocks (colatitude and
# longitude edges in degrees, radius edges in km), seeded values in km/s and g/cm^3
PERT_BASE = "prem"


def _pert_blocks(rng):
    subs = []
    for edges in (([38.0, 45.0, 50.5, 61.0], [9.0, 16.0, 24.0, 31.0], [2900.0, 4000.0, 5200.0, 6000.0, 6400.0]),
                  ([49.0, 55.0], [12.0, 20.0, 26.0], [5800.0, 6100.0, 6371.0])):
        bx, by, bz = (np.array(e) for e in edges)
        n = (len(bx) - 1) * (len(by) - 1) * (len(bz) - 1)
        vals = {k: np.round(rng.normal(0.0, 0.1, n), 4) for k in ("dvsh", "dvsv", "dvp", "drho")}
        subs.append((bx, by, bz, vals))
    return subs


def _write_pert_files(d3, subs):
    os.makedirs(d3, exist_ok=True)
    for axis, idx in (("x", 0), ("y", 1), ("z", 2)):
        with open(os.path.join(d3, f"block_{axis}"), "w") as fh:
            fh.write(f"{len(subs)}\n")
            for s in subs:
                fh.write(f"{len(s[idx])}\n" + "".join(f"{float(v)!r}\n" for v in s[idx]))
    for name in ("dvsh", "dvsv", "dvp", "drho"):
        with open(os.path.join(d3, name), "w") as fh:
            fh.write(f"{len(subs)}\n")
            for s in subs:
                fh.write(f"{len(s[3][name])}\n" + "".join(f"{float(v)!r}\n" for v in s[3][name]))


def _tree(setup: F.Setup) -> str:
    root = tempfile.mkdtemp(prefix="golden_models_")
    for d in ("INPUT", "MODELS/MAIN", "MODELS/MODELS", "MODELS/MODELS_3D", "DATA/COORDINATES"):
        os.makedirs(os.path.join(root, d), exist_ok=True)
    setup.write(os.path.join(root, "INPUT", "setup"))
    return root


def _run(exe: str, root: str, nprocs: int) -> None:
    cmd = [exe] if nprocs == 1 else [shutil.which("mpiexec") or "/opt/conda/bin/mpiexec", "-n", str(nprocs), exe]
    r = subprocess.run(cmd, cwd=os.path.join(root, "MODELS", "MAIN"), capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def _collect(out: dict, name: str, setup: F.Setup, root: str) -> None:
    boxes = F.read_boxfile(os.path.join(root, "MODELS", "MODELS", "boxfile"))
    out[f"boxes_{name}"] = np.array([[b.nx_min, b.nx_max, b.ny_min, b.ny_max, b.nz_min, b.nz_max]
                                     for b in boxes], np.int32)
    out[f"bounds_{name}"] = np.array([[b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax] for b in boxes], np.float32)
    for r, b in enumerate(boxes):
        shape = (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
        for f in FIELDS:
            if f == "Q" and not setup.is_diss:
                continue
            out[f"{name}_{f}_{r}"] = F.read_field(os.path.join(root, "MODELS", "MODELS", f"{f}{r}"), shape)


def _setup_array(s: F.Setup) -> np.ndarray:
    return np.array([s.theta_min, s.theta_max, s.phi_min, s.phi_max, s.z_min, s.z_max, s.is_diss, s.model_type,
                     s.nx_global, s.ny_global, s.nz_global, s.px, s.py, s.pz], np.float64)


def _example(out: dict) -> None:
    """The reference's own example setup (INPUT/setup: model type 1 on 3 x 4 x 4
    rank boxes): its boxfile, and rank 0's model fields."""
    setup = F.Setup.read("/root/reference/INPUT/setup")
    root = _tree(setup)
    try:
        _run(GM, root, setup.px * setup.py * setup.pz)
        out["setup_example"] = _setup_array(setup)
        boxes = F.read_boxfile(os.path.join(root, "MODELS", "MODELS", "boxfile"))
        out["boxes_example"] = np.array([[b.nx_min, b.nx_max, b.ny_min, b.ny_max, b.nz_min, b.nz_max]
                                         for b in boxes], np.int32)
        out["bounds_example"] = np.array([[b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax] for b in boxes],
                                         np.float32)
        b = boxes[0]
        for f in F.MODEL_FIELDS:
            out[f"example_{f}_0"] = F.read_field(os.path.join(root, "MODELS", "MODELS", f"{f}0"),
                                                 (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5))
    finally:
        shutil.rmtree(root, ignore_errors=True)


def make() -> str:
    if not (os.path.exists(GM) and os.path.exists(AP)):
        raise RuntimeError("build the reference model tools first: make -C oracle ref")
    out = {}
    _example(out)
    for name, kw in CASES:
        setup = F.Setup(**kw)
        root = _tree(setup)
        try:
            _run(GM, root, setup.px * setup.py * setup.pz)
            out[f"setup_{name}"] = _setup_array(setup)
            _collect(out, name, setup, root)
            if name == PERT_BASE:
                subs = _pert_blocks(np.random.default_rng(7))
                _write_pert_files(os.path.join(root, "MODELS", "MODELS_3D"), subs)
                _run(AP, root, setup.px * setup.py * setup.pz)
                _collect(out, "perturbed", setup, root)
                out["setup_perturbed"] = _setup_array(setup)
                for k, (bx, by, bz, vals) in enumerate(subs):
                    out[f"pert_bx_{k}"], out[f"pert_by_{k}"], out[f"pert_bz_{k}"] = bx, by, bz
                    for vn, v in vals.items():
                        out[f"pert_{vn}_{k}"] = v
        finally:
            shutil.rmtree(root, ignore_errors=True)
    np.savez_compressed(OUT, **out)
    return OUT


if __name__ == "__main__":
    print(make())
