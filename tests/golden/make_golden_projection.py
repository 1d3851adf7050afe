"""Reference block-grid projections (TOOLS/SOURCE/project_kernel.f90) of the
reference's own kernels in tests/golden/anelastic_adjoint.npz.

Runs oracle/_ref/project_kernel.exe (built by ``make -C oracle ref`` from the
reference's TOOLS/SOURCE/project_kernel.f90 and SOURCE/ses3d_modules.f90) on
the fixture's float32 kernels written as the reference writes kernel files,
for one block grid that cuts elements at arbitrary positions (colatitude
40-60 deg in 11 edges, longitude 10-30 deg in 9, radius 5371-6371 km in 11),
and stores the four projected fields with the block edges.

    python tests/golden/make_golden_projection.py
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

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import formats as F  # noqa: E402
from tests.golden import fixture_case as FC  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FULL = (23, 28, 8, 5, 5, 5)
KERNELS = ("grad_rho", "grad_cp", "grad_csh", "grad_csv")
OUTPUTS = ("gradient_csh", "gradient_csv", "gradient_cp", "gradient_rho")
EDGES = {"bx": np.linspace(40.0, 60.0, 11), "by": np.linspace(10.0, 30.0, 9), "bz": np.linspace(5371.0, 6371.0, 11)}


def main() -> str:
    fx = dict(np.load(os.path.join(HERE, "anelastic_adjoint.npz")))
    root = tempfile.mkdtemp(prefix="golden_projection_")
    try:
        FC.write_tree(fx, root)
        kd, od, m3 = (os.path.join(root, d) for d in ("KERNELS", "OUT", os.path.join("MODELS", "MODELS_3D")))
        for d in (kd, od, m3):
            os.makedirs(d, exist_ok=True)
        for g in KERNELS:
            a = np.zeros(FULL, np.float32)
            k = fx[g]
            a[: k.shape[0], : k.shape[1], : k.shape[2]] = k
            F.write_field(os.path.join(kd, g + "_0"), a)
        for ax, name in (("bx", "block_x"), ("by", "block_y"), ("bz", "block_z")):
            F.write_block_file(os.path.join(m3, name), [EDGES[ax]])
        exe = os.path.join(ROOT, "oracle", "_ref", "project_kernel.exe")
        # list-directed reads stop at '/', so the directories are quoted
        r = subprocess.run([exe], cwd=os.path.join(root, "MAIN"), input="0\n'../KERNELS/'\n'../OUT/'\n",
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        shape = tuple(len(EDGES[a]) - 1 for a in ("bx", "by", "bz"))
        out = {o: F.read_projected(os.path.join(od, o))[0].reshape(shape) for o in OUTPUTS}
    finally:
        shutil.rmtree(root, ignore_errors=True)
    path = os.path.join(HERE, "projection_anelastic_adjoint.npz")
    np.savez_compressed(path, **out, **EDGES)
    print(f"wrote {path}: {shape} blocks, max |gradient_csh| {np.abs(out['gradient_csh']).max():.3e}")
    return path


if __name__ == "__main__":
    main()
