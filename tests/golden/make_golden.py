"""Write the golden fixtures tests/golden/<case>.npz from the reference solver.

Runs the reference Fortran binary oracle/_ref/ses3d.exe (compiled from the
reference sources by ``make -C oracle ref``) on each compact case of
fixture_case.CASES and stores the case inputs together with the reference's
outputs as read back from its own output files:
  seismograms  (3, nr, nt)        from <ofd>/<station>.x/.y/.z
  vx, vy, vz   (nx+1,...,5,5,5)   snapshot files vx_0_<nt> (velocity gathered at step nt)
  grad_*       (nx+1,...,5,5,5)   kernel files (adjoint cases)
  grad_Q_mu    (nx+1,...,5,5,5)   anelastic adjoint cases: the shear-Q kernel of
                                  SOURCE/ses3d_grad_Q.f90, written by the grad_Q
                                  build (oracle/_ref/ses3d_grad_q.exe), which runs
                                  the adjoint pass of every adjoint case
  history_*    stored forward fields of the first and last stored steps
               (1D-model adjoint cases)
  saving_vector

    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as PO  # noqa: E402
from ses3d_amd import formats as F, inputs as I  # noqa: E402
from tests.golden import fixture_case as FC  # noqa: E402

# the reference's static array extents (SOURCE/ses3d_modules.f90:13-15)
FULL = (23, 28, 8, 5, 5, 5)
HERE = os.path.dirname(os.path.abspath(__file__))


def _crop(a, run):
    return a[: run.nx + 1, : run.ny + 1, : run.nz + 1]


def make(name: str) -> str:
    fx = FC.compact_inputs(name)
    root = tempfile.mkdtemp(prefix=f"golden_{name}_")
    try:
        FC.write_tree(fx, root)
        r = PO.run_ref(root)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        run = I.load_run(root)
        e = run.events[0]
        ofd = os.path.join(root, "MAIN", e.event.ofd)
        nt = e.nt
        seis = np.stack([F.read_seismogram(ofd, n)["trace"].astype(np.float32) for n in e.receivers.names], 1)
        out = dict(fx)
        out["seismograms"] = seis
        for c in ("vx", "vy", "vz"):
            out[c] = _crop(F.read_field(os.path.join(ofd, f"{c}_0_{nt}"), FULL), run)
        if int(fx["adjoint"]):
            ffd = run.ffd("1")
            sv = F.read_saving_vector(os.path.join(ffd, "saving_vector_0"), nt)
            out["saving_vector"] = sv
            stored = [int(s) for s in sv if s > 0]
            for tag, step in (() if name in FC.MODEL_3D else (("first", stored[0]), ("last", stored[-1]))):
                recs = []
                for fname in ("vx", "vy", "vz", "exx", "eyy", "ezz", "exy", "exz", "eyz"):
                    data = F.read_records(os.path.join(ffd, f"{fname}_0"))
                    rec = np.frombuffer(data[stored.index(step)], np.float32).reshape(FULL, order="F")
                    recs.append(_crop(rec, run))
                out[f"history_{tag}"] = np.stack(recs)
                out[f"history_{tag}_step"] = step
            FC.write_adjoint(fx, root, seis)
            # the grad_Q build: elastic kernels as ses3d_grad.f90 forms them, plus grad_Q_mu
            r = PO.run_ref(root, exe=PO.REF_GRADQ_EXE)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
            for g in ("grad_rho", "grad_cp", "grad_csh", "grad_csv") + (("grad_Q_mu",) if int(fx["is_diss"]) else ()):
                out[g] = _crop(F.read_field(os.path.join(ofd, f"{g}_0"), FULL), run)
            out["adjoint_vx"] = _crop(F.read_field(os.path.join(ofd, f"vx_0_{nt}"), FULL), run)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        return path
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    if not PO.build_ref():
        sys.exit("reference binary unavailable (needs /root/reference and amdflang)")
    for n in (sys.argv[1:] or [c for c in FC.CASES if c not in FC.MULTIBOX]):
        print(make(n))
