"""The benchmark's contract pieces that need no GPU: the refusal of a
multi-GPU run and the algorithmic byte counts the roofline uses (the per-point
figures: 178.9 B elastic at 24x24x12 and 326.5 B anelastic per whole
step; the element kernel's share without the velocity write and mass read)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_algorithmic_bytes():
    sys.path.insert(0, ROOT)
    import bench
    e = bench.algorithmic_bytes((24, 24, 12), False)
    a = bench.algorithmic_bytes((96, 96, 48), True)
    assert abs(e["step"] - 178.9) < 0.05 and abs(e["element"] - 170.4) < 0.05
    assert abs(a["step"] - 326.5) < 0.1 and abs(a["element"] - 318.2) < 0.05
    # rho_g = global nodes per GLL point
    assert abs(a["rho_g"] - (385 * 385 * 193) / (96 * 96 * 48 * 125)) < 1e-12
    # the implemented element kernel moves less than the reference algorithm
    assert bench.implemented_bytes((96, 96, 48), 3, True) < a["element"]


def test_small_gradient_workload_builds():
    """The 8x8x4 gradient workload of the batched-vs-plain legs builds on the
    host for 32 events: the mesh and model of 8x8x4 elements, one moment
    source and a receiver list inside the section per event, the first event
    on the 10x10 surface grid, the others with 10-400 receivers."""
    import numpy as np
    from ses3d_amd import configs
    w = configs.WORKLOADS["gradient_8x8x4"]
    assert w.nel == (8, 8, 4) and w.n_points == 8 * 8 * 4 * 125
    run = configs.build(w, seed=0, nt=20, n_events=32)
    assert (run.nx, run.ny, run.nz) == (7, 7, 3)
    assert run.model["mu"].size == w.n_points
    assert len(run.events) == 32
    counts = [len(e.rec32) for e in run.events]
    assert counts[0] == 100 and all(10 <= c <= 400 for c in counts[1:])
    for e in run.events:
        assert e.nt == 20 and np.all(np.isfinite(e.stf))
