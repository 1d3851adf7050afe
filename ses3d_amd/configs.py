"""The benchmark workloads, built in memory as loaded run inputs.

Each builder returns a :class:`ses3d_amd.inputs.RunInput` equal to what
``load_run`` would return for the run tree the same case writes, so the GPU
path sees exactly the float32 values the reference would read.

Workloads (spherical section colat 40-60 deg, lon 10-30 deg, r 5371-6371 km):
  elastic_24x24x12     1D PREM-like, is_diss=0, dt=0.2, nt=2000, 20 receivers
  anelastic_96x96x48   + Gaussian random 3D (200 km, 3% rms), Q_mu 80-600, pml=3,
                       dt=0.1, nt=4000, 200 receivers
  gradient_48x48x24    anelastic, true = 1D + 2% blobs, 100 receivers, dt=0.15, nt=3000
  gradient_8x8x4       the same physics on 8x8x4 elements (one event per step is launch-bound)
  gradient_16x16x8     the same on 16x16x8 elements
  batched_64x64x32     32 events, anelastic, 10-400 receivers each, nt=3000
  + EXTRA_SECTIONS     six more anelastic sections of uneven element counts (one
                       event each) advanced with the 32 events in one launch per step
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import formats as F, inputs as I, models as M


@dataclass
class Workload:
    name: str
    nel: tuple
    is_diss: bool
    pml: int
    dt: float
    nt: int
    n_rec: int
    perturbation: str          # "none" | "grf" | "blobs"
    period: float = 20.0
    n_events: int = 1

    @property
    def n_points(self) -> int:
        return int(np.prod(self.nel)) * 125


WORKLOADS = {
    "elastic_24x24x12": Workload("elastic_24x24x12", (24, 24, 12), False, 2, 0.2, 2000, 20, "none"),
    "anelastic_96x96x48": Workload("anelastic_96x96x48", (96, 96, 48), True, 3, 0.1, 4000, 200, "grf"),
    "gradient_48x48x24": Workload("gradient_48x48x24", (48, 48, 24), True, 2, 0.15, 3000, 100, "blobs"),
    # a section small enough that one event's step is launch-bound (batching pays)
    "gradient_8x8x4": Workload("gradient_8x8x4", (8, 8, 4), True, 2, 0.15, 3000, 100, "blobs"),
    "gradient_16x16x8": Workload("gradient_16x16x8", (16, 16, 8), True, 2, 0.15, 3000, 100, "blobs"),
    "batched_64x64x32": Workload("batched_64x64x32", (64, 64, 32), True, 2, 0.1, 3000, 200, "grf",
                                 n_events=32),
}


# the uneven sections batched with the 64x64x32 events (elements per axis)
EXTRA_SECTIONS = [(16, 16, 16), (24, 24, 12), (32, 32, 16), (40, 40, 20), (48, 48, 24), (56, 56, 28)]


def section_workload(nel, nt: int = 3000, seed: int = 0) -> Workload:
    """One extra anelastic section (3D random model, dt=0.1, pml=2) with a
    seeded receiver count in 10..400."""
    rng = np.random.default_rng(1000 + seed)
    return Workload(f"section_{nel[0]}x{nel[1]}x{nel[2]}", tuple(nel), True, 2, 0.1, nt,
                    int(rng.integers(10, 401)), "grf")


def _event(nt, dt, rng, colat, lon, depth, rec: F.Receivers, stf, name="1") -> I.EventInput:
    mom = M.random_moment_tensor(rng)
    ev = F.Event(nt=nt, dt=dt, xxs=colat, yys=lon, zzs=depth, source_type=10, moment=tuple(mom))
    return I.EventInput(name, ev, nt, F.f32(repr(dt)), 10, F.deg2rad32(F.f32(repr(colat))),
                        F.deg2rad32(F.f32(repr(lon))), F.f32(repr(depth)),
                        np.array([F.f32(repr(float(m))) for m in mom], np.float32), stf[:nt].copy(),
                        rec, rec.loc32())


def build(w: Workload, seed: int = 0, nt: Optional[int] = None, n_events: Optional[int] = None,
          model: Optional[dict] = None) -> I.RunInput:
    """RunInput for workload ``w`` (optionally truncated to ``nt`` steps)."""
    nt = nt or w.nt
    n_events = n_events or w.n_events
    rng = np.random.default_rng(seed)
    setup = M.section_setup(*w.nel, is_diss=w.is_diss)
    box = F.make_boxes(setup)[0]
    if model is None:
        pert = None
        if w.perturbation == "grf":
            pert = M.gaussian_random_field(box, seed + 1, corr_len=200e3, rms=0.03)
        elif w.perturbation == "blobs":
            pert = M.gaussian_blobs(box, seed + 1, radius=300e3, amp=0.02)
        model = M.make_model(box, is_diss=w.is_diss, perturbation=pert)
    tau, D = (np.array([1.7308, 14.3961, 22.9973], np.float32), np.array([2.51, 2.4354, 0.0879], np.float32))
    run = I.RunInput(None, setup, box, F.f32(repr(box.xmin)), F.f32(repr(box.xmax)), F.f32(repr(box.ymin)),
                     F.f32(repr(box.ymax)), F.f32(repr(box.zmin)), F.f32(repr(box.zmax)), setup.zmax_global,
                     int(w.is_diss), 0, 1, tau, D, model)
    stf = M.ricker(nt, w.dt, w.period)
    for k in range(n_events):
        if k == 0:
            colat, lon, depth = 50.0, 20.0, 10000.0
            # exactly n_rec receivers on the squarest grid: side = largest divisor <= sqrt(n_rec)
            side = max(d for d in range(1, int(np.sqrt(w.n_rec)) + 1) if w.n_rec % d == 0)
            rec = M.surface_receiver_grid(setup, side, w.n_rec // side, pml=w.pml)
        else:
            colat, lon = rng.uniform(42.0, 58.0), rng.uniform(12.0, 28.0)
            depth = rng.uniform(5e3, 600e3)
            rec = M.random_receivers(setup, int(rng.integers(10, 401)), rng, pml=w.pml)
        run.events.append(_event(nt, w.dt, rng, round(float(colat), 4), round(float(lon), 4),
                                 round(float(depth), 1), rec, stf, name=str(k + 1)))
    return run
