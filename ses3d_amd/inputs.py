"""Load a SES3D run tree the way ``ses3d_input`` does.

A tree split into rank boxes (px*py*pz of them, each with its own model
records) is loaded as the one section the boxes tile: one GPU holds it whole.

Produces the float32 values the reference holds in module state after
SOURCE/ses3d_input.f90:11-627: box bounds from the boxfile (rad, rad, m),
``zmax_global`` from the setup file, the source position converted to
radians with the reference's float32 ``pi`` (:319-320), receiver and adjoint
source positions likewise (:541-542, :600-601), the STF after its 4 header
lines (:419-428), the relaxation parameters (:338-353) and the model fields
(:472-499).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import formats as F


@dataclass
class EventInput:
    name: str
    event: F.Event
    nt: int
    dt: np.float32
    source_type: int
    xxs: np.float32          # rad
    yys: np.float32          # rad
    zzs: np.float32          # m depth
    moment: np.ndarray       # (6,) float32
    stf: np.ndarray          # (nt,) float32
    receivers: Optional[F.Receivers] = None
    rec32: Optional[np.ndarray] = None       # (nr,3) rad, rad, m
    ad_loc32: Optional[np.ndarray] = None    # (n,3)
    ad_stf: Optional[np.ndarray] = None      # (n, nt, 3) rows as stored (reversed time)
    saving_vector: Optional[np.ndarray] = None


@dataclass
class RunInput:
    root: str
    setup: F.Setup
    box: F.Box
    xmin: np.float32
    xmax: np.float32
    ymin: np.float32
    ymax: np.float32
    zmin: np.float32
    zmax: np.float32
    zmax_global: np.float32
    is_diss: int
    adjoint_flag: int
    samp_ad: int
    tau_p: np.ndarray
    D_p: np.ndarray
    model: dict
    events: List[EventInput] = field(default_factory=list)
    # the rank boxes of the boxfile and each box's element slices in the section
    boxes: List[F.Box] = field(default_factory=list)
    box_slices: List[tuple] = field(default_factory=list)

    def box_part(self, a: np.ndarray, k: int) -> np.ndarray:
        """Rank box k's part of a section field in reference layout (the
        extent of the field files that box's rank reads and writes)."""
        return np.asfortranarray(a[self.box_slices[k]])

    @property
    def nx(self) -> int:
        return self.box.nx

    @property
    def ny(self) -> int:
        return self.box.ny

    @property
    def nz(self) -> int:
        return self.box.nz

    @property
    def shape_local(self):
        return (self.nx + 1, self.ny + 1, self.nz + 1, 5, 5, 5)

    def ffd(self, event_name: str) -> str:
        """Forward field directory, ``ffd//'/'//event//'/'`` (ses3d_input.f90:309-310),
        resolved relative to MAIN/."""
        return os.path.normpath(os.path.join(self.root, "MAIN", self.setup.ffd.strip(), event_name))


def load_event(root: str, name: str, adjoint_flag: int, ffd: Optional[str] = None) -> EventInput:
    inp = os.path.join(root, "INPUT")
    ev = F.Event.read(os.path.join(inp, f"event_{name}"))
    nt = ev.nt
    xxs = F.deg2rad32(F.f32(repr(ev.xxs)))
    yys = F.deg2rad32(F.f32(repr(ev.yys)))
    e = EventInput(name, ev, nt, F.f32(repr(ev.dt)), ev.source_type, xxs, yys,
                   F.f32(repr(ev.zzs)), np.array([F.f32(repr(float(m))) for m in ev.moment], np.float32),
                   np.zeros(nt, np.float32))
    if adjoint_flag in (0, 1):
        e.stf = F.read_stf(os.path.join(inp, "stf"), nt)
        e.receivers = F.Receivers.read(os.path.join(inp, f"recfile_{name}"))
        e.rec32 = e.receivers.loc32()
    else:
        d = os.path.join(root, "ADJOINT", name)
        e.ad_loc32 = F.read_ad_srcfile(os.path.join(d, "ad_srcfile"))
        e.ad_stf = np.stack([F.read_ad_src(os.path.join(d, f"ad_src_{k + 1}"), nt)
                             for k in range(len(e.ad_loc32))]) if len(e.ad_loc32) else np.zeros((0, nt, 3), np.float32)
        if ffd is not None and os.path.exists(os.path.join(ffd, "saving_vector_0")):
            e.saving_vector = F.read_saving_vector(os.path.join(ffd, "saving_vector_0"), nt)
    return e


def load_run(root: str, events: Optional[List[str]] = None) -> RunInput:
    inp = os.path.join(root, "INPUT")
    setup = F.Setup.read(os.path.join(inp, "setup"))
    boxes = F.read_boxfile(os.path.join(root, "MODELS", "MODELS", "boxfile"))
    b, slices = section_box(boxes)
    tau, D = F.read_relax(os.path.join(inp, "relax"))
    shape = (b.nx + 1, b.ny + 1, b.nz + 1, 5, 5, 5)
    mdir = os.path.join(root, "MODELS", "MODELS")
    if len(boxes) == 1:
        model = F.read_model(mdir, shape, 0, bool(setup.is_diss))
    else:
        # each rank's records in its part of the section (SOURCE/ses3d_input.f90:472-501)
        model = {}
        for k, (bk, sl) in enumerate(zip(boxes, slices)):
            part = F.read_model(mdir, (bk.nx + 1, bk.ny + 1, bk.nz + 1, 5, 5, 5), bk.index - 1,
                                bool(setup.is_diss))
            for name, a in part.items():
                if name not in model:
                    model[name] = np.zeros(shape, np.float32, order="F")
                model[name][sl] = a
    run = RunInput(root, setup, b, F.f32(repr(b.xmin)), F.f32(repr(b.xmax)), F.f32(repr(b.ymin)),
                   F.f32(repr(b.ymax)), F.f32(repr(b.zmin)), F.f32(repr(b.zmax)), setup.zmax_global,
                   int(setup.is_diss), int(setup.adjoint_flag), int(setup.samp_ad), tau, D, model,
                   boxes=boxes, box_slices=slices)
    names = events if events is not None else F.read_event_list(os.path.join(inp, "event_list"))
    for n in names:
        run.events.append(load_event(root, n, run.adjoint_flag, run.ffd(n)))
    return run


def section_box(boxes: List[F.Box]):
    """The section a boxfile's rank boxes tile, as one box, and each rank
    box's element slices in it.  Box (i, j, k) holds elements nx_min+i-1 ..
    nx_max+i-1 along theta, likewise along phi, counted from the section's
    minimum, and radially nz_min+k-1 .. nz_max+k-1 counted from zmin while the
    arrays' element index runs down from zmax
    (MODELS/SOURCE/generate_models.f90:200-282, SOURCE/ses3d_init.f90:239)."""
    from .models import rank_slices
    if len(boxes) == 1:
        b = boxes[0]
        return b, [tuple(slice(0, n + 1) for n in (b.nx, b.ny, b.nz))]
    n = [max(getattr(b, f"n{a}_max") + b.multi[q] for b in boxes) for q, a in enumerate("xyz")]
    if sum((b.nx + 1) * (b.ny + 1) * (b.nz + 1) for b in boxes) != n[0] * n[1] * n[2]:
        raise ValueError(f"the boxes do not tile a section of {n[0]}x{n[1]}x{n[2]} elements")
    px, py, pz = (max(b.multi[q] for b in boxes) for q in range(3))
    lo = {q: next(b for b in boxes if b.multi[q] == 1) for q in range(3)}
    hi = {0: next(b for b in boxes if b.multi[0] == px), 1: next(b for b in boxes if b.multi[1] == py),
          2: next(b for b in boxes if b.multi[2] == pz)}
    sec = F.Box(1, (1, 1, 1), 0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1, lo[0].xmin, hi[0].xmax,
                lo[1].ymin, hi[1].ymax, lo[2].zmin, hi[2].zmax)
    return sec, [rank_slices(b, n[2]) for b in boxes]
