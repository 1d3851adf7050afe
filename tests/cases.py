"""Small seeded SES3D run trees for the tests (inputs in the reference formats)."""
from __future__ import annotations

import numpy as np

from ses3d_amd import formats as F, inputs as I, models as M


def make_case(root: str, nel=(6, 7, 4), nt=60, dt=0.2, is_diss=False, adjoint_flag=0, samp_ad=3,
              source_type=10, seed=0, src=(49.0, 21.0, 30000.0), n_rec=(2, 2), perturb=True,
              period=8.0, n_events=1):
    """Write a run tree under ``root`` and return the loaded RunInput."""
    rng = np.random.default_rng(seed)
    setup = M.section_setup(*nel, is_diss=is_diss, adjoint_flag=adjoint_flag, samp_ad=samp_ad)
    box = F.make_boxes(setup)[0]
    pert = M.gaussian_random_field(box, seed=seed + 100, corr_len=200e3, rms=0.03) if perturb else None
    model = M.make_model(box, is_diss=is_diss, perturbation=pert)
    events, recs = [], []
    for k in range(n_events):
        if k == 0:
            s = src
        else:
            s = (rng.uniform(42, 58), rng.uniform(12, 28), rng.uniform(5e3, 200e3))
        events.append(F.Event(nt=nt, dt=dt, xxs=round(float(s[0]), 3), yys=round(float(s[1]), 3),
                              zzs=round(float(s[2]), 1), source_type=source_type,
                              moment=tuple(M.random_moment_tensor(rng)), ofd=f"../DATA/OUTPUT/{k + 1}/",
                              ssamp=nt, output_displacement=1))
        recs.append(M.surface_receiver_grid(setup, *n_rec) if k == 0 else M.random_receivers(setup, 3 + k, rng))
    stf = M.ricker(nt + 10, dt, period, t0=0.75 * period)
    case = M.RunCase(setup, events, recs, stf, model)
    M.write_run_tree(root, case)
    return I.load_run(root)


def oracle_for(run, ev, adjoint_flag=None, precision=32, nt=None, pml=2):
    from oracle import pyoracle as PO
    af = run.adjoint_flag if adjoint_flag is None else adjoint_flag
    o = PO.Oracle(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin,
                  ymax=run.ymax, zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global,
                  nt=nt or ev.nt, dt=ev.dt, is_diss=run.is_diss, adjoint_flag=af, samp_ad=run.samp_ad,
                  pml=pml, tau_p=run.tau_p, D_p=run.D_p, precision=precision)
    o.set_model(run.model)
    if af in (0, 1):
        o.set_source(ev.source_type, ev.xxs, ev.yys, ev.zzs, ev.moment, ev.stf)
        o.set_receivers(ev.rec32)
    else:
        o.set_adjoint_sources(ev.ad_loc32, ev.ad_stf)
    o.init()
    return o
