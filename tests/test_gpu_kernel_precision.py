"""Kernel accumulators in float32 (ses3d_hip_set_kernel_precision(ctx, 32)).

That is the type of the reference's grad_* arrays
(SOURCE/ses3d_modules.f90:221-222), summed in its statement order
(SOURCE/ses3d_grad.f90:77-83, SOURCE/ses3d_grad_Q.f90:111-128, normalised as
:151-153), so the kernels equal the float32 oracle's bit for bit, grad_Q_mu
included, on every path that accumulates them: stored history, HBM
checkpoints, lower-degree stored fields, a group of sections.  The reference
binary's own kernels are matched bit for bit in test_gpu_golden.py.  The
float64 default differs from them by accumulation rounding only (relative L2
<= 1e-5).
"""
import numpy as np
import pytest

from tests.cases import make_case, oracle_for
from tests.test_gpu_adjoint import _adjoint_sources, _rel

pytestmark = pytest.mark.gpu


def _ctx(run, af, bits=64, hdeg=4):
    from ses3d_amd.solver import Solver
    s = Solver.from_run(run, adjoint_flag=af)
    if hdeg != 4:
        s.set_history_degree(hdeg)
    if af == 2:
        s.set_kernel_precision(bits)
    s.set_model(run.model)
    return s


def _oracle_kernels(run, e, loc, rows, hdeg=4):
    from oracle import pyoracle as PO
    of = oracle_for(run, e, adjoint_flag=1)
    if hdeg != 4:
        of.set_history_degree(hdeg)
    of.run(1, e.nt)
    oa = PO.Oracle(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin,
                   ymax=run.ymax, zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global, nt=e.nt,
                   dt=e.dt, is_diss=run.is_diss, adjoint_flag=2, samp_ad=run.samp_ad, pml=2,
                   tau_p=run.tau_p, D_p=run.D_p)
    oa.set_model(run.model)
    oa.set_adjoint_sources(loc, rows)
    oa.init()
    oa.run(1, e.nt, fwd=of)
    return oa


def _stored_history_kernels(run, e, loc, rows, bits, hdeg=4):
    f = _ctx(run, 1, hdeg=hdeg)
    f.load_event(0, e)
    f.init()
    f.run()
    a = _ctx(run, 2, bits, hdeg)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    for st in np.flatnonzero(f.saving_vector()) + 1:
        a.put_history(int(st), f.history(int(st)))
    a.run()
    K = a.kernels()
    f.close()
    a.close()
    return K


def _checkpointed_kernels(run, e, loc, rows, bits, k, hdeg=4):
    f = _ctx(run, 1, hdeg=hdeg)
    f.load_event(0, e)
    f.init()
    f.forward_checkpointed(k)
    a = _ctx(run, 2, bits, hdeg)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    a.adjoint_checkpointed(f)
    K = a.kernels()
    f.close()
    a.close()
    return K


@pytest.mark.parametrize("is_diss,aniso,hdeg", [(1, 0, 4), (0, 0, 4), (1, 1, 4), (1, 0, 2)])
def test_float32_kernels_equal_oracle32(tmp_path, is_diss, aniso, hdeg):
    run = make_case(str(tmp_path), nel=(5, 6, 4), nt=40, is_diss=bool(is_diss), adjoint_flag=1, samp_ad=3,
                    seed=80 + is_diss + 2 * aniso + hdeg)
    if aniso:
        rng = np.random.default_rng(5)
        for c in ("A", "B", "C"):
            run.model[c] = np.asfortranarray((run.model["mu"] * rng.uniform(-0.05, 0.05, run.model["mu"].shape))
                                             .astype(np.float32))
    e = run.events[0]
    print(f"float32 kernels: diss={is_diss} aniso={aniso} degree={hdeg}", flush=True)
    f0 = _ctx(run, 1)
    f0.load_event(0, e)
    f0.init()
    f0.run()
    loc, rows = _adjoint_sources(e, f0.seismograms())
    f0.close()
    K32 = _stored_history_kernels(run, e, loc, rows, 32, hdeg)
    oa = _oracle_kernels(run, e, loc, rows, hdeg)
    assert ("grad_Q_mu" in K32) == bool(is_diss)
    for g, v in K32.items():
        ref = oa.get(g)
        assert np.linalg.norm(ref) > 0, g
        np.testing.assert_array_equal(v, ref.astype(np.float64), err_msg=g)
    # the checkpointed path accumulates the same float32 sums
    Kc = _checkpointed_kernels(run, e, loc, rows, 32, 7, hdeg)
    for g in K32:
        np.testing.assert_array_equal(Kc[g], K32[g], err_msg=f"checkpointed {g}")
    # float64 accumulation: the same increments, rounded less
    K64 = _stored_history_kernels(run, e, loc, rows, 64, hdeg)
    for g in K32:
        rel = _rel(K64[g], K32[g])
        print(f"  {g}: float64 vs float32 accumulation rel L2 {rel:.2e}", flush=True)
        assert rel <= 1e-5, (g, rel)


def test_float32_kernels_in_a_group(tmp_path):
    """A group's adjoint accumulates each member's kernels in that member's
    type: float32 members give the float32 single-context kernels exactly."""
    from ses3d_amd.solver import Group
    runs = [make_case(str(tmp_path / "a"), nel=(6, 5, 4), nt=30, is_diss=True, adjoint_flag=1, samp_ad=2, seed=91),
            make_case(str(tmp_path / "b"), nel=(4, 4, 5), nt=30, is_diss=True, adjoint_flag=1, samp_ad=2, seed=92)]
    singles, fwd, srcs = [], [], []
    for r in runs:
        e = r.events[0]
        f = _ctx(r, 1)
        f.load_event(0, e)
        f.init()
        f.run()
        loc, rows = _adjoint_sources(e, f.seismograms())
        fwd.append(f)
        srcs.append((loc, rows))
        singles.append(_stored_history_kernels(r, e, loc, rows, 32))
    ads = []
    for r, f, (loc, rows) in zip(runs, fwd, srcs):
        a = _ctx(r, 2, 32)
        a.set_adjoint_sources(0, loc, rows)
        a.init()
        for st in np.flatnonzero(f.saving_vector()) + 1:
            a.put_history(int(st), f.history(int(st)))
        ads.append(a)
    print("group adjoint with float32 accumulators", flush=True)
    with Group(ads) as g:
        g.run()
    for a, K in zip(ads, singles):
        Kg = a.kernels()
        for name in K:
            np.testing.assert_array_equal(Kg[name], K[name], err_msg=name)
    for s in ads + fwd:
        s.close()


def test_kernel_precision_rejects_other_types(tmp_path):
    from ses3d_amd.solver import Ses3dError
    run = make_case(str(tmp_path), nel=(3, 3, 3), nt=6, is_diss=True, adjoint_flag=1, samp_ad=2, seed=4)
    a = _ctx(run, 2)
    for bits in (16, 0, 128):
        with pytest.raises(Ses3dError, match="precision"):
            a.set_kernel_precision(bits)
    a.close()


def test_autograd_problem_float32_kernels():
    """Problem(kernel_bits=32): the misfit is unchanged and the gradient moves
    by float32 accumulation rounding only; Problem(ckpt_keep=3) gives the same
    gradient bit for bit."""
    import torch
    from dataclasses import replace
    from ses3d_amd import configs
    from ses3d_amd.autograd import Problem, misfit
    from ses3d_amd.solver import Solver
    w = configs.Workload("kb", (10, 10, 5), True, 2, 0.25, 240, 16, "blobs", period=25.0)
    true = configs.build(w, seed=2)
    cur = configs.build(replace(w, perturbation="none"), seed=2)
    s = Solver.from_run(true, adjoint_flag=0, pml=2)
    s.set_model(true.model)
    s.load_event(0, true.events[0])
    s.init()
    s.run()
    data = [s.seismograms()]
    s.close()
    m = cur.model
    rho0 = 1.0 / m["rhoinv"].astype(np.float64)
    vs0 = np.sqrt(m["mu"] / rho0)
    vp0 = np.sqrt((m["lambda"] + 2 * m["mu"]) / rho0)
    out = {}
    for bits, keep in ((64, 1), (32, 1), (32, 3)):
        prob = Problem(cur, data, pml=2, ckpt=40, taper_width=20.0, kernel_bits=bits, ckpt_keep=keep)
        x = [torch.tensor(a, requires_grad=True) for a in (rho0, vp0, vs0, vs0.copy())]
        chi = misfit(prob, *x)
        chi.backward()
        out[bits, keep] = (float(chi.detach()), np.concatenate([t.grad.numpy().ravel() for t in x]))
    # kept checkpoint segments change nothing
    assert out[32, 3][0] == out[32, 1][0]
    np.testing.assert_array_equal(out[32, 3][1], out[32, 1][1])
    out = {bits: out[bits, 1] for bits in (64, 32)}
    assert out[32][0] == out[64][0]
    g32, g64 = out[32][1], out[64][1]
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    print(f"gradient, float32 vs float64 kernel accumulators: rel L2 {rel:.2e}", flush=True)
    assert np.linalg.norm(g64) > 0 and rel <= 1e-4
