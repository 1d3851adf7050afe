"""The torch.autograd.Function's structure: the adjoint runs in backward(),
never in forward(); its gradient is bit-identical to the eager
misfit_and_gradient path (one batch, and several batches whose checkpoints
are recomputed in backward); model tensors on the GPU give bit-identical
results without a host round trip of the model or the kernels."""
import numpy as np
import pytest
import torch
from dataclasses import replace

from ses3d_amd import configs

pytestmark = pytest.mark.gpu


def _problem(events_per_batch, n_events=2):
    from ses3d_amd.autograd import Problem
    from ses3d_amd.solver import Solver
    w = configs.Workload("bw", (8, 8, 4), True, 2, 0.25, 200, 16, "blobs", period=25.0, n_events=n_events)
    true = configs.build(w, seed=3)
    cur = configs.build(replace(w, perturbation="none"), seed=3)
    data = []
    for k in range(n_events):
        s = Solver.from_run(true, adjoint_flag=0, pml=2, n_events=1)
        s.set_model(true.model)
        s.load_event(0, true.events[k])
        s.init()
        s.run()
        data.append(s.seismograms())
        s.close()
    prob = Problem(cur, data, pml=2, ckpt=40, taper_width=20.0, events_per_batch=events_per_batch)
    m = cur.model
    rho0 = 1.0 / m["rhoinv"].astype(np.float64)
    vs0 = np.sqrt(m["mu"] / rho0)
    vp0 = np.sqrt((m["lambda"] + 2 * m["mu"]) / rho0)
    return prob, [rho0, vp0, vs0, vs0.copy(), m["Q"].astype(np.float64)]


class _Count:
    def __init__(self, monkeypatch):
        from ses3d_amd.solver import Solver
        self.n = 0
        orig = Solver.adjoint_checkpointed

        def counted(s, fwd):
            self.n += 1
            return orig(s, fwd)
        monkeypatch.setattr(Solver, "adjoint_checkpointed", counted)


@pytest.mark.parametrize("per_batch", [2, 1])
def test_adjoint_runs_in_backward_only(monkeypatch, per_batch):
    from ses3d_amd.autograd import misfit
    prob, base = _problem(per_batch)
    count = _Count(monkeypatch)
    print(f"events per batch {per_batch}", flush=True)
    x = [torch.tensor(a, requires_grad=True) for a in base]
    chi = misfit(prob, *x)
    assert count.n == 0, "forward ran an adjoint"
    chi.backward()
    assert count.n == 2 // per_batch
    # today's eager path: forward + adjoint per batch, kernels summed in event order
    c_ref, K = prob.misfit_and_gradient(prob.fields(*base[:4], Q=base[4]))
    g_ref = prob.gradient_from_kernels(K, prob.fields(*base[:4], Q=base[4]))
    assert float(chi.detach()) == c_ref
    for name, t in zip(("rho", "vp", "vsh", "vsv", "Q"), x):
        np.testing.assert_array_equal(t.grad.numpy(), g_ref[name], err_msg=name)


def test_backward_frees_its_contexts_and_a_second_backward_recomputes(monkeypatch):
    """The autograd node lives as long as the misfit tensor: backward must free
    the contexts (and the checkpoints they hold) itself; a second backward
    (retain_graph) recomputes every batch and gives the same gradient."""
    from ses3d_amd.autograd import misfit
    from ses3d_amd.solver import Solver
    prob, base = _problem(1)
    made = []
    orig = Solver.from_run.__func__

    def tracked(cls, *a, **k):
        s = orig(cls, *a, **k)
        made.append(s)
        return s
    monkeypatch.setattr(Solver, "from_run", classmethod(tracked))
    x = [torch.tensor(a, requires_grad=True) for a in base]
    chi = misfit(prob, *x)
    chi.backward(retain_graph=True)
    assert made and all(s.h is None for s in made), "contexts left open after backward"
    g1 = [t.grad.clone() for t in x]
    for t in x:
        t.grad = None
    n0 = len(made)
    chi.backward()
    assert len(made) > n0 and all(s.h is None for s in made)
    for name, a, t in zip(("rho", "vp", "vsh", "vsv", "Q"), g1, x):
        np.testing.assert_array_equal(t.grad.numpy(), a.numpy(), err_msg=name)


def test_forward_without_grad_runs_no_adjoint(monkeypatch):
    from ses3d_amd.autograd import misfit
    prob, base = _problem(2)
    count = _Count(monkeypatch)
    chi = misfit(prob, *[torch.tensor(a) for a in base])
    assert count.n == 0 and float(chi) == prob.misfit(prob.fields(*base[:4], Q=base[4]))


def test_device_tensors_match_host_path():
    from ses3d_amd.autograd import misfit
    prob, base = _problem(2)
    xh = [torch.tensor(a, requires_grad=True) for a in base]
    chi_h = misfit(prob, *xh)
    chi_h.backward()
    xd = [torch.tensor(a, device="cuda", requires_grad=True) for a in base]
    chi_d = misfit(prob, *xd)
    assert chi_d.is_cuda and float(chi_d.detach().cpu()) == float(chi_h.detach())
    chi_d.backward()
    for name, th, td in zip(("rho", "vp", "vsh", "vsv", "Q"), xh, xd):
        assert td.grad.is_cuda
        np.testing.assert_array_equal(td.grad.cpu().numpy(), th.grad.numpy(), err_msg=name)
    # the model fields formed on the device equal the host ones bit for bit
    fh = prob.fields(*base[:4], Q=base[4])
    fd = prob.fields_torch(*[torch.tensor(a, device="cuda") for a in base[:4]], Q=torch.tensor(base[4], device="cuda"))
    for k in fh:
        np.testing.assert_array_equal(fd[k].cpu().numpy(), fh[k], err_msg=k)
