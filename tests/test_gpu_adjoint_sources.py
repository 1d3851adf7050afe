"""The reference's adjoint-source measurements (ses3d_amd.adjoint_source)
driving the GPU adjoint: the gradient of a measurement equals its central
finite difference.

For a fixed displacement-form source f of a measurement on velocity
seismograms (TOOLS/adjoint_source.py with integrate=False, before its final
differentiation), J(m) = sum_r sum_n f[n] v_r(m)[n] dt is linear in the
velocity seismograms v, and its adjoint source is the tool's output, the
differentiated -(f[n+1]-f[n])/dt.  So the kernels from that adjoint source,
mapped to d/d(vp, vsh, vsv), must match (J(m + h dm) - J(m - h dm)) / 2h up
to the time-discretisation error that tests/test_gpu_autograd.py bounds for
the waveform misfit (0.2% for the velocities at dt = 0.25 s).  The windows
end inside the record, so f ends at zero and the summation by parts behind
the differentiation has no boundary term.  The multitaper sources keep one
undifferentiated sample at the end of the window (adjoint_source.py:194-197);
the tapers are nearly zero there, which the bound for them allows for.
"""
import numpy as np
import pytest
import torch

from ses3d_amd import adjoint_source as A, configs, models as M

pytestmark = pytest.mark.gpu

BOUND = {"cc_L2amp": 0.008, "multitaper": 0.004}
MEASUREMENTS = {
    "cc_L2amp": lambda a, r: (a.make_adsrc_cc("z", 15.0, 65.0, 8.0), a.make_adsrc_L2amp("x", 10.0, 70.0, 8.0)),
    "multitaper": lambda a, r: (a.make_adsrc_mttime("z", 15.0, 65.0, 0.05), a.make_adsrc_mtamp("y", 15.0, 65.0, 0.05)),
}


def _displacement_form(u, dt, make):
    """f per receiver: the same measurement with the differentiation left out."""
    f = np.zeros((3,) + u.shape[1:])
    for r in range(u.shape[1]):
        a = A.AdjointSource()
        s = A.Seismogram.from_traces(u[:, r], dt)
        s.integrate = True             # same traces, no final differentiation
        a.fetch_seismogram(s)
        make(a, r)
        f[:, r] = a.forward_time()
    return f


@pytest.mark.parametrize("kind", sorted(MEASUREMENTS))
def test_measurement_gradient_matches_finite_difference(kind):
    from ses3d_amd.autograd import Problem, misfit
    from ses3d_amd.solver import Solver
    dt = 0.25
    w = configs.Workload("fd", (10, 10, 5), False, 2, dt, 320, 16, "none", period=25.0)
    run = configs.build(w, seed=0)
    print(f"measurement {kind}: 10x10x5 elastic, nt=320, dt={dt}", flush=True)
    s = Solver.from_run(run, adjoint_flag=0, pml=2)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    s.run()
    u0 = s.seismograms().astype(np.float64)
    s.close()
    make = MEASUREMENTS[kind]
    rows = A.event_rows(u0, dt, make)
    f = _displacement_form(u0, dt, make)
    assert rows.any() and np.isfinite(rows).all()

    def measure(ev, u, dt_):
        return float(np.sum(f * u) * dt_), rows

    prob = Problem(run, [], pml=2, ckpt=40, measure=measure)
    m = run.model
    rho0 = 1.0 / m["rhoinv"].astype(np.float64)
    vs0 = np.sqrt(m["mu"] / rho0)
    vp0 = np.sqrt((m["lambda"] + 2 * m["mu"]) / rho0)
    base0 = [rho0, vp0, vs0, vs0.copy()]
    base = [torch.tensor(a, requires_grad=True) for a in base0]
    J = misfit(prob, *base)
    # the moduli are re-formed from (rho, vp, vs) in float32: the same model to rounding
    assert np.isclose(J.item(), np.sum(f * u0) * dt, rtol=1e-5)
    J.backward()
    grads = [b.grad.numpy() for b in base]
    dirs = [M.gaussian_random_field(run.box, 90 + k, corr_len=300e3, rms=1.0) for k in range(4)]
    h = 1e-3
    dm = [None, dirs[1] * vp0, dirs[2] * vs0, dirs[3] * vs0]
    pred = sum(float(np.sum(grads[k] * dm[k])) for k in (1, 2, 3))
    vals = []
    for sgn in (1, -1):
        t = [torch.tensor(base0[0])] + [torch.tensor(base0[k] + sgn * h * dm[k]) for k in (1, 2, 3)]
        vals.append(float(misfit(prob, *t)))
    fd = (vals[0] - vals[1]) / (2 * h)
    err = abs(pred - fd) / abs(fd)
    print(f"{kind}: fd={fd:.6e} adjoint={pred:.6e} rel.err={err:.2e}", flush=True)
    assert err < BOUND[kind]
