"""Differentiable SES3D misfit: ``torch.autograd.Function`` over the GPU solver.

forward:  model (rho, vp, vsh, vsv[, Q]) at the GLL points -> SES3D moduli
          (mu = rho vsh^2, B = rho vsv^2 - mu, lambda = rho vp^2 - 2 mu, as
          MODELS/SOURCE/add_perturbation.f90:512-514) -> forward solve of every
          event -> velocity seismograms -> chi = sum_events 1/2 sum (u - d)^2 dt.
          When a gradient is wanted, the last event batch is solved with
          restart states every ``ckpt`` steps kept in HBM, the others plainly;
          the residuals' adjoint sources are kept.  No adjoint runs here.
backward: the checkpointed adjoint solve with kernel accumulation
          (SOURCE/ses3d_grad.f90, ses3d_grad_Q.f90) of every batch -- the last
          one from the checkpoints the forward kept, the others after a
          checkpointed forward solve of their own (one batch's checkpoints are
          held at a time) -- and the map from SES3D kernels to the gradient of
          chi with respect to the per-point model values:

              dchi/dm_p = -K_p * w_l w_m w_n * Jac^2 * r_p^2 sin(theta_p)

          (GLL quadrature weight of the point in spherical coordinates,
          w_l w_m w_n Jac r^2 sin(theta), times the Jac the reference divides its
          kernels by, SOURCE/ses3d_grad.f90:77-83; the sign follows from the
          adjoint being run forward in reversed time).  K is K_rho, K_cp, K_csh,
          K_csv; for Q the relative-perturbation kernel of ses3d_grad_Q.f90
          (normalised with 2 mu / Q) is divided by Q.

Model tensors on the GPU stay there: the moduli are formed on the device and
read by the solver from device memory (ses3d_hip_set_model_device), and the
kernels come back as device tensors (ses3d_hip_get_kernels_device); results
are bit-identical to the host path.

The kernels are continuous-adjoint, time-sampled quantities, so the gradient
equals the derivative of the discrete misfit up to the time-discretisation
error, O(omega dt); tests/test_gpu_autograd.py checks it against central finite
differences and checks that the error falls when dt is halved.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import misfit as MF
from . import models as M
from .solver import Solver

GLL_W = np.array([0.1000000000000, 0.5444444444444, 0.7111111111111, 0.5444444444444, 0.1000000000000])


def quadrature_weights(run) -> np.ndarray:
    """w_l w_m w_n Jac^2 r^2 sin(theta) per GLL point, reference layout, float64."""
    box = run.box
    x, y, z = M.gll_coords(box)
    dx = (box.xmax - box.xmin) / (box.nx + 1)
    dy = (box.ymax - box.ymin) / (box.ny + 1)
    dz = (box.zmax - box.zmin) / (box.nz + 1)
    jac = dx * dy * dz / 8.0
    w3 = GLL_W[:, None, None] * GLL_W[None, :, None] * GLL_W[None, None, :]
    X = x[:, None, None, :, None, None]
    Z = z[None, None, :, None, None, :]
    W = w3[None, None, None] * (jac * jac) * Z ** 2 * np.sin(X)
    # Fortran order, as the kernels come back, so products with them stream
    return np.asfortranarray(np.broadcast_to(W, (box.nx + 1, box.ny + 1, box.nz + 1, 5, 5, 5)))


class Problem:
    """Fixed geometry, sources, receivers and observed data for a set of events.

    ``run``: a RunInput (geometry, relaxation, events); ``data``: one observed
    seismogram array (3, nr, nt) per event, receivers in the order the solver
    locates them.  Events are solved ``events_per_batch`` at a time in one
    context (they advance in the same launches).

    ``measure``: optional callable ``(ev, u, dt) -> (chi_ev, rows)`` that
    replaces the L2 waveform misfit: ``u`` is the event's synthetic velocity
    seismograms (3, nr, nt), ``rows`` the adjoint-source rows (nr, nt, 3),
    time-reversed as ADJOINT/<ev>/ad_src_<k> stores them (for instance
    ses3d_amd.adjoint_source.event_rows of the reference's measurements)."""

    def __init__(self, run, data: Sequence[np.ndarray], pml: int = 2, ckpt: int = 100,
                 events_per_batch: Optional[int] = None, q_model: Optional[np.ndarray] = None,
                 taper_width: float = 0.0, progress=None, history_degree: int = 4,
                 kernel_bits: int = 64, ckpt_keep: int = 1, measure=None):
        self.run = run
        self.measure = measure
        self.data = [np.asarray(d, np.float32) for d in data]
        self.pml = pml
        self.ckpt = ckpt
        # trailing checkpoint segments whose sampled steps stay in HBM (no replay)
        self.ckpt_keep = ckpt_keep
        self.n_events = len(run.events)
        self.events_per_batch = events_per_batch or self.n_events
        self.W = quadrature_weights(run)
        self.dt = float(run.events[0].dt)
        self.is_diss = bool(run.is_diss)
        self.q_model = q_model if q_model is not None else run.model.get("Q")
        self.last_seismograms: List[np.ndarray] = []
        nt = run.events[0].nt
        self.taper = MF.cosine_taper(nt, self.dt, taper_width) if taper_width > 0 else None
        self.progress = progress          # optional callable(str), told when a batch is done
        # polynomial degree of the stored forward fields (4: as the reference
        # stores them; lower: (d+1)^3/125 of the segment memory, lossy)
        self.history_degree = history_degree
        # kernel accumulators: 64 (float64) or 32 (float32, as the reference)
        self.kernel_bits = kernel_bits

    def fields_torch(self, rho, vp, vsh, vsv, Q=None) -> dict:
        """fields() on the device: the same float64 operations in the same
        order, rounded to float32; tensors indexed as the reference arrays."""
        import torch
        rho, vp, vsh, vsv = (t.detach().to(torch.float64) for t in (rho, vp, vsh, vsv))
        mu = rho * vsh * vsh
        out = {"rhoinv": 1.0 / rho, "mu": mu, "lambda": rho * vp * vp - 2.0 * mu,
               "A": torch.zeros_like(rho), "B": rho * vsv * vsv - mu, "C": torch.zeros_like(rho)}
        if self.is_diss:
            q = Q.detach() if Q is not None else torch.as_tensor(np.asarray(self.q_model), device=rho.device)
            out["Q"] = q.to(torch.float64)
        return {k: v.to(torch.float32).contiguous() for k, v in out.items()}

    def fields(self, rho, vp, vsh, vsv, Q=None) -> dict:
        rho, vp, vsh, vsv = (np.asarray(a, np.float64) for a in (rho, vp, vsh, vsv))
        mu = rho * vsh * vsh
        out = {"rhoinv": 1.0 / rho, "mu": mu, "lambda": rho * vp * vp - 2.0 * mu,
               "A": np.zeros_like(rho), "B": rho * vsv * vsv - mu, "C": np.zeros_like(rho)}
        if self.is_diss:
            out["Q"] = np.asarray(Q if Q is not None else self.q_model, np.float64)
        return {k: np.asfortranarray(v.astype(np.float32)) for k, v in out.items()}

    def _measure(self, ev: int, u: np.ndarray, rows: bool):
        """(chi of event ev, its adjoint rows or None)."""
        if self.measure is not None:
            chi, r = self.measure(ev, u, self.dt)
            return float(chi), (np.ascontiguousarray(r, np.float32) if rows else None)
        chi = MF.l2_misfit(u, self.data[ev], self.dt, self.taper)
        return chi, (MF.l2_adjoint_rows(u, self.data[ev], self.dt, taper=self.taper) if rows else None)

    def _batches(self):
        for b0 in range(0, self.n_events, self.events_per_batch):
            yield list(range(b0, min(b0 + self.events_per_batch, self.n_events)))

    def _solver(self, model, af, n):
        s = Solver.from_run(self.run, adjoint_flag=af, pml=self.pml, n_events=n)
        if af:
            s.set_history_degree(self.history_degree)
        if af == 2:
            s.set_kernel_precision(self.kernel_bits)
        if _is_tensor(model["rhoinv"]):
            s.set_model_tensors(model)
        else:
            s.set_model(model)
        return s

    # ------------------------------------------------ forward, then backward
    def forward_for_gradient(self, model):
        """chi, with what backward needs kept: the checkpointed forward
        contexts of the last event batch and every event's adjoint sources.
        Runs no adjoint."""
        chi = 0.0
        self.last_seismograms = []
        st = _GradState(self, model)
        batches = list(self._batches())
        plain = {}
        try:
            for bi, evs in enumerate(batches):
                n = len(evs)
                if bi == len(batches) - 1:
                    f, _ = st.pair(n)
                    s = f
                    for k, ev in enumerate(evs):
                        f.load_event(k, self.run.events[ev])
                    f.init()
                    f.forward_checkpointed(self.ckpt, keep=self.ckpt_keep)
                    st.held = bi
                else:
                    if n not in plain:
                        plain[n] = self._solver(model, 0, n)
                    s = plain[n]
                    for k, ev in enumerate(evs):
                        s.load_event(k, self.run.events[ev])
                    s.init()
                    s.run()
                for k, ev in enumerate(evs):
                    u = s.seismograms(k)
                    self.last_seismograms.append(u)
                    st.seismograms[ev] = u
                    c, rows = self._measure(ev, u, True)
                    chi += c
                    st.sources[ev] = (self.run.events[ev].rec32[s.receiver_index(k)], rows)
                if self.progress:
                    self.progress(f"forward solve of events {evs[0]}..{evs[-1]} done")
        except BaseException:
            st.close()
            raise
        finally:
            for p in plain.values():
                p.close()
        return chi, st

    def gradient_kernels(self, st: "_GradState"):
        """The kernels summed over events, in event order, from the state
        forward_for_gradient kept (numpy float64 arrays, or float64 device
        tensors when the model is on the device)."""
        batches = list(self._batches())
        dev = _is_tensor(st.model["rhoinv"])

        def adjoint(bi, recompute):
            evs = batches[bi]
            f, a = st.pair(len(evs))
            if recompute:
                for k, ev in enumerate(evs):
                    f.load_event(k, self.run.events[ev])
                f.init()
                f.forward_checkpointed(self.ckpt, keep=self.ckpt_keep)
                for k, ev in enumerate(evs):
                    if not np.array_equal(f.seismograms(k), st.seismograms[ev]):
                        raise RuntimeError(f"event {ev}: the checkpointed forward solve differs from the first")
            for k, ev in enumerate(evs):
                a.set_adjoint_sources(k, *st.sources[ev])
            a.init()
            a.adjoint_checkpointed(f)
            out = [a.kernels_tensors(k) if dev else a.kernels(k) for k in range(len(evs))]
            if self.progress:
                self.progress(f"gradient of events {evs[0]}..{evs[-1]} done")
            return out

        # the held batch first (its checkpoints are overwritten by any other
        # batch of its size), added in its place, so the sum runs in event
        # order; once released (a second backward), every batch is recomputed
        held = adjoint(st.held, recompute=False) if st.held is not None else None
        K = None
        for bi in range(len(batches)):
            for kk in (held if bi == st.held else adjoint(bi, recompute=True)):
                if K is None:
                    K = dict(kk)
                else:
                    for name in K:
                        K[name] += kk[name]
        return K

    def misfit(self, model) -> float:
        """Plain forward solve of every event -> chi (no state kept)."""
        chi = 0.0
        self.last_seismograms = []
        ctx = {}    # one context per batch size, reused by every batch of that size
        try:
            for evs in self._batches():
                n = len(evs)
                if n not in ctx:
                    ctx[n] = self._solver(model, 0, n)
                s = ctx[n]
                for k, ev in enumerate(evs):
                    s.load_event(k, self.run.events[ev])
                s.init()
                s.run()
                for k, ev in enumerate(evs):
                    u = s.seismograms(k)
                    self.last_seismograms.append(u)
                    chi += self._measure(ev, u, False)[0]
                if self.progress:
                    self.progress(f"forward solve of events {evs[0]}..{evs[-1]} done")
        finally:
            for s in ctx.values():
                s.close()
        return chi

    def misfit_and_gradient(self, model):
        """chi and the SES3D kernels summed over events (float64 dict).  The
        forward and adjoint contexts of a batch size are created once and
        reused by every batch of that size, so their checkpoint and segment
        buffers are allocated once."""
        chi = 0.0
        K = None
        self.last_seismograms = []
        ctx = {}
        try:
            for evs in self._batches():
                n = len(evs)
                if n not in ctx:
                    ctx[n] = (self._solver(model, 1, n), self._solver(model, 2, n))
                f, a = ctx[n]
                for k, ev in enumerate(evs):
                    f.load_event(k, self.run.events[ev])
                f.init()
                f.forward_checkpointed(self.ckpt, keep=self.ckpt_keep)
                for k, ev in enumerate(evs):
                    u = f.seismograms(k)
                    self.last_seismograms.append(u)
                    c, rows = self._measure(ev, u, True)
                    chi += c
                    a.set_adjoint_sources(k, self.run.events[ev].rec32[f.receiver_index(k)], rows)
                a.init()
                a.adjoint_checkpointed(f)
                for k in range(n):
                    kk = a.kernels(k)      # fresh arrays, Fortran order
                    if K is None:
                        K = dict(kk)
                    else:
                        for name in K:
                            K[name] += kk[name]
                if self.progress:
                    self.progress(f"gradient of events {evs[0]}..{evs[-1]} done")
        finally:
            for f, a in ctx.values():
                f.close()
                a.close()
        return chi, K

    def gradient_from_kernels(self, K: dict, model: dict) -> dict:
        """dchi/d(rho, vp, vsh, vsv[, Q]) per GLL point from summed kernels
        (numpy arrays, or device tensors with the weights moved to the device)."""
        if _is_tensor(K["grad_rho"]):
            import torch
            if getattr(self, "_W_dev", None) is None:
                self._W_dev = torch.from_numpy(np.ascontiguousarray(self.W)).to(K["grad_rho"].device)
            s = -self._W_dev
            q = model["Q"].to(torch.float64) if "Q" in model else None
        else:
            s = -self.W
            q = np.asarray(model["Q"], np.float64) if "Q" in model else None
        g = {"rho": s * K["grad_rho"], "vp": s * K["grad_cp"], "vsh": s * K["grad_csh"], "vsv": s * K["grad_csv"]}
        if self.is_diss and "grad_Q_mu" in K:
            g["Q"] = s * K["grad_Q_mu"] / q
        return g


def _is_tensor(x) -> bool:
    return type(x).__module__.startswith("torch")


class _GradState:
    """What backward needs from forward: the forward/adjoint context pair per
    batch size (the pair of the held batch carries its checkpoints), every
    event's adjoint sources and seismograms.  Contexts are closed with it."""

    def __init__(self, problem: Problem, model: dict):
        self.problem = problem
        self.model = model
        self.ctx = {}
        self.sources = {}
        self.seismograms = {}
        self.held = None

    def pair(self, n: int):
        if n not in self.ctx:
            self.ctx[n] = (self.problem._solver(self.model, 1, n), self.problem._solver(self.model, 2, n))
        return self.ctx[n]

    def close(self):
        """Free the contexts (and the held batch's checkpoints with them)."""
        for f, a in self.ctx.values():
            f.close()
            a.close()
        self.ctx = {}
        self.held = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


try:
    import torch

    class Ses3dMisfit(torch.autograd.Function):
        """chi(rho, vp, vsh, vsv, Q); backward is the adjoint time loop.

        Forward solves every event (checkpointed for the last batch when a
        gradient is wanted) and keeps the adjoint sources; backward runs the
        adjoint solves and maps the kernels to the gradient.  A forward
        without backward runs no adjoint.  Tensors on the GPU stay on the GPU."""

        @staticmethod
        def forward(ctx, rho, vp, vsh, vsv, Q, problem: Problem):
            if rho.is_cuda:
                model = problem.fields_torch(rho, vp, vsh, vsv, Q)
            else:
                arrs = [t.detach().to("cpu", torch.float64).numpy() for t in (rho, vp, vsh, vsv)]
                q = Q.detach().to("cpu", torch.float64).numpy() if Q is not None else None
                model = problem.fields(*arrs, Q=q)
            ctx.problem = problem
            ctx.state = None
            if any(ctx.needs_input_grad[:5]):
                chi, ctx.state = problem.forward_for_gradient(model)
            else:
                chi = problem.misfit(model)
            ctx.meta = [(t.dtype, t.device) for t in (rho, vp, vsh, vsv)] + \
                       [((Q.dtype, Q.device) if Q is not None else None)]
            return torch.tensor(chi, dtype=torch.float64, device=rho.device)

        @staticmethod
        def backward(ctx, grad_out):
            problem, st = ctx.problem, ctx.state
            try:
                K = problem.gradient_kernels(st)
            finally:
                # the autograd node outlives backward as long as the misfit
                # tensor does: its contexts and checkpoint storage are freed
                # here (a second backward recomputes every batch)
                st.close()
            grads = problem.gradient_from_kernels(K, st.model)
            out = []
            for name, meta in zip(("rho", "vp", "vsh", "vsv", "Q"), ctx.meta):
                if meta is None or name not in grads or not ctx.needs_input_grad[len(out)]:
                    out.append(None)
                    continue
                dtype, dev = meta
                g = grads[name]
                g = g if _is_tensor(g) else torch.from_numpy(g)
                out.append((grad_out.to(g.device, torch.float64) * g).to(device=dev, dtype=dtype))
            return (*out, None)

    def misfit(problem: Problem, rho, vp, vsh, vsv=None, Q=None):
        """Differentiable misfit; vsv defaults to vsh (isotropic)."""
        if vsv is None:
            vsv = vsh
        return Ses3dMisfit.apply(rho, vp, vsh, vsv, Q, problem)

except ImportError:   # torch is optional for the numpy API above
    pass
