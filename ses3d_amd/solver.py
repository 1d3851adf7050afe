"""Python host side of the MI355X SES3D time-stepper (ctypes over libses3d_hip.so).

Mirrors the reference's run flow (SOURCE/ses3d_main.f90:77-182): inputs as
``ses3d_input`` holds them (see :mod:`ses3d_amd.inputs`), ``init`` as
``ses3d_init``, ``run`` as the time loop, and outputs in the reference's file
formats (:mod:`ses3d_amd.formats`).  The library is required: there is no CPU
fallback, and loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SES3D_HIP_LIB", os.path.join(HERE, "libses3d_hip.so"))

_FIELDS_LOCAL = ("vx", "vy", "vz", "dxux", "dyuy", "dzuz", "dxuy", "dyux", "dxuz", "dzux", "dyuz",
                 "dzuy", "sxx", "syy", "szz", "sxy", "syx", "sxz", "szx", "syz", "szy")
_FIELDS_GLOBAL = ("vx_global", "vy_global", "vz_global", "MM_global")
_FIELDS_MEM = ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz")


class Config(C.Structure):
    """``ses3d_hip_config`` (include/ses3d_hip.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("lpd", C.c_int32),
        ("pml", C.c_int32), ("is_diss", C.c_int32), ("adjoint_flag", C.c_int32),
        ("samp_ad", C.c_int32), ("nt", C.c_int32), ("dt", C.c_float),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("zmin", C.c_float), ("zmax", C.c_float), ("zmax_global", C.c_float),
        ("tau_p", C.c_float * 3), ("D_p", C.c_float * 3),
    ]


class Ses3dError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libses3d_hip.so (built by ``python -m ses3d_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Ses3dError(f"{path} is missing: build it with `python -m ses3d_amd.build` "
                         "(there is no CPU fallback)")
    lib = C.CDLL(path)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    i32 = C.c_int32
    sig = {
        "ses3d_hip_create": ([C.POINTER(Config), i32, C.POINTER(P)], C.c_int),
        "ses3d_hip_destroy": ([P], None),
        "ses3d_hip_last_error": ([], C.c_char_p),
        "ses3d_hip_set_model": ([P, fp, fp, fp, fp, fp, fp, fp], C.c_int),
        "ses3d_hip_set_source": ([P, i32, i32, C.c_float, C.c_float, C.c_float, fp, fp], C.c_int),
        "ses3d_hip_set_receivers": ([P, i32, i32, fp], C.c_int),
        "ses3d_hip_set_adjoint_sources": ([P, i32, i32, fp, fp], C.c_int),
        "ses3d_hip_init": ([P], C.c_int),
        "ses3d_hip_run": ([P, i32, i32], C.c_int),
        "ses3d_hip_synchronize": ([P], C.c_int),
        "ses3d_hip_get_nr": ([P, i32, ip], C.c_int),
        "ses3d_hip_get_receiver_index": ([P, i32, ip], C.c_int),
        "ses3d_hip_get_seismograms": ([P, i32, fp], C.c_int),
        "ses3d_hip_get_field": ([P, i32, C.c_char_p, fp], C.c_int),
        "ses3d_hip_get_kernels": ([P, i32, dp, dp, dp, dp, dp], C.c_int),
        "ses3d_hip_get_history": ([P, i32, i32, fp], C.c_int),
        "ses3d_hip_put_history": ([P, i32, i32, fp], C.c_int),
        "ses3d_hip_get_saving_vector": ([P, ip], C.c_int),
        "ses3d_hip_set_history_degree": ([P, i32], C.c_int),
        "ses3d_hip_history_size": ([P, C.POINTER(C.c_int64)], C.c_int),
        "ses3d_hip_set_kernel_precision": ([P, i32], C.c_int),
        "ses3d_hip_forward_checkpointed": ([P, i32], C.c_int),
        "ses3d_hip_adjoint_checkpointed": ([P, P], C.c_int),
        "ses3d_hip_set_checkpoint_keep": ([P, i32], C.c_int),
        "ses3d_hip_prepare_checkpoints": ([P, i32], C.c_int),
        "ses3d_hip_memory_bytes": ([P, C.POINTER(C.c_int64)], C.c_int),
        "ses3d_hip_set_profiling": ([P, i32], C.c_int),
        "ses3d_hip_set_graph_steps": ([P, i32], C.c_int),
        "ses3d_hip_get_info": ([P, C.POINTER(C.c_int64), i32], C.c_int),
        "ses3d_hip_get_timing": ([P, dp, dp, dp, ip], C.c_int),
        "ses3d_hip_set_schedule": ([P, ip], C.c_int),
        "ses3d_hip_get_schedule": ([P, ip, fp], C.c_int),
        "ses3d_hip_autotune": ([P, i32, i32], C.c_int),
        "ses3d_hip_set_streaming": ([P, i32], C.c_int),
        "ses3d_hip_get_streaming": ([P, ip, ip], C.c_int),
        "ses3d_hip_group_create": ([C.POINTER(P), i32, C.POINTER(P)], C.c_int),
        "ses3d_hip_group_destroy": ([P], None),
        "ses3d_hip_group_run": ([P, i32, i32], C.c_int),
        "ses3d_hip_group_autotune": ([P, i32, i32], C.c_int),
        "ses3d_hip_group_synchronize": ([P], C.c_int),
        "ses3d_hip_group_set_schedule": ([P, ip], C.c_int),
        "ses3d_hip_group_get_schedule": ([P, ip, fp], C.c_int),
        "ses3d_hip_group_set_streaming": ([P, i32], C.c_int),
        "ses3d_hip_group_get_streaming": ([P, ip, ip], C.c_int),
        "ses3d_hip_group_set_graph_steps": ([P, i32], C.c_int),
        "ses3d_hip_group_set_profiling": ([P, i32], C.c_int),
        "ses3d_hip_group_get_timing": ([P, dp, dp, dp, ip], C.c_int),
        "ses3d_hip_project": ([P, fp, i32, fp, i32, fp, i32, fp, fp], C.c_int),
        "ses3d_hip_project_noalloc": ([P, fp, i32, fp, i32, fp, i32, fp, fp], C.c_int),
        "ses3d_hip_set_model_device": ([P] + [C.c_void_p] * 7, C.c_int),
        "ses3d_hip_get_kernels_device": ([P, i32] + [C.c_void_p] * 5, C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load_library().ses3d_hip_last_error().decode(errors="replace")
        raise Ses3dError(f"ses3d_hip error {rc}: {msg}")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ref_f32(a) -> np.ndarray:
    """Flatten a reference-layout array in Fortran order as contiguous float32."""
    a = np.asarray(a, np.float32)
    return np.ascontiguousarray(a.ravel(order="F"))


class Solver:
    """One section (mesh + model) and ``n_events`` events on the current GPU.

    Parameters are the float32 values the reference holds after ses3d_input:
    box bounds in rad (theta, phi) and m (radius), ``zmax_global`` in m."""

    # steps per captured step graph for every new context (None: the library's default)
    default_graph_steps: Optional[int] = None

    def __init__(self, *, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, zmax_global, nt, dt,
                 is_diss=0, adjoint_flag=0, samp_ad=1, pml=2, tau_p=(1.0, 1.0, 1.0),
                 D_p=(0.0, 0.0, 0.0), n_events=1):
        self.lib = load_library()
        self.cfg = Config(nx, ny, nz, 4, pml, is_diss, adjoint_flag, samp_ad, nt, dt, xmin, xmax, ymin,
                          ymax, zmin, zmax, zmax_global, (C.c_float * 3)(*[float(v) for v in tau_p]),
                          (C.c_float * 3)(*[float(v) for v in D_p]))
        h = C.c_void_p()
        _check(self.lib.ses3d_hip_create(C.byref(self.cfg), n_events, C.byref(h)))
        self.h = h
        if Solver.default_graph_steps is not None:
            self.set_graph_steps(Solver.default_graph_steps)
        self.n_events = n_events
        self.nt = nt
        self.shape_local = (nx + 1, ny + 1, nz + 1, 5, 5, 5)
        self.shape_global = (4 * (nx + 1) + 1, 4 * (ny + 1) + 1, 4 * (nz + 1) + 1)
        self.n_points = int(np.prod(self.shape_local))

    @classmethod
    def from_run(cls, run, adjoint_flag: Optional[int] = None, pml: int = 2, n_events: Optional[int] = None,
                 nt: Optional[int] = None):
        """Context for a loaded run tree (:func:`ses3d_amd.inputs.load_run`)."""
        ev = run.events[0]
        return cls(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin,
                   ymax=run.ymax, zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global,
                   nt=nt or ev.nt, dt=ev.dt, is_diss=run.is_diss,
                   adjoint_flag=run.adjoint_flag if adjoint_flag is None else adjoint_flag,
                   samp_ad=run.samp_ad, pml=pml, tau_p=run.tau_p, D_p=run.D_p,
                   n_events=n_events or len(run.events))

    # ---------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, "h", None):
            self.lib.ses3d_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- inputs
    def set_model(self, model: dict):
        arrs = [_ref_f32(model[k]) for k in ("rhoinv", "mu", "lambda", "A", "B", "C")]
        q = _ref_f32(model["Q"]) if (self.cfg.is_diss and "Q" in model) else None
        _check(self.lib.ses3d_hip_set_model(self.h, *[_fptr(a) for a in arrs], _fptr(q) if q is not None else None))

    def set_model_tensors(self, model: dict):
        """set_model from torch tensors on this context's GPU, indexed as the
        reference arrays (nx+1, ny+1, nz+1, 5, 5, 5): they are brought into the
        reference's memory layout on the device and read there
        (ses3d_hip_set_model_device), with no host round trip."""
        import torch
        names = ("rhoinv", "mu", "lambda", "A", "B", "C") + (("Q",) if self.cfg.is_diss else ())
        ref = {}
        for k in names:
            t = model[k]
            if t.dtype != torch.float32 or not t.is_cuda or tuple(t.shape) != self.shape_local:
                raise ValueError(f"{k}: a float32 GPU tensor of shape {self.shape_local} is needed")
            ref[k] = t.permute(5, 4, 3, 2, 1, 0).contiguous()   # Fortran order of the reference arrays
        torch.cuda.current_stream().synchronize()                # their contents are complete
        ptr = [ref[k].data_ptr() for k in ("rhoinv", "mu", "lambda", "A", "B", "C")]
        _check(self.lib.ses3d_hip_set_model_device(self.h, *ptr, ref["Q"].data_ptr() if "Q" in ref else None))

    def kernels_tensors(self, ev: int = 0, with_q: Optional[bool] = None) -> dict:
        """kernels() as float64 torch tensors on this context's GPU, indexed as
        the reference arrays, written there by ses3d_hip_get_kernels_device."""
        import torch
        if with_q is None:
            with_q = bool(self.cfg.is_diss)
        names = ["grad_rho", "grad_cp", "grad_csh", "grad_csv"] + (["grad_Q_mu"] if with_q else [])
        rev = tuple(reversed(self.shape_local))
        outs = {n: torch.empty(rev, dtype=torch.float64, device="cuda") for n in names}
        torch.cuda.current_stream().synchronize()
        ptrs = [outs[n].data_ptr() if n in outs else None
                for n in ("grad_rho", "grad_cp", "grad_csh", "grad_csv", "grad_Q_mu")]
        _check(self.lib.ses3d_hip_get_kernels_device(self.h, ev, *ptrs))
        return {n: t.permute(5, 4, 3, 2, 1, 0) for n, t in outs.items()}

    def set_source(self, ev: int, source_type: int, xxs, yys, zzs, moment, so):
        mom = np.ascontiguousarray(moment, np.float32)
        so = np.ascontiguousarray(np.asarray(so, np.float32)[: self.nt])
        if len(so) < self.nt:
            raise ValueError("source time function shorter than nt")
        _check(self.lib.ses3d_hip_set_source(self.h, ev, int(source_type), float(np.float32(xxs)),
                                             float(np.float32(yys)), float(np.float32(zzs)), _fptr(mom), _fptr(so)))

    def set_receivers(self, ev: int, rec32: np.ndarray):
        rec = np.ascontiguousarray(rec32, np.float32).reshape(-1, 3)
        _check(self.lib.ses3d_hip_set_receivers(self.h, ev, len(rec), _fptr(rec)))

    def set_adjoint_sources(self, ev: int, loc32: np.ndarray, stf_rows: np.ndarray):
        loc = np.ascontiguousarray(loc32, np.float32).reshape(-1, 3)
        stf = np.ascontiguousarray(np.asarray(stf_rows, np.float32)[:, : self.nt, :])
        _check(self.lib.ses3d_hip_set_adjoint_sources(self.h, ev, len(loc), _fptr(loc), _fptr(stf)))

    def load_event(self, ev: int, e) -> None:
        """Set source, receivers or adjoint sources of one loaded event."""
        if self.cfg.adjoint_flag in (0, 1):
            self.set_source(ev, e.source_type, e.xxs, e.yys, e.zzs, e.moment, e.stf)
            self.set_receivers(ev, e.rec32)
        else:
            self.set_adjoint_sources(ev, e.ad_loc32, e.ad_stf)

    def init(self):
        _check(self.lib.ses3d_hip_init(self.h))

    # ---------------------------------------------------------------- stepping
    def run(self, it_begin: int = 1, n_steps: Optional[int] = None, sync: bool = True):
        if n_steps is None:
            n_steps = self.nt - it_begin + 1
        _check(self.lib.ses3d_hip_run(self.h, it_begin, n_steps))
        if sync:
            self.synchronize()

    def autotune(self, it_begin: int = 1, n_steps: int = 24, sync: bool = True) -> dict:
        """Advance steps it_begin..it_begin+n_steps-1 (as ``run`` would) while
        timing the trial launch schedules; keeps the fastest.  Returns
        ``schedule()``."""
        _check(self.lib.ses3d_hip_autotune(self.h, it_begin, n_steps))
        if sync:
            self.synchronize()
        return self.schedule()

    def set_schedule(self, epb: int = 2, node_threads: int = 256, remap_elem: int = 1, remap_node: int = 1):
        v = np.array([epb, node_threads, remap_elem, remap_node], np.int32)
        _check(self.lib.ses3d_hip_set_schedule(self.h, v.ctypes.data_as(C.POINTER(C.c_int32))))

    def schedule(self) -> dict:
        return _schedule(self.lib.ses3d_hip_get_schedule, self.h)

    def set_streaming(self, mode: int = -1):
        """Element-kernel state streams: -1 non-temporal when they exceed the
        last-level cache (default), 0 always cached, 1 always non-temporal."""
        _check(self.lib.ses3d_hip_set_streaming(self.h, int(mode)))

    def streaming(self) -> dict:
        return _streaming(self.lib.ses3d_hip_get_streaming, self.h)

    def synchronize(self):
        _check(self.lib.ses3d_hip_synchronize(self.h))

    # ---------------------------------------------------------------- outputs
    def nr(self, ev: int = 0) -> int:
        n = C.c_int32()
        _check(self.lib.ses3d_hip_get_nr(self.h, ev, C.byref(n)))
        return n.value

    def receiver_index(self, ev: int = 0) -> np.ndarray:
        out = np.zeros(max(self.nr(ev), 1), np.int32)
        _check(self.lib.ses3d_hip_get_receiver_index(self.h, ev, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out[: self.nr(ev)]

    def seismograms(self, ev: int = 0) -> np.ndarray:
        nr = self.nr(ev)
        out = np.zeros(3 * max(nr, 1) * self.nt, np.float32)
        _check(self.lib.ses3d_hip_get_seismograms(self.h, ev, _fptr(out)))
        return out[: 3 * nr * self.nt].reshape(3, nr, self.nt)

    def field(self, name: str, ev: int = 0) -> np.ndarray:
        if name in _FIELDS_GLOBAL:
            shape = self.shape_global
            out = np.zeros(int(np.prod(shape)), np.float32)
        elif name in _FIELDS_MEM:
            out = np.zeros(3 * self.n_points, np.float32)
            _check(self.lib.ses3d_hip_get_field(self.h, ev, name.encode(), _fptr(out)))
            n = self.n_points
            return np.stack([out[k * n:(k + 1) * n].reshape(self.shape_local, order="F") for k in range(3)])
        else:
            shape = self.shape_local
            out = np.zeros(self.n_points, np.float32)
        _check(self.lib.ses3d_hip_get_field(self.h, ev, name.encode(), _fptr(out)))
        return out.reshape(shape, order="F")

    def kernels(self, ev: int = 0, with_q: Optional[bool] = None) -> dict:
        if with_q is None:
            with_q = bool(self.cfg.is_diss)
        names = ["grad_rho", "grad_cp", "grad_csh", "grad_csv"] + (["grad_Q_mu"] if with_q else [])
        outs = {n: np.zeros(self.n_points, np.float64) for n in names}
        ptrs = [outs[n].ctypes.data_as(C.POINTER(C.c_double)) if n in outs else None
                for n in ("grad_rho", "grad_cp", "grad_csh", "grad_csv", "grad_Q_mu")]
        _check(self.lib.ses3d_hip_get_kernels(self.h, ev, *ptrs))
        return {n: a.reshape(self.shape_local, order="F") for n, a in outs.items()}

    def set_history_degree(self, deg: int):
        """Polynomial degree (0..4) of the stored forward fields, as the `deg`
        argument of SOURCE/ses3d_output.f90:5-267 (the reference passes 4):
        with deg < 4 a stored step keeps each element's field at the (deg+1)^3
        degree-deg knots and the kernels read it interpolated back (lossy,
        (deg+1)^3/125 of the memory).  Before init; the same in the forward
        and the adjoint context."""
        _check(self.lib.ses3d_hip_set_history_degree(self.h, int(deg)))
        self.hdeg = int(deg)

    def set_kernel_precision(self, bits: int):
        """Kernel accumulators of an adjoint context: 64 (float64, default) or
        32 (float32 like the reference's grad_* arrays: its kernels bit for
        bit).  Takes effect at the next init."""
        _check(self.lib.ses3d_hip_set_kernel_precision(self.h, int(bits)))
        self.kernel_bits = int(bits)

    def _history_shape(self):
        d = getattr(self, "hdeg", 4)
        return self.shape_local if d == 4 else self.shape_local[:3] + (d + 1,) * 3

    def history(self, step: int, ev: int = 0) -> np.ndarray:
        """Stored forward step: (9,) + shape_local, or with degree d < 4 the
        (9, nx+1, ny+1, nz+1, d+1, d+1, d+1) values ses3d_store writes."""
        shape = self._history_shape()
        n = int(np.prod(shape))
        size = C.c_int64()
        _check(self.lib.ses3d_hip_history_size(self.h, C.byref(size)))
        assert size.value == 9 * n
        out = np.zeros(9 * n, np.float32)
        _check(self.lib.ses3d_hip_get_history(self.h, ev, step, _fptr(out)))
        return np.stack([out[k * n:(k + 1) * n].reshape(shape, order="F") for k in range(9)])

    def put_history(self, step: int, fields9: np.ndarray, ev: int = 0):
        shape = self._history_shape()
        if tuple(np.shape(fields9)) != (9,) + tuple(shape):
            raise Ses3dError(f"put_history: expected shape {(9,) + tuple(shape)}, got {np.shape(fields9)}")
        buf = np.concatenate([_ref_f32(fields9[k]) for k in range(9)])
        _check(self.lib.ses3d_hip_put_history(self.h, ev, step, _fptr(buf)))

    def saving_vector(self) -> np.ndarray:
        out = np.zeros(self.nt, np.int32)
        _check(self.lib.ses3d_hip_get_saving_vector(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def forward_checkpointed(self, k: int, keep: int = 1):
        """Forward pass with restart states every k steps; the sampled steps of
        the last ``keep`` segments stay in HBM, so the adjoint replays only the
        earlier segments (kernels identical for every ``keep``)."""
        _check(self.lib.ses3d_hip_set_checkpoint_keep(self.h, int(keep)))
        _check(self.lib.ses3d_hip_forward_checkpointed(self.h, int(k)))
        self.synchronize()

    def prepare_checkpoints(self, k: int, keep: int = 1):
        """Allocate and clear the restart states for interval k and the
        sampled-step buffer for ``keep`` kept segments ahead of
        forward_checkpointed(k, keep), which then allocates nothing; discards
        an earlier checkpointed forward."""
        _check(self.lib.ses3d_hip_set_checkpoint_keep(self.h, int(keep)))
        _check(self.lib.ses3d_hip_prepare_checkpoints(self.h, int(k)))

    def adjoint_checkpointed(self, fwd: "Solver"):
        _check(self.lib.ses3d_hip_adjoint_checkpointed(self.h, fwd.h))
        self.synchronize()

    def project(self, field, bx_deg, by_deg, bz_km, no_alloc: bool = False) -> np.ndarray:
        """Block-grid projection of a reference-layout field as
        TOOLS/SOURCE/project_kernel.f90 projects a kernel (one subvolume);
        returns (nbx-1, nby-1, nbz-1) float32.  no_alloc stages the data in
        the library's fixed per-GPU scratch instead of a per-call device
        buffer (ses3d_hip_project_noalloc; at most 64 MiB)."""
        f = _ref_f32(field)
        e = [np.ascontiguousarray(a, np.float32) for a in (bx_deg, by_deg, bz_km)]
        out = np.zeros((len(e[0]) - 1) * (len(e[1]) - 1) * (len(e[2]) - 1), np.float32)
        fn = self.lib.ses3d_hip_project_noalloc if no_alloc else self.lib.ses3d_hip_project
        _check(fn(self.h, _fptr(f), len(e[0]), _fptr(e[0]), len(e[1]), _fptr(e[1]), len(e[2]), _fptr(e[2]),
                  _fptr(out)))
        return out.reshape(len(e[0]) - 1, len(e[1]) - 1, len(e[2]) - 1)

    def project_kernels(self, bx_deg, by_deg, bz_km, ev: int = 0) -> dict:
        """This context's kernels, rounded to float32 as the reference writes
        them, projected onto the block grid: {'gradient_csh', 'gradient_csv',
        'gradient_cp', 'gradient_rho'[, 'gradient_Q_mu']}."""
        K = self.kernels(ev)
        names = {"grad_csh": "gradient_csh", "grad_csv": "gradient_csv", "grad_cp": "gradient_cp",
                 "grad_rho": "gradient_rho", "grad_Q_mu": "gradient_Q_mu"}
        return {names[k]: self.project(v.astype(np.float32), bx_deg, by_deg, bz_km) for k, v in K.items()}

    def memory_bytes(self) -> int:
        b = C.c_int64()
        _check(self.lib.ses3d_hip_memory_bytes(self.h, C.byref(b)))
        return b.value

    def info(self) -> dict:
        out = (C.c_int64 * 8)()
        _check(self.lib.ses3d_hip_get_info(self.h, out, 8))
        return {"fast_div_tau": bool(out[0]), "fast_div_3": bool(out[1]), "div_check_failures": list(out[2:6]),
                "anisotropic": bool(out[6]), "graph_steps": int(out[7])}

    def set_graph_steps(self, steps: int):
        _check(self.lib.ses3d_hip_set_graph_steps(self.h, int(steps)))

    def set_profiling(self, on: bool):
        _check(self.lib.ses3d_hip_set_profiling(self.h, int(bool(on))))

    def timing(self) -> dict:
        t, e, n = C.c_double(), C.c_double(), C.c_double()
        k = C.c_int32()
        _check(self.lib.ses3d_hip_get_timing(self.h, C.byref(t), C.byref(e), C.byref(n), C.byref(k)))
        return {"total_ms": t.value, "element_ms": e.value, "node_ms": n.value, "n_element": k.value}


def _schedule(fn, h) -> dict:
    v = np.zeros(4, np.int32)
    ms = np.zeros(12, np.float32)
    _check(fn(h, v.ctypes.data_as(C.POINTER(C.c_int32)), _fptr(ms)))
    return {"epb": int(v[0]), "node_threads": int(v[1]), "remap_elem": int(v[2]), "remap_node": int(v[3]),
            "trial_ms": ms.reshape(6, 2).tolist() if ms[0] >= 0 else None}


def _streaming(fn, h) -> dict:
    mode, in_use = C.c_int32(), C.c_int32()
    _check(fn(h, C.byref(mode), C.byref(in_use)))
    return {"mode": mode.value, "non_temporal": bool(in_use.value)}


class Group:
    """Sections (initialised Solvers, each with its own mesh and events)
    advanced together: one element-kernel and one node-kernel launch per step
    for all of them (ses3d_hip_group_* in include/ses3d_hip.h).  Each member
    gets exactly the results ``member.run`` would give; outputs are read from
    the members."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        _check(self.lib.ses3d_hip_group_create(arr, len(self.members), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ses3d_hip_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, it_begin: int = 1, n_steps: Optional[int] = None, sync: bool = True):
        if n_steps is None:
            n_steps = min(m.nt for m in self.members) - it_begin + 1
        _check(self.lib.ses3d_hip_group_run(self.h, it_begin, n_steps))
        if sync:
            self.synchronize()

    def autotune(self, it_begin: int = 1, n_steps: int = 24, sync: bool = True) -> dict:
        _check(self.lib.ses3d_hip_group_autotune(self.h, it_begin, n_steps))
        if sync:
            self.synchronize()
        return self.schedule()

    def synchronize(self):
        _check(self.lib.ses3d_hip_group_synchronize(self.h))

    def set_schedule(self, epb: int = 2, node_threads: int = 256, remap_elem: int = 1, remap_node: int = 1):
        v = np.array([epb, node_threads, remap_elem, remap_node], np.int32)
        _check(self.lib.ses3d_hip_group_set_schedule(self.h, v.ctypes.data_as(C.POINTER(C.c_int32))))

    def schedule(self) -> dict:
        return _schedule(self.lib.ses3d_hip_group_get_schedule, self.h)

    def set_streaming(self, mode: int = -1):
        _check(self.lib.ses3d_hip_group_set_streaming(self.h, int(mode)))

    def streaming(self) -> dict:
        return _streaming(self.lib.ses3d_hip_group_get_streaming, self.h)

    def set_graph_steps(self, steps: int):
        _check(self.lib.ses3d_hip_group_set_graph_steps(self.h, int(steps)))

    def set_profiling(self, on: bool):
        _check(self.lib.ses3d_hip_group_set_profiling(self.h, int(bool(on))))

    def timing(self) -> dict:
        t, e, n = C.c_double(), C.c_double(), C.c_double()
        k = C.c_int32()
        _check(self.lib.ses3d_hip_group_get_timing(self.h, C.byref(t), C.byref(e), C.byref(n), C.byref(k)))
        return {"total_ms": t.value, "element_ms": e.value, "node_ms": n.value, "n_element": k.value}
