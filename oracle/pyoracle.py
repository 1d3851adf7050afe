"""ctypes wrapper of the C oracle (libses3d_oracle32.so / libses3d_oracle64.so)
and a runner for the reference binary built under oracle/_ref.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke check and the
CPU-baseline leg of the benchmark; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_EXE = os.path.join(HERE, "_ref", "ses3d.exe")
# the reference objects with SOURCE/ses3d_grad_Q.f90 in place of ses3d_grad.f90,
# driven by ref_grad_q_harness.f90 (adjoint passes; also writes grad_Q_mu_<rank>)
REF_GRADQ_EXE = os.path.join(HERE, "_ref", "ses3d_grad_q.exe")


class OParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("pml", C.c_int32), ("is_diss", C.c_int32), ("adjoint_flag", C.c_int32),
        ("samp_ad", C.c_int32), ("nt", C.c_int32), ("dt", C.c_float),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("zmin", C.c_float), ("zmax", C.c_float), ("zmax_global", C.c_float),
        ("tau_p", C.c_float * 3), ("D_p", C.c_float * 3),
    ]


_LIBS = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(precision: int):
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(HERE, f"libses3d_oracle{precision}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    lib.o_create.restype = P
    lib.o_create.argtypes = [C.POINTER(OParams)]
    lib.o_destroy.argtypes = [P]
    lib.o_set_model.argtypes = [P] + [fp] * 7
    lib.o_set_source.argtypes = [P, C.c_int, C.c_float, C.c_float, C.c_float, fp, fp]
    lib.o_set_receivers.argtypes = [P, C.c_int, fp]
    lib.o_set_adjoint_sources.argtypes = [P, C.c_int, fp, fp]
    lib.o_init.argtypes = [P]
    lib.o_step.argtypes = [P, C.c_int, P]
    lib.o_run.argtypes = [P, C.c_int, C.c_int, P]
    lib.o_np.restype = C.c_int64
    lib.o_np.argtypes = [P]
    lib.o_ng.restype = C.c_int64
    lib.o_ng.argtypes = [P]
    lib.o_nr.argtypes = [P]
    lib.o_nad.argtypes = [P]
    lib.o_get.argtypes = [P, C.c_char_p, dp]
    lib.o_get.restype = C.c_int
    lib.o_get_setup.argtypes = [P, dp]
    lib.o_get_receivers.argtypes = [P, dp]
    lib.o_get_history.argtypes = [P, C.c_int, fp]
    lib.o_get_history.restype = C.c_int
    rp = fp if precision == 32 else dp
    lib.o_compress.argtypes = [C.c_int, C.c_int64, rp, rp]
    lib.o_restore.argtypes = [C.c_int, C.c_int64, rp, rp]
    lib.o_set_history_degree.argtypes = [P, C.c_int]
    lib.o_set_history_degree.restype = C.c_int
    lib.o_history_size.argtypes = [P]
    lib.o_history_size.restype = C.c_int64
    lib.o_saving_vector.argtypes = [P, C.c_int]
    lib.o_lgll.restype = C.c_double
    lib.o_lgll.argtypes = [C.c_int, C.c_double]
    lib.o_dlgll.restype = C.c_double
    lib.o_dlgll.argtypes = [C.c_int, C.c_int]
    _LIBS[precision] = lib
    return lib


def _f(a: Optional[np.ndarray]):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F") if a.ndim > 1 else np.asarray(a, np.float32))
    return a


class Oracle:
    """One reference-style simulation (one event, one rank box) on the CPU.

    ``model`` arrays are in reference layout (nx+1,ny+1,nz+1,5,5,5)."""

    def __init__(self, *, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, zmax_global, nt, dt,
                 is_diss=0, adjoint_flag=0, samp_ad=1, pml=2, tau_p=(1.0, 1.0, 1.0),
                 D_p=(0.0, 0.0, 0.0), precision=32):
        self.lib = _lib(precision)
        p = OParams(nx, ny, nz, pml, is_diss, adjoint_flag, samp_ad, nt, dt, xmin, xmax, ymin, ymax,
                    zmin, zmax, zmax_global, (C.c_float * 3)(*tau_p), (C.c_float * 3)(*D_p))
        self.p = p
        self.h = self.lib.o_create(C.byref(p))
        self.shape_local = (nx + 1, ny + 1, nz + 1, 5, 5, 5)
        self.shape_global = (4 * (nx + 1) + 1, 4 * (ny + 1) + 1, 4 * (nz + 1) + 1)
        self._keep = []

    def close(self):
        if self.h:
            self.lib.o_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptr(self, a):
        if a is None:
            return None
        self._keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def set_model(self, m: dict):
        arrs = [_f(m[k]) for k in ("rhoinv", "mu", "lambda", "A", "B", "C")]
        q = _f(m.get("Q")) if "Q" in m else None
        self.lib.o_set_model(self.h, *[self._ptr(a) for a in arrs], self._ptr(q))

    def set_source(self, source_type, xxs_rad, yys_rad, zzs, mom6, so):
        self.lib.o_set_source(self.h, int(source_type), float(np.float32(xxs_rad)), float(np.float32(yys_rad)),
                              float(np.float32(zzs)), self._ptr(np.asarray(mom6, np.float32)),
                              self._ptr(np.ascontiguousarray(so, np.float32)))

    def set_receivers(self, rec32: np.ndarray):
        rec32 = np.ascontiguousarray(rec32, np.float32).reshape(-1, 3)
        self.lib.o_set_receivers(self.h, len(rec32), self._ptr(rec32))

    def set_adjoint_sources(self, loc32: np.ndarray, stf_rows: np.ndarray):
        """loc32 (n,3); stf_rows (n, nt, 3) as read from ad_src files."""
        loc32 = np.ascontiguousarray(loc32, np.float32).reshape(-1, 3)
        stf_rows = np.ascontiguousarray(stf_rows, np.float32)
        self.lib.o_set_adjoint_sources(self.h, len(loc32), self._ptr(loc32), self._ptr(stf_rows))

    def init(self):
        self.lib.o_init(self.h)

    def run(self, it0: int, n: int, fwd: "Oracle" = None):
        self.lib.o_run(self.h, it0, n, fwd.h if fwd is not None else None)

    def get(self, name: str) -> np.ndarray:
        if name in ("vxg", "vyg", "vzg", "MM_global", "prof_global", "taper_global"):
            out = np.zeros(int(np.prod(self.shape_global)))
            shape = self.shape_global
        elif name == "seismograms":
            nr = self.lib.o_nr(self.h)
            out = np.zeros(3 * nr * self.p.nt)
            rc = self.lib.o_get(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
            return out.reshape(3, nr, self.p.nt)
        elif name in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"):
            n = int(np.prod(self.shape_local))
            out = np.zeros(3 * n)
            rc = self.lib.o_get(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
            if rc != 0:
                raise KeyError(name)
            return np.stack([out[k * n:(k + 1) * n].reshape(self.shape_local, order="F") for k in range(3)])
        else:
            out = np.zeros(int(np.prod(self.shape_local)))
            shape = self.shape_local
        rc = self.lib.o_get(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise KeyError(name)
        return out.reshape(shape, order="F")

    def setup_values(self) -> dict:
        out = np.zeros(64)
        self.lib.o_get_setup(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {"w": out[0:5], "dl": out[5:30].reshape(5, 5), "dx": out[30], "dy": out[31],
                "dz": out[32], "Jac": out[33], "isx": int(out[34]), "isy": int(out[35]),
                "isz": int(out[36]), "xxs_loc": out[37], "yys_loc": out[38], "zzs_loc": out[39],
                "is_source": int(out[40])}

    def receivers(self) -> np.ndarray:
        nr = self.lib.o_nr(self.h)
        out = np.zeros(6 * max(nr, 1))
        self.lib.o_get_receivers(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:6 * nr].reshape(nr, 6)

    def set_history_degree(self, deg: int):
        """Degree of the stored forward fields (ses3d_store's `deg`); before the first step."""
        if self.lib.o_set_history_degree(self.h, int(deg)) != 0:
            raise ValueError(f"history degree {deg}")
        self.hdeg = int(deg)

    def history(self, step: int) -> np.ndarray:
        """Stored step: (9,) + shape_local, or with degree d < 4 the (9, nx+1,
        ny+1, nz+1, d+1, d+1, d+1) values ses3d_store writes."""
        d = getattr(self, "hdeg", 4)
        shape = self.shape_local if d == 4 else self.shape_local[:3] + (d + 1,) * 3
        n = int(np.prod(shape))
        assert 9 * n == self.lib.o_history_size(self.h)
        out = np.zeros(9 * n, np.float32)
        if self.lib.o_get_history(self.h, step, out.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise KeyError(step)
        return np.stack([out[k * n:(k + 1) * n].reshape(shape, order="F") for k in range(9)])

    def saving_vector(self) -> np.ndarray:
        return np.array([self.lib.o_saving_vector(self.h, it) for it in range(1, self.p.nt + 1)])


def lgll(i: int, x: float, precision: int = 32) -> float:
    return _lib(precision).o_lgll(i, x)


def dlgll(i: int, j: int, precision: int = 32) -> float:
    return _lib(precision).o_dlgll(i, j)


def _real(precision):
    return (np.float32, C.c_float) if precision == 32 else (np.float64, C.c_double)


def compress(deg: int, field: np.ndarray, precision: int = 32) -> np.ndarray:
    """ses3d_store's degree-deg values of one field: (nx+1, ny+1, nz+1, 5, 5, 5)
    -> (nx+1, ny+1, nz+1, deg+1, deg+1, deg+1), Fortran order."""
    dt, ct = _real(precision)
    f = np.asarray(field, dt).ravel(order="F")
    ne = f.size // 125
    c = np.zeros(ne * (deg + 1) ** 3, dt)
    _lib(precision).o_compress(deg, ne, f.ctypes.data_as(C.POINTER(ct)), c.ctypes.data_as(C.POINTER(ct)))
    return c.reshape(field.shape[:3] + (deg + 1,) * 3, order="F")


def restore(deg: int, values: np.ndarray, precision: int = 32) -> np.ndarray:
    """ses3d_restore: degree-deg values back to the 125 GLL nodes."""
    dt, ct = _real(precision)
    c = np.asarray(values, dt).ravel(order="F")
    ne = c.size // (deg + 1) ** 3
    f = np.zeros(ne * 125, dt)
    _lib(precision).o_restore(deg, ne, c.ctypes.data_as(C.POINTER(ct)), f.ctypes.data_as(C.POINTER(ct)))
    return f.reshape(values.shape[:3] + (5, 5, 5), order="F")


# ---------------------------------------------------------------------------
# reference binary
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_EXE) and os.path.exists(REF_GRADQ_EXE)


def build_ref() -> bool:
    """Build oracle/_ref/ses3d.exe from the reference sources where they lie.
    Returns False when the reference tree is absent (a machine that received
    only this repository)."""
    if not os.path.isdir("/root/reference/SOURCE"):
        return ref_available()
    r = subprocess.run(["make", "-s", "-C", HERE, "ref"], capture_output=True, text=True)
    return r.returncode == 0 and ref_available()


def build_ref_multirank(per_rank, grad_q: bool = False, pml: Optional[int] = None) -> str:
    """Build the reference solver for runs whose rank boxes hold per_rank =
    (ex, ey, ez) elements each (its static extents nx_max, ny_max, nz_max set
    to ex-1, ey-1, ez-1, as the reference's workflow sets them per run; `make
    ref-mr`).  Returns the path of the solver, or with grad_q of the grad_Q
    build (ses3d_grad_Q.f90 and the harness that writes grad_Q_mu).  ``pml``
    also sets the build's absorbing width (SOURCE/ses3d_modules.f90:21,
    default 2)."""
    tag = "_".join(str(int(v) - 1) for v in per_rank) + (f"_{int(pml)}" if pml is not None else "")
    exe = os.path.join(HERE, "_ref", f"ses3d_{'grad_q_' if grad_q else ''}mr_{tag}.exe")
    r = subprocess.run(["make", "-s", "-C", HERE, "ref-mr", f"MR={tag}"], capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(exe):
        raise RuntimeError(f"reference build for {per_rank} elements per rank failed:\n{r.stdout[-2000:]}{r.stderr[-2000:]}")
    return exe


def run_ref(root: str, timeout: float = 600.0, nprocs: int = 1, exe: str = REF_EXE) -> subprocess.CompletedProcess:
    """Run the reference binary (or the grad_Q build, exe=REF_GRADQ_EXE) on a
    run tree (cwd = root/MAIN)."""
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", "1")
    cmd = [exe]
    if nprocs > 1:
        cmd = [shutil.which("mpiexec") or "/opt/conda/bin/mpiexec", "-n", str(nprocs), exe]
    def _stack():
        # the reference keeps large automatic arrays on the stack
        # (e.g. SOURCE/ses3d_evolution.f90:19-21)
        import resource
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    return subprocess.run(cmd, cwd=os.path.join(root, "MAIN"), capture_output=True, text=True,
                          timeout=timeout, env=env, preexec_fn=_stack)
