"""Readers and writers for the SES3D on-disk formats.

Every text input is read the way the reference's list-directed Fortran reads
read it: only the first whitespace-separated token of a line counts, so
trailing ``! comments`` are ignored; lines read with ``(A1000)`` keep the
whole line.  Binary model, snapshot and kernel files are Fortran unformatted
sequential records: ``int32 nbytes | payload | int32 nbytes``, with the
payload a float32 array in Fortran (column-major) order.

Reference readers/writers followed here (file:line):
  setup        SOURCE/ses3d_input.f90:259-317
  event_<i>    SOURCE/ses3d_input.f90:196-251
  relax        SOURCE/ses3d_input.f90:333-357
  stf          SOURCE/ses3d_input.f90:415-430 (4 header lines skipped)
  recfile_<i>  SOURCE/ses3d_input.f90:531-546
  event_list   SOURCE/ses3d_main.f90:35-43
  boxfile      SOURCE/ses3d_input.f90:49-146, writer MODELS/SOURCE/generate_models.f90:210-282
  model files  SOURCE/ses3d_input.f90:472-501, writer MODELS/SOURCE/generate_models.f90:429-452
  seismograms  SOURCE/ses3d_output.f90:398-451, reader TOOLS/seismograms.py:37-98
  ad_srcfile   SOURCE/ses3d_input.f90:592-611
  ad_src_<k>   SOURCE/ses3d_init.f90:475-492, writer TOOLS/adjoint_source.py:76-93
  saving_vector SOURCE/ses3d_output.f90:379-392, reader SOURCE/ses3d_input.f90:562-578
  snapshots/kernels SOURCE/ses3d_output.f90:293-343
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Sequence

import numpy as np

# Single-precision pi exactly as the reference declares it
# (SOURCE/ses3d_modules.f90:25, ``real, parameter :: pi=3.1415926535898``).
PI32 = np.float32(3.1415926535898)
LPD = 4


def _tok(line: str) -> str:
    s = line.strip()
    if not s:
        raise ValueError("empty line where a value was expected")
    return s.split()[0]


def f32_array(tokens: Sequence[str]) -> np.ndarray:
    """Parse tokens to float32 as a Fortran ``real`` read does: the decimal
    value rounded once, to nearest float32.  Rounding to float64 first and
    then to float32 gives the same result except where the float64 value lies
    exactly halfway between two float32 values (the decimal can lie on either
    side of it); those tokens are decided on their exact decimal value."""
    toks = [t.strip().replace("d", "e").replace("D", "e") for t in tokens]
    x = np.array([float(t) for t in toks], np.float64)
    f = x.astype(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        up = np.nextafter(f, np.float32(np.inf))
        dn = np.nextafter(f, np.float32(-np.inf))
        other = np.where(x > f.astype(np.float64), up, dn)
        half = (f.astype(np.float64) + other.astype(np.float64)) * 0.5 == x
    for i in np.flatnonzero(half & np.isfinite(f) & (x != f.astype(np.float64))):
        exact, mid = Fraction(toks[i]), Fraction(float(x[i]))
        if exact != mid and (exact > mid) == (other[i] > f[i]):
            f[i] = other[i]
    return f


def f32(s: str) -> np.float32:
    """Parse one token to float32 (Fortran list-directed ``real`` read)."""
    return f32_array([s])[0]


def deg2rad32(v) -> np.float32:
    """``x*pi/180`` in float32, left to right (SOURCE/ses3d_input.f90:319-325)."""
    return np.float32(np.float32(np.float32(v) * PI32) / np.float32(180.0))


def rad2deg32(v) -> np.float32:
    """``x*180/pi`` in float32 (SOURCE/ses3d_output.f90:425)."""
    return np.float32(np.float32(np.float32(v) * np.float32(180.0)) / PI32)


class _Lines:
    def __init__(self, path: str):
        with open(path, "r") as fh:
            self.lines = fh.read().splitlines()
        self.pos = 0
        self.path = path

    def next(self) -> str:
        while True:
            if self.pos >= len(self.lines):
                raise EOFError(f"{self.path}: unexpected end of file")
            line = self.lines[self.pos]
            self.pos += 1
            # list-directed reads skip blank records when looking for a value
            return line

    def value(self) -> str:
        while True:
            line = self.next()
            if line.strip():
                return _tok(line)

    def skip(self, n: int = 1) -> None:
        """n header reads ``read(u,*) junk`` of a character item: list-directed
        input passes over blank records, so each consumes the next non-blank
        record (a blank first line of an ad_src file is not a header line)."""
        for _ in range(n):
            self.value()


# ---------------------------------------------------------------------------
# setup
# ---------------------------------------------------------------------------

@dataclass
class Setup:
    """Content of ``INPUT/setup`` (SOURCE/ses3d_input.f90:264-313).

    Angles are kept in degrees as written; ``*_rad`` give the float32 values
    the reference holds after its conversion at ses3d_input.f90:322-325.
    """
    theta_min: float = 40.0
    theta_max: float = 60.0
    phi_min: float = 10.0
    phi_max: float = 30.0
    z_min: float = 5371000.0
    z_max: float = 6371000.0
    is_diss: int = 0
    model_type: int = 1
    nx_global: int = 23
    ny_global: int = 23
    nz_global: int = 11
    lpd: int = 4
    px: int = 1
    py: int = 1
    pz: int = 1
    adjoint_flag: int = 0
    samp_ad: int = 15
    ffd: str = "../DATA/OUTPUT/ADJOINT"

    def write(self, path: str) -> None:
        lines = [
            "MODEL " + "=" * 60,
            f"{self.theta_min!r:<44}! theta_min (colatitude) in degrees",
            f"{self.theta_max!r:<44}! theta_max (colatitude) in degrees",
            f"{self.phi_min!r:<44}! phi_min (longitude) in degrees",
            f"{self.phi_max!r:<44}! phi_max (longitude) in degrees",
            f"{self.z_min!r:<44}! z_min (radius) in m",
            f"{self.z_max!r:<44}! z_max (radius) in m",
            f"{self.is_diss:<44d}! is_diss",
            f"{self.model_type:<44d}! model_type",
            "COMPUTATIONAL SETUP (PARALLELISATION) " + "=" * 40,
            f"{self.nx_global:<44d}! nx_global, (nx_global+px = global # elements in theta direction)",
            f"{self.ny_global:<44d}! ny_global, (ny_global+py = global # elements in phi direction)",
            f"{self.nz_global:<44d}! nz_global, (nz_global+pz = global # of elements in r direction)",
            f"{self.lpd:<44d}! lpd, LAGRANGE polynomial degree",
            f"{self.px:<44d}! px, processors in theta direction",
            f"{self.py:<44d}! py, processors in phi direction",
            f"{self.pz:<44d}! pz, processors in r direction",
            "ADJOINT PARAMETERS " + "=" * 60,
            f"{self.adjoint_flag:<44d}! adjoint_flag (0=normal simulation, 1=adjoint forward, 2=adjoint reverse)",
            f"{self.samp_ad:<44d}! samp_ad, sampling rate of forward field",
            self.ffd,
        ]
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")

    @classmethod
    def read(cls, path: str) -> "Setup":
        L = _Lines(path)
        s = cls()
        L.skip()
        s.theta_min = float(L.value()); s.theta_max = float(L.value())
        s.phi_min = float(L.value()); s.phi_max = float(L.value())
        s.z_min = float(L.value()); s.z_max = float(L.value())
        s.is_diss = int(L.value()); s.model_type = int(L.value())
        L.skip()
        s.nx_global = int(L.value()); s.ny_global = int(L.value()); s.nz_global = int(L.value())
        s.lpd = int(L.value())
        s.px = int(L.value()); s.py = int(L.value()); s.pz = int(L.value())
        L.skip()
        s.adjoint_flag = int(L.value()); s.samp_ad = int(L.value())
        s.ffd = L.next().rstrip()
        return s

    @property
    def zmax_global(self) -> np.float32:
        return f32(repr(self.z_max))


# ---------------------------------------------------------------------------
# event
# ---------------------------------------------------------------------------

@dataclass
class Event:
    """Content of ``INPUT/event_<i>`` (SOURCE/ses3d_input.f90:201-249).

    ``moment`` is (M_tt, M_pp, M_rr, M_tp, M_tr, M_pr) in N m.
    """
    nt: int = 2000
    dt: float = 0.2
    xxs: float = 50.0           # colatitude [deg]
    yys: float = 20.0           # longitude [deg]
    zzs: float = 10000.0        # depth [m]
    source_type: int = 10
    moment: Sequence[float] = (1e18, 1e18, 1e18, 0.0, 0.0, 0.0)
    ofd: str = "../DATA/OUTPUT/"
    ssamp: int = 100000
    output_displacement: int = 0

    def write(self, path: str) -> None:
        m = [repr(float(v)) for v in self.moment]
        lines = [
            "SIMULATION PARAMETERS " + "=" * 60,
            f"{self.nt:<40d}! nt, number of time steps",
            f"{self.dt!r:<40}! dt in sec, time increment",
            "SOURCE " + "=" * 60,
            f"{self.xxs!r:<40}! xxs, theta-coord. center of source in degrees",
            f"{self.yys!r:<40}! yys, phi-coord. center of source in degrees",
            f"{self.zzs!r:<40}! zzs, source depth in (m)",
            f"{self.source_type:<40d}! srctype, 1:f_x, 2:f_y, 3:f_z, 10:M_ij",
            f"{m[0]:<40}! M_theta_theta",
            f"{m[1]:<40}! M_phi_phi",
            f"{m[2]:<40}! M_r_r",
            f"{m[3]:<40}! M_theta_phi",
            f"{m[4]:<40}! M_theta_r",
            f"{m[5]:<40}! M_phi_r",
            "OUTPUT DIRECTORY " + "=" * 60,
            self.ofd,
            "OUTPUT FLAGS " + "=" * 60,
            f"{self.ssamp:<40d}! ssamp, snapshot sampling",
            f"{self.output_displacement:<40d}! output_displacement (1=yes,0=no)",
        ]
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")

    @classmethod
    def read(cls, path: str) -> "Event":
        L = _Lines(path)
        e = cls()
        L.skip()
        e.nt = int(L.value()); e.dt = float(L.value())
        L.skip()
        e.xxs = float(L.value()); e.yys = float(L.value()); e.zzs = float(L.value())
        e.source_type = int(L.value())
        e.moment = tuple(float(L.value()) for _ in range(6))
        L.skip()
        e.ofd = L.next().rstrip()
        L.skip()
        e.ssamp = int(L.value()); e.output_displacement = int(L.value())
        return e


# ---------------------------------------------------------------------------
# relax, stf, event_list
# ---------------------------------------------------------------------------

def write_relax(path: str, tau_p: Sequence[float], D_p: Sequence[float]) -> None:
    with open(path, "w") as fh:
        fh.write("RELAXATION TIMES [s] =====================\n")
        for t in tau_p:
            fh.write(f"{float(t)!r}\n")
        fh.write("WEIGHTS OF RELAXATION MECHANISMS =========\n")
        for d in D_p:
            fh.write(f"{float(d)!r}\n")


def read_relax(path: str):
    """-> (tau_p[3], D_p[3]) float32 (SOURCE/ses3d_input.f90:338-351)."""
    L = _Lines(path)
    L.skip()
    tau = np.array([f32(L.value()) for _ in range(3)], np.float32)
    L.skip()
    D = np.array([f32(L.value()) for _ in range(3)], np.float32)
    return tau, D


def write_stf(path: str, samples: np.ndarray, dt: float = 0.0, note: str = "") -> None:
    """STF with the 4-line header the reader skips (TOOLS/make_stf.py:66-72)."""
    samples = np.asarray(samples, np.float32)
    with open(path, "w") as fh:
        fh.write("source time function, ses3d version 4.1\n")
        fh.write(f"nt= {len(samples)}, dt={dt}\n")
        fh.write((note or "synthetic") + "\n")
        fh.write("-- samples --\n")
        for v in samples:
            fh.write(f"{float(v)!r}\n")


def read_stf(path: str, nt: int) -> np.ndarray:
    """4 header reads, then nt samples (SOURCE/ses3d_input.f90:419-428).  A file
    without the header (as INPUT/stf of the reference is) loses its first 4
    samples, as in the reference."""
    L = _Lines(path)
    L.skip(4)
    return f32_array([L.value() for _ in range(nt)])


def write_event_list(path: str, events: Sequence[str]) -> None:
    with open(path, "w") as fh:
        fh.write(f"{len(events)}\n")
        for e in events:
            fh.write(f"{e}\n")


def read_event_list(path: str) -> List[str]:
    L = _Lines(path)
    n = int(L.value())
    return [L.value() for _ in range(n)]


# ---------------------------------------------------------------------------
# receivers
# ---------------------------------------------------------------------------

@dataclass
class Receivers:
    names: List[str] = field(default_factory=list)
    colat: np.ndarray = field(default_factory=lambda: np.zeros(0))   # deg
    lon: np.ndarray = field(default_factory=lambda: np.zeros(0))     # deg
    depth: np.ndarray = field(default_factory=lambda: np.zeros(0))   # m

    def write(self, path: str) -> None:
        with open(path, "w") as fh:
            fh.write(f"{len(self.names)}\n")
            for n, a, b, c in zip(self.names, self.colat, self.lon, self.depth):
                fh.write(f"{n[:12]:<12}\n{float(a)!r} {float(b)!r} {float(c)!r}\n")

    @classmethod
    def read(cls, path: str) -> "Receivers":
        """SOURCE/ses3d_input.f90:533-544: count, then (name, colat lon depth)."""
        L = _Lines(path)
        n = int(L.value())
        names, rows = [], []
        for _ in range(n):
            names.append(L.value()[:12])
            while True:
                line = L.next()
                if line.strip():
                    break
            rows.append([float(t) for t in line.split()[:3]])
        rows = np.array(rows, np.float64).reshape(-1, 3)
        return cls(names, rows[:, 0], rows[:, 1], rows[:, 2])

    def loc32(self) -> np.ndarray:
        """(nr,3) float32 (colat rad, lon rad, depth m) as the reference holds
        ``recloc_global`` after ses3d_input.f90:541-542."""
        out = np.zeros((len(self.names), 3), np.float32)
        for i in range(len(self.names)):
            out[i, 0] = deg2rad32(f32(repr(float(self.colat[i]))))
            out[i, 1] = deg2rad32(f32(repr(float(self.lon[i]))))
            out[i, 2] = f32(repr(float(self.depth[i])))
        return out


# ---------------------------------------------------------------------------
# boxfile
# ---------------------------------------------------------------------------

@dataclass
class Box:
    """One rank box of ``MODELS/MODELS/boxfile``; x/y in rad, z in m."""
    index: int
    multi: tuple
    nx_min: int
    nx_max: int
    ny_min: int
    ny_max: int
    nz_min: int
    nz_max: int
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    zmin: float
    zmax: float

    @property
    def nx(self) -> int:
        return self.nx_max - self.nx_min

    @property
    def ny(self) -> int:
        return self.ny_max - self.ny_min

    @property
    def nz(self) -> int:
        return self.nz_max - self.nz_min


def make_boxes(setup: Setup) -> List[Box]:
    """Rank boxes as MODELS/SOURCE/generate_models.f90:200-282 lays them out."""
    px, py, pz = setup.px, setup.py, setup.pz
    nx, ny, nz = setup.nx_global, setup.ny_global, setup.nz_global
    x0, x1 = np.deg2rad(setup.theta_min), np.deg2rad(setup.theta_max)
    y0, y1 = np.deg2rad(setup.phi_min), np.deg2rad(setup.phi_max)
    z0, z1 = setup.z_min, setup.z_max
    dz = (z1 - z0) / (nz + pz)
    dy = (y1 - y0) / (ny + py)
    dx = (x1 - x0) / (nx + px)
    bnx, bny, bnz = (nx + 1) // px, (ny + 1) // py, (nz + 1) // pz
    boxes = []
    for k in range(1, pz + 1):
        for j in range(1, py + 1):
            for i in range(1, px + 1):
                mi = (k - 1) * py * px + (j - 1) * px + i
                lo = ((i - 1) * bnx, (j - 1) * bny, (k - 1) * bnz)
                hi = (i * bnx if i != px else nx, j * bny if j != py else ny, k * bnz if k != pz else nz)
                boxes.append(Box(
                    mi, (i, j, k), lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                    float(x0 + (lo[0] + i - 1) * dx), float(x0 + (hi[0] + i) * dx),
                    float(y0 + (lo[1] + j - 1) * dy), float(y0 + (hi[1] + j) * dy),
                    float(z0 + (lo[2] + k - 1) * dz), float(z0 + (hi[2] + k) * dz)))
    return boxes


def write_boxfile(path: str, setup: Setup) -> List[Box]:
    boxes = make_boxes(setup)
    write_boxes(path, boxes, setup.px, setup.py, setup.pz)
    return boxes


def write_boxes(path: str, boxes: Sequence[Box], px: int, py: int, pz: int) -> None:
    """MODELS/MODELS/boxfile (MODELS/SOURCE/generate_models.f90:208-282):
    14 header lines, the process counts, then per box its index, multi-index,
    index bounds and physical bounds (values written so they read back exactly)."""
    hdr = ["- format description -", "total number of processes",
           "number of processes in x direction", "number of processes in y direction",
           "number of processes in z direction", "single index (npz-1)*py*px+(npy-1)*px+npx",
           "multi index (npx, npy, npz)", "index boundaries x (x_min, x_max)",
           "index boundaries y (y_min, y_max)", "index boundaries z (z_min, z_max)",
           "physical boundaries x", "physical boundaries y", "phyiscal boundaries z",
           "-------------------"]
    out = [" " + h for h in hdr]
    out += [f" {px * py * pz}", f" {px}", f" {py}", f" {pz}", " -------------------"]
    for b in boxes:
        out += [f" {b.index}", f" {b.multi[0]} {b.multi[1]} {b.multi[2]}",
                f" {b.nx_min} {b.nx_max}", f" {b.ny_min} {b.ny_max}", f" {b.nz_min} {b.nz_max}",
                f" {b.xmin!r} {b.xmax!r}", f" {b.ymin!r} {b.ymax!r}", f" {b.zmin!r} {b.zmax!r}",
                " -------------------"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read_boxfile(path: str) -> List[Box]:
    """SOURCE/ses3d_input.f90:51-99: 14 header lines, pp, px, py, pz, ---, boxes."""
    L = _Lines(path)
    L.skip(14)
    pp = int(L.value())
    L.value(); L.value(); L.value()
    L.skip()
    boxes = []
    for _ in range(pp):
        idx = int(L.value())
        multi = tuple(int(t) for t in L.next().split()[:3])
        a = [int(t) for t in L.next().split()[:2]]
        b = [int(t) for t in L.next().split()[:2]]
        c = [int(t) for t in L.next().split()[:2]]
        xs = [float(t) for t in L.next().split()[:2]]
        ys = [float(t) for t in L.next().split()[:2]]
        zs = [float(t) for t in L.next().split()[:2]]
        L.skip()
        boxes.append(Box(idx, multi, a[0], a[1], b[0], b[1], c[0], c[1], *xs, *ys, *zs))
    return boxes


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records
# ---------------------------------------------------------------------------

def write_record(fh, arr: np.ndarray) -> None:
    data = np.asarray(arr, dtype=np.float32).ravel(order="F").tobytes()
    n = np.array([len(data)], np.int32).tobytes()
    fh.write(n + data + n)


def read_records(path: str) -> List[bytes]:
    out = []
    with open(path, "rb") as fh:
        buf = fh.read()
    pos = 0
    while pos < len(buf):
        n = int(np.frombuffer(buf, np.int32, 1, pos)[0])
        payload = buf[pos + 4:pos + 4 + n]
        n2 = int(np.frombuffer(buf, np.int32, 1, pos + 4 + n)[0])
        if n2 != n:
            raise ValueError(f"{path}: corrupt record markers {n} != {n2}")
        out.append(payload)
        pos += n + 8
    return out


def write_field(path: str, arr: np.ndarray) -> None:
    """One record holding a (nx+1,ny+1,nz+1,5,5,5) float32 field, Fortran order."""
    with open(path, "wb") as fh:
        write_record(fh, arr)


def read_field(path: str, shape, dtype=np.float32) -> np.ndarray:
    """Read the first record as a Fortran-order array of ``shape``.

    Files written from the full static ``nx_max``-sized arrays (snapshots and
    kernels, SOURCE/ses3d_output.f90:304-337) are read with that full shape;
    model files (SOURCE/ses3d_input.f90:479-484) hold exactly (nx+1,...).
    """
    rec = read_records(path)[0]
    a = np.frombuffer(rec, dtype=dtype)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{path}: record holds {a.size} values, expected {int(np.prod(shape))}")
    return a.reshape(shape, order="F").copy()


MODEL_FIELDS = ("rhoinv", "mu", "lambda", "A", "B", "C")


def write_model(model_dir: str, model: dict, rank: int = 0, is_diss: bool = False) -> None:
    os.makedirs(model_dir, exist_ok=True)
    names = MODEL_FIELDS + (("Q",) if is_diss else ())
    for n in names:
        write_field(os.path.join(model_dir, f"{n}{rank}"), model[n])


def read_model(model_dir: str, shape, rank: int = 0, is_diss: bool = False) -> dict:
    names = MODEL_FIELDS + (("Q",) if is_diss else ())
    return {n: read_field(os.path.join(model_dir, f"{n}{rank}"), shape) for n in names}


# ---------------------------------------------------------------------------
# seismograms
# ---------------------------------------------------------------------------

def _fortran_real(v) -> str:
    """List-directed style float text (value round-trips through float32)."""
    return f"{float(np.float32(v)):.9G}"


def write_seismogram(ofd: str, name: str, trace_xyz: np.ndarray, nt: int, dt: float,
                     rec_deg, src_deg) -> None:
    """Three text files ``<name>.x/.y/.z`` (SOURCE/ses3d_output.f90:400-447).

    ``rec_deg`` = (colat deg, lon deg, depth m); ``src_deg`` likewise.
    The station name is padded to 12 characters as the reference's
    ``character(len=12)`` does.
    """
    sname = f"{name[:12]:<12}"
    titles = ("theta component seismograms", "phi component seismograms", "r component seismograms")
    for c, ext in enumerate("xyz"):
        with open(os.path.join(ofd, sname + "." + ext), "w") as fh:
            fh.write(f" {titles[c]}\n")
            fh.write(f" nt= {nt}\n")
            fh.write(f" dt= {_fortran_real(dt)}\n")
            fh.write(" receiver location (colat [deg],lon [deg],depth [m])\n")
            fh.write(f" x= {_fortran_real(rec_deg[0])} y= {_fortran_real(rec_deg[1])} z= {_fortran_real(rec_deg[2])}\n")
            fh.write(" source location (colat [deg],lon [deg],depth [m])\n")
            fh.write(f" x= {_fortran_real(src_deg[0])} y= {_fortran_real(src_deg[1])} z= {_fortran_real(src_deg[2])}\n")
            fh.write("".join(f" {_fortran_real(v)}\n" for v in trace_xyz[c, :nt]))


def read_seismogram(ofd: str, name: str):
    """-> dict(nt, dt, rec=(x,y,z), src=(x,y,z), trace=(3,nt) float64).

    Parses as TOOLS/seismograms.py:56-83 does (split on '=')."""
    sname = f"{name[:12]:<12}"
    out = {}
    traces = []
    for c, ext in enumerate("xyz"):
        with open(os.path.join(ofd, sname + "." + ext)) as fh:
            lines = fh.read().splitlines()
        if c == 0:
            out["nt"] = int(lines[1].strip().split("=")[1])
            out["dt"] = float(lines[2].strip().split("=")[1])
            ln = lines[4].strip().split("=")
            out["rec"] = (float(ln[1].split("y")[0]), float(ln[2].split("z")[0]), float(ln[3]))
            ln = lines[6].strip().split("=")
            out["src"] = (float(ln[1].split("y")[0]), float(ln[2].split("z")[0]), float(ln[3]))
        traces.append([float(v) for v in lines[7:7 + out["nt"]]])
    out["trace"] = np.array(traces, np.float64)
    return out


# ---------------------------------------------------------------------------
# adjoint sources, saving vector
# ---------------------------------------------------------------------------

def write_ad_srcfile(path: str, locs_deg: np.ndarray) -> None:
    locs_deg = np.asarray(locs_deg, np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{len(locs_deg)}\n")
        for a, b, c in locs_deg:
            fh.write(f"{float(a)!r} {float(b)!r} {float(c)!r}\n")


def read_ad_srcfile(path: str) -> np.ndarray:
    """-> (n,3) float32 (colat rad, lon rad, depth m), SOURCE/ses3d_input.f90:594-601."""
    L = _Lines(path)
    n = int(L.value())
    out = np.zeros((n, 3), np.float32)
    for k in range(n):
        while True:
            line = L.next()
            if line.strip():
                break
        t = line.split()
        out[k, 0] = deg2rad32(f32(t[0]))
        out[k, 1] = deg2rad32(f32(t[1]))
        out[k, 2] = f32(t[2])
    return out


def write_ad_src(path: str, loc_deg, adsrc_xyz_forward_time: np.ndarray) -> None:
    """Adjoint source file; samples written time-REVERSED as
    TOOLS/adjoint_source.py:84-91 does.  ``adsrc`` is (3, nt) in forward time."""
    a = np.asarray(adsrc_xyz_forward_time, np.float64)
    nt = a.shape[1]
    with open(path, "w") as fh:
        fh.write("-- adjoint source ------------------\n")
        fh.write("-- source coordinates (colat,lon,depth)\n")
        fh.write(f"{float(loc_deg[0])!r} {float(loc_deg[1])!r} {float(loc_deg[2])!r}\n")
        fh.write("-- source time function (x, y, z) --\n")
        for k in range(nt - 1, -1, -1):
            fh.write(f"{float(a[0, k])!r} {float(a[1, k])!r} {float(a[2, k])!r}\n")


def read_ad_src(path: str, nt: int) -> np.ndarray:
    """-> (nt,3) float32 rows as stored (already reversed), SOURCE/ses3d_init.f90:479-490:
    4 header reads (blank records skipped), then nt rows."""
    L = _Lines(path)
    L.skip(4)
    toks = []
    for i in range(nt):
        while True:
            line = L.next()
            if line.strip():
                break
        toks += line.split()[:3]
    return f32_array(toks).reshape(nt, 3)


def write_saving_vector(path: str, sv: np.ndarray) -> None:
    with open(path, "w") as fh:
        for v in sv:
            fh.write(f" {int(v):11d}\n")


def read_saving_vector(path: str, nt: int) -> np.ndarray:
    L = _Lines(path)
    return np.array([int(L.value()) for _ in range(nt)], np.int32)


# ---------------------------------------------------------------------------
# block grids and projected kernels (TOOLS/SOURCE/project_kernel.f90)
# ---------------------------------------------------------------------------

def write_block_file(path: str, edges: Sequence[np.ndarray]) -> None:
    """MODELS/MODELS_3D/block_{x,y,z}: the number of subvolumes, then per
    subvolume the number of block edges and one edge per line (degrees for
    x = colatitude and y = longitude, km of radius for z), as
    TOOLS/SOURCE/project_kernel.f90:306-308,350-372 reads them."""
    out = [f"{len(edges)}"]
    for e in edges:
        out.append(f"{len(e)}")
        out += [repr(float(v)) for v in e]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read_block_file(path: str) -> List[np.ndarray]:
    toks = [_tok(l) for l in open(path) if l.strip()]
    n, k, out = int(toks[0]), 1, []
    for _ in range(n):
        m = int(toks[k]); k += 1
        out.append(np.array([float(t) for t in toks[k:k + m]]))
        k += m
    return out


def read_projected(path: str) -> List[np.ndarray]:
    """gradient_<param> written by project_kernel.f90:735-750: the number of
    subvolumes, then per subvolume the block count and one value per block
    (x slowest, z fastest)."""
    toks = [_tok(l) for l in open(path) if l.strip()]
    n, k, out = int(toks[0]), 1, []
    for _ in range(n):
        m = int(toks[k]); k += 1
        out.append(f32_array(toks[k:k + m]))
        k += m
    return out


def write_projected(path: str, values: Sequence[np.ndarray]) -> None:
    out = [f"{len(values)}"]
    for v in values:
        v = np.asarray(v, np.float32).ravel()
        out.append(f"{v.size}")
        out += [_fortran_real(x) for x in v]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
