#!/usr/bin/env python3
"""Benchmark: GLL-point updates per second of the SES3D forward time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload anelastic_96x96x48]

A step is one iteration of the reference time loop (SOURCE/ses3d_main.f90:95-178)
over the whole section: element kernel, point forces, node kernel and the
seismogram sampling of 200 receivers.  Inputs are resident in HBM before the
timed region.  updates/s = N_elements * 125 * K / wall.

The JSON line also carries
  roofline      the element kernel (the dominant kernel): algorithmic bytes per
                launch / mean launch duration measured with HIP events on the
                kernel's stream over the timed region, against 8 TB/s HBM;
  cpu_baseline  the reference Fortran solver (oracle/_ref, compiled from the
                reference sources) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLL-pt updates/s fwd; s per event-gradient, batched vs plain, 1 MI355X"
HBM_PEAK = 8.0e12


def algorithmic_bytes(nel, is_diss: bool) -> dict:
    """Algorithmic HBM bytes per GLL-point update (float32).

    Whole step: state read+written once (9 displacement gradients, 9 split
    stresses, 18 memory variables if anelastic), coefficients read once
    (lambda, mu, A, B, C [, Q]), velocity read+written and the mass read once
    per global node (rho_g = global nodes / GLL points).  Element kernel: all
    of it except the velocity write and the mass read (node kernel)."""
    nx, ny, nz = nel
    n_p = nx * ny * nz * 125
    n_g = (4 * nx + 1) * (4 * ny + 1) * (4 * nz + 1)
    rho_g = n_g / n_p
    state = 18 + 18 + (36 if is_diss else 0)
    coef = 6 if is_diss else 5
    step = 4 * (state + coef + 7 * rho_g)
    elem = 4 * (state + coef + 3 * rho_g)
    return {"rho_g": rho_g, "step": step, "element": elem, "node": 4 * 4 * rho_g}


def implemented_bytes(nel, pml: int, is_diss: bool, aniso: bool = False) -> float:
    """Bytes per GLL point the implemented element kernel must move (forward
    mode).  Elements outside every absorbing layer keep 6 split stresses and no
    displacement gradient; elements in a layer keep 9 + 9.  Coefficients:
    lambda, mu (+Q) (+A, B, C only for anisotropic models); the velocity is
    gathered once per global node; the element force (3 floats) is written."""
    nx, ny, nz = nel
    n_p = nx * ny * nz * 125
    rho_g = (4 * nx + 1) * (4 * ny + 1) * (4 * nz + 1) / n_p
    inner = max(nx - 2 * pml, 0) * max(ny - 2 * pml, 0) * max(nz - pml, 0) / (nx * ny * nz)
    mem = 36 if is_diss else 0
    coef = 2 + (1 if is_diss else 0) + (3 if aniso else 0)
    common = 4 * (mem + coef + 3 * rho_g + 3)
    return inner * (common + 4 * 12) + (1 - inner) * (common + 4 * 36)


def device_copy_gbps(mib: int = 2048, reps: int = 10) -> float:
    """HBM bandwidth a plain device-to-device copy reaches on this GPU (bytes
    read + written per second, GB/s): the practical ceiling a streaming
    kernel is compared with, next to the 8 TB/s data-sheet peak."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * 4 * n / (ms * 1e-3) / 1e9


def host_cpu() -> dict:
    """The host's CPU model, logical CPUs and the CPUs this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline_reference(is_diss: bool, steps=(10, 50), timeout=300.0, decomp=(1, 1, 1)):
    """Time the reference binary on px*py*pz MPI ranks (decomp), each rank
    holding the largest box its static arrays allow (23 x 28 x 8 elements,
    SOURCE/ses3d_modules.f90:13-15), so the section is 23px x 28py x 8pz;
    per-step cost = difference of two runs of different length."""
    from oracle import pyoracle as PO
    from ses3d_amd import formats as F, models as M
    if not PO.ref_available():
        return None
    nprocs = int(np.prod(decomp))
    nel = (23 * decomp[0], 28 * decomp[1], 8 * decomp[2])
    times = []
    for nt in steps:
        root = tempfile.mkdtemp(prefix="ses3d_cpu_")
        try:
            setup = M.section_setup(*nel, is_diss=is_diss)
            box = F.make_boxes(setup)[0]
            pert = M.gaussian_random_field(box, 1, corr_len=200e3, rms=0.03)
            model = M.make_model(box, is_diss=is_diss, perturbation=pert)
            rng = np.random.default_rng(0)
            ev = F.Event(nt=nt, dt=0.1, xxs=50.0, yys=20.0, zzs=10000.0, source_type=10,
                         moment=tuple(M.random_moment_tensor(rng)), ofd="../DATA/OUTPUT/1/", ssamp=10 ** 6)
            rec = M.surface_receiver_grid(setup, 10, 10)
            if nprocs > 1:
                setup = M.decompose(setup, *decomp)
            M.write_run_tree(root, M.RunCase(setup, [ev], [rec], M.ricker(nt + 4, 0.1, 20.0), model))
            t0 = time.perf_counter()
            r = PO.run_ref(root, timeout=timeout, nprocs=nprocs)
            t1 = time.perf_counter()
            if r.returncode != 0:
                return None
            times.append(t1 - t0)
        finally:
            shutil.rmtree(root, ignore_errors=True)
    n_p = int(np.prod(nel)) * 125
    per_step = (times[1] - times[0]) / (steps[1] - steps[0])
    ranks = f"{nprocs} MPI ranks ({decomp[0]}x{decomp[1]}x{decomp[2]} boxes of 23x28x8)" if nprocs > 1 else \
        "1 MPI rank"
    return {"value": n_p / per_step, "unit": "GLL-pt updates/s", "cores": nprocs, "kind": "reference",
            "sample": (f"reference Fortran (amdflang -O3, MPICH) on {ranks}: {nel[0]}x{nel[1]}x{nel[2]} "
                       f"{'anelastic' if is_diss else 'elastic'} elements, per-step time = difference of "
                       f"{steps[1]}- and {steps[0]}-step runs ({times[1]:.1f} s and {times[0]:.1f} s)"),
            **host_cpu()}


def cpu_baseline_workload(w, decomp=(4, 4, 1), steps=(2, 12), timeout=900.0):
    """Time the reference solver on the benchmark section itself: the
    workload's run tree (same mesh, model, source, receivers, absorbing
    width) split into px*py*pz rank boxes and run on as many MPI ranks by the
    reference build sized for those boxes (oracle/_ref/ses3d_baseline_<w>.exe,
    static extents = the boxes' element counts, pml as the workload's).
    Per-step cost = difference of two runs of different length."""
    from dataclasses import replace
    from oracle import pyoracle as PO
    from ses3d_amd import configs, formats as F, models as M
    exe = os.path.join(ROOT, "oracle", "_ref", f"ses3d_baseline_{w.name.split('_', 1)[1]}.exe")
    if not os.path.exists(exe) or any(n % d for n, d in zip(w.nel, decomp)):
        return None
    nprocs = int(np.prod(decomp))
    run = configs.build(w, seed=0, nt=max(steps))
    e = run.events[0]
    root = tempfile.mkdtemp(prefix="ses3d_cpu_w_")
    times = []
    try:
        setup = M.decompose(run.setup, *decomp)
        ev = replace(e.event, nt=max(steps), ssamp=10 ** 6, output_displacement=0)
        M.write_run_tree(root, M.RunCase(setup, [ev], [e.receivers], M.ricker(max(steps) + 4, w.dt, w.period),
                                         run.model))
        del run
        for nt in steps:
            replace(ev, nt=nt).write(os.path.join(root, "INPUT", "event_1"))
            t0 = time.perf_counter()
            r = PO.run_ref(root, timeout=timeout, nprocs=nprocs, exe=exe)
            t1 = time.perf_counter()
            if r.returncode != 0:
                return None
            times.append(t1 - t0)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    per_step = (times[1] - times[0]) / (steps[1] - steps[0])
    bx = [n // d for n, d in zip(w.nel, decomp)]
    return {"value": w.n_points / per_step, "unit": "GLL-pt updates/s", "cores": nprocs, "kind": "reference",
            "sample": (f"reference Fortran (amdflang -O3, MPICH) on {nprocs} MPI ranks: the {w.name} section "
                       f"({w.nel[0]}x{w.nel[1]}x{w.nel[2]} elements, pml {w.pml}) as {decomp[0]}x{decomp[1]}x"
                       f"{decomp[2]} rank boxes of {bx[0]}x{bx[1]}x{bx[2]} elements (a build sized for them), "
                       f"per-step time = difference of {steps[1]}- and {steps[0]}-step runs "
                       f"({times[1]:.1f} s and {times[0]:.1f} s)"),
            **host_cpu()}


def source_sha256() -> str:
    """sha256 of the HIP source the library is built from (stamps measured
    counter files with the code they describe)."""
    import hashlib
    with open(os.path.join(ROOT, "ses3d_amd", "csrc", "ses3d_hip.hip"), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def traffic_file(workload: str) -> str:
    return os.path.join(ROOT, "profiles", f"k_element_{workload}_traffic.json")


def _pmc_mean_kib(csv_path: str, kernel: str) -> float:
    import csv
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(csv_path)) if kernel in r["Kernel_Name"]]
    if not vals:
        raise RuntimeError(f"no {kernel} dispatches in {csv_path}")
    return sum(vals) / len(vals)


def measure_traffic(workload: str, steps: int = 8) -> dict:
    """HBM bytes per element-kernel launch from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over a short run of the
    workload, each scaled by what the same counter reads for a plain device
    copy of known size; written with the HIP source's sha256 to the
    workload's traffic file, which the benchmark reports only while that
    source is unchanged."""
    import subprocess
    out = tempfile.mkdtemp(prefix="ses3d_pmc_")
    py = sys.executable
    runs = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for tag, prog in (("kernel", [py, os.path.abspath(__file__), "--traffic-probe", workload,
                                      "--steps", str(steps)]),
                          ("copy", [py, os.path.abspath(__file__), "--calibrate-copy", "1024"])):
            d = os.path.join(out, f"{tag}_{counter}")
            cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--", *prog]
            print("# " + " ".join(cmd), file=sys.stderr, flush=True)
            subprocess.run(cmd, check=True, timeout=300, cwd=out, stdout=subprocess.DEVNULL)
            runs[(tag, counter)] = os.path.join(d, "run_counter_collection.csv")
    copy_bytes = 1024 << 20
    fetch_scale = copy_bytes / 1024 / _pmc_mean_kib(runs[("copy", "FETCH_SIZE")], "copyBuffer")
    write_scale = copy_bytes / 1024 / _pmc_mean_kib(runs[("copy", "WRITE_SIZE")], "copyBuffer")
    fetch = _pmc_mean_kib(runs[("kernel", "FETCH_SIZE")], "k_element")
    write = _pmc_mean_kib(runs[("kernel", "WRITE_SIZE")], "k_element")
    nfetch = _pmc_mean_kib(runs[("kernel", "FETCH_SIZE")], "k_node")
    nwrite = _pmc_mean_kib(runs[("kernel", "WRITE_SIZE")], "k_node")
    res = {"kernel": "k_element", "workload": workload,
           "hbm_bytes_per_launch": 1024 * (fetch * fetch_scale + write * write_scale),
           "fetch_kib": fetch, "write_kib": write, "fetch_scale": fetch_scale, "write_scale": write_scale,
           "node_kernel_bytes_per_launch": 1024 * (nfetch * fetch_scale + nwrite * write_scale),
           "source_sha256": source_sha256(),
           "command": f"python bench.py --measure-traffic {workload} (rocprofv3 --pmc FETCH_SIZE, then "
                      f"WRITE_SIZE, on bench.py --traffic-probe {workload} --steps {steps} and on a "
                      f"1 GiB device copy)"}
    with open(traffic_file(workload), "w") as fh:
        json.dump(res, fh, indent=1)
    shutil.rmtree(out, ignore_errors=True)
    return res


def traffic_probe(workload: str, steps: int) -> None:
    """A short forward run of the workload for counter passes."""
    from ses3d_amd import configs
    from ses3d_amd.solver import Solver
    w = configs.WORKLOADS[workload]
    run = configs.build(w, seed=0, nt=steps)
    s = Solver.from_run(run, adjoint_flag=0, pml=w.pml, nt=steps)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    s.run()
    s.synchronize()
    s.close()


def configs_nt(name: str) -> int:
    from ses3d_amd import configs
    return configs.WORKLOADS[name].nt


def gradient_leg(k: int = 100, nt=None, n_events: int = 1, history_degree: int = 4, kernel_bits: int = 64,
                 keep: int = 1, workload: str = "gradient_48x48x24") -> dict:
    """Seconds per event-gradient on the differentiation workload (48x48x24
    anelastic, dt=0.15): data from the 'true' model (1D + 2% Gaussian blobs),
    gradient of the L2 waveform misfit at the 1D model by the checkpointed
    forward (restart states every k steps in HBM) and the adjoint with kernel
    accumulation, for n_events events solved together in one context (event 1:
    100 surface receivers; the others seeded sources 5-600 km deep with 10-400
    receivers).  The timed region is forward + adjoint of all events.
    history_degree < 4 stores the sampled forward fields at that polynomial
    degree (ses3d_store's `deg`; lossy, smaller segment buffer); kernel_bits=32
    accumulates the kernels in float32 as the reference does (its kernels bit
    for bit) instead of float64; keep > 1 keeps the sampled steps of the last
    `keep` segments of the checkpointed forward in HBM, so the adjoint replays
    (keep-1)*k fewer forward steps."""
    from dataclasses import replace
    from ses3d_amd import configs, misfit
    from ses3d_amd.solver import Solver
    w = configs.WORKLOADS[workload]
    nt = nt or w.nt
    true = configs.build(w, seed=0, nt=nt, n_events=n_events)
    cur = configs.build(replace(w, perturbation="none"), seed=0, nt=nt, n_events=n_events)
    f0 = Solver.from_run(true, adjoint_flag=0, pml=w.pml, nt=nt, n_events=n_events)
    f0.set_model(true.model)
    for j in range(n_events):
        f0.load_event(j, true.events[j])
    f0.init()
    f0.run()
    data = [f0.seismograms(j) for j in range(n_events)]
    f0.close()
    del true
    f = Solver.from_run(cur, adjoint_flag=1, pml=w.pml, nt=nt, n_events=n_events)
    f.set_history_degree(history_degree)
    f.set_model(cur.model)
    for j in range(n_events):
        f.load_event(j, cur.events[j])
    f.init()
    a = Solver.from_run(cur, adjoint_flag=2, pml=w.pml, nt=nt, n_events=n_events)
    a.set_history_degree(history_degree)
    a.set_kernel_precision(kernel_bits)
    a.set_model(cur.model)
    dt = float(cur.events[0].dt)
    taper = misfit.cosine_taper(nt, dt, 40.0)
    # checkpoint storage is set-up (allocated and first touched before the timed region)
    f.prepare_checkpoints(k, keep=keep)
    t0 = time.perf_counter()
    f.forward_checkpointed(k, keep=keep)
    t_fwd = time.perf_counter()
    chi = 0.0
    for j in range(n_events):
        u = f.seismograms(j)
        chi += misfit.l2_misfit(u, data[j], dt, taper)
        rows = misfit.l2_adjoint_rows(u, data[j], dt, taper=taper)
        a.set_adjoint_sources(j, cur.events[j].rec32[f.receiver_index(j)], rows)
    a.init()
    a.adjoint_checkpointed(f)
    kern = [a.kernels(j) for j in range(n_events)]
    t1 = time.perf_counter()
    ok = all(np.isfinite(v).all() and np.abs(v).max() > 0 for kk in kern for v in kk.values())
    out = {"s_per_event_gradient": (t1 - t0) / n_events, "events": n_events,
           "s_forward_checkpointed": t_fwd - t0, "s_adjoint": t1 - t_fwd,
           "workload": f"gradient {w.name} anelastic nt={nt} dt={w.dt}",
           "checkpoint_every": k, "receivers": [int(f.nr(j)) for j in range(n_events)],
           "gll_points": w.n_points, "misfit": chi, "kernels_finite_nonzero": bool(ok),
           "device_GiB": (f.memory_bytes() + a.memory_bytes()) / 2 ** 30}
    if history_degree < 4:
        out["history_degree"] = history_degree
    out["kernel_accumulators"] = f"float{kernel_bits}"
    out["kept_segments"] = keep
    f.close()
    a.close()
    return out


def small_forward_leg(name: str = "elastic_24x24x12", warmup: int = 20) -> dict:
    """The reference's CPU-runnable case on the GPU: the whole time loop of the
    elastic 24x24x12 workload (864000 GLL points, 20 receivers, nt=2000), the
    first `warmup` steps autotuning the launch schedule, the rest timed.  At
    ~150 MB of state per step it lives in the GPU's caches, so it measures
    launch and latency costs rather than HBM bandwidth."""
    from ses3d_amd import configs
    from ses3d_amd.solver import Solver
    w = configs.WORKLOADS[name]
    run = configs.build(w, seed=0)
    nt = run.events[0].nt
    s = Solver.from_run(run, adjoint_flag=0, pml=w.pml)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    s.autotune(1, warmup)
    steps = nt - warmup
    wall = _timed_run(s, warmup + 1, steps)
    ok = bool(np.isfinite(s.seismograms()).all())
    s.close()
    return {"workload": f"plain forward {name} (elastic, {w.n_points} GLL points, nt={nt})",
            "updates_per_s": w.n_points * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "timed_steps": steps, "seismograms_finite": ok}


def _timed_run(x, it0: int, steps: int) -> float:
    x.synchronize()
    t0 = time.perf_counter()
    x.run(it0, steps, sync=False)
    x.synchronize()
    return time.perf_counter() - t0


def batched_leg(n_events: int = 32, steps: int = 30) -> dict:
    """Uneven batch: n_events events on the 64x64x32 anelastic mesh plus six
    extra anelastic sections (16^3 ... 56x56x28, one event each) advanced as one
    group, one element-kernel and one node-kernel launch per step (autotuned
    schedule, captured step graphs).  Compared with one event alone on the
    64x64x32 mesh (per-event throughput) and with the same sections stepped one
    context after another."""
    from ses3d_amd import configs
    from ses3d_amd.solver import Group, Solver
    w = configs.WORKLOADS["batched_64x64x32"]
    tune, warm = 12, 16                    # autotune steps, then steps that capture the graph
    nt = tune + warm + steps
    runs = [(configs.build(w, seed=0, nt=nt, n_events=n_events), w)]
    for k, nel in enumerate(configs.EXTRA_SECTIONS):
        ws = configs.section_workload(nel, nt=nt, seed=k)
        runs.append((configs.build(ws, seed=10 + k, nt=nt), ws))

    def make(run, ww, n):
        s = Solver.from_run(run, adjoint_flag=0, pml=ww.pml, nt=nt, n_events=n)
        s.set_model(run.model)
        for k in range(n):
            s.load_event(k, run.events[k])
        s.init()
        return s

    # one event alone on the main mesh
    s = make(runs[0][0], w, 1)
    s.autotune(1, tune)
    s.run(tune + 1, warm)
    plain = w.n_points * steps / _timed_run(s, tune + warm + 1, steps)
    s.close()
    sols = [make(runs[0][0], w, n_events)] + [make(r, ww, 1) for r, ww in runs[1:]]
    pts = w.n_points * n_events + sum(ww.n_points for _, ww in runs[1:])
    # the same sections one context after another
    seq = 0.0
    for x in sols:
        x.autotune(1, tune)
        x.run(tune + 1, warm)
        seq += _timed_run(x, tune + warm + 1, steps)
    for x in sols:
        x.init()
    g = Group(sols)
    sched = g.autotune(1, tune)
    g.run(tune + 1, warm)
    grouped = pts * steps / _timed_run(g, tune + warm + 1, steps)
    ok = all(np.isfinite(x.seismograms(k)).all() for x in sols for k in range(x.n_events))
    g.close()
    for x in sols:
        x.close()
    return {"workload": (f"forward, {n_events} events on {w.name} (seeded sources 5-600 km deep, 10-400 "
                         f"receivers each) + sections " + ", ".join("x".join(map(str, n)) for n in
                                                                    configs.EXTRA_SECTIONS) + ", anelastic"),
            "gll_points_per_step": pts, "steps": steps,
            "plain_updates_per_s": plain, "batched_updates_per_s": grouped,
            "batched_per_event_over_plain": grouped / plain,
            "sequential_updates_per_s": pts * steps / seq, "group_over_sequential": grouped / (pts * steps / seq),
            "schedule": {k: v for k, v in sched.items() if k != "trial_ms"}, "seismograms_finite": bool(ok)}


def size_sweep(sizes=((16, 16, 8), (24, 24, 12), (32, 32, 16), (48, 48, 24), (64, 64, 32), (80, 80, 40),
                      (96, 96, 48), (112, 112, 56), (128, 128, 64))) -> list:
    """Plain forward step against section size, anelastic (3D random model,
    pml=3) and elastic (1D, pml=2), one event with 20 receivers: 18 autotuning
    steps, 30 steps timed kernel by kernel with HIP events, 16 steps that
    capture the step graph, then 80 graph-replayed steps timed as a whole."""
    from ses3d_amd import configs
    from ses3d_amd.solver import Solver
    out = []
    for diss in (True, False):
        for nel in sizes:
            w = configs.Workload(f"sweep_{nel[0]}x{nel[1]}x{nel[2]}", tuple(nel), diss, 3 if diss else 2, 0.1, 150,
                                 20, "grf" if diss else "none")
            run = configs.build(w, seed=0)
            s = Solver.from_run(run, adjoint_flag=0, pml=w.pml)
            s.set_model(run.model)
            s.load_event(0, run.events[0])
            s.init()
            del run
            sch = s.autotune(1, 18)
            s.set_profiling(True)
            s.run(19, 30)
            tm = s.timing()
            s.set_profiling(False)
            s.run(49, 16)
            wall = _timed_run(s, 65, 80)
            ok = bool(np.isfinite(s.seismograms()).all())
            s.close()
            ab = algorithmic_bytes(nel, diss)
            n_p = w.n_points
            e = tm["element_ms"] / tm["n_element"]
            r = {"elements": list(nel), "anelastic": diss, "gll_points": n_p, "ms_per_step": 1e3 * wall / 80,
                 "updates_per_s": n_p * 80 / wall, "element_ms": e, "node_ms": tm["node_ms"] / tm["n_element"],
                 "element_frac": ab["element"] * n_p / (e * 1e-3) / HBM_PEAK,
                 "step_frac": ab["step"] * n_p * 80 / wall / HBM_PEAK,
                 "schedule": {k: sch[k] for k in ("epb", "node_threads", "remap_elem", "remap_node")},
                 "seismograms_finite": ok}
            print(json.dumps(r), flush=True)
            out.append(r)
    return out


def inversion_leg(n_events: int = 16, nt: int = 4000, k: int = 100, step_frac: float = 0.005) -> dict:
    """One inversion iteration on the 64x64x32 anelastic mesh: data from the
    'true' model (1D + 2% Gaussian blobs of 300 km) for n_events events, the
    summed L2 misfit gradient at the 1D model through the torch.autograd.Function
    (checkpointed forward every k steps + adjoint with float64 kernel
    accumulation, one event per solve: the checkpoints of one event take ~160 GB),
    a steepest-descent update of vp, vsh, vsv (step_frac of their means along the
    max-normalised gradient), and the misfit before and after it, measured by plain
    single-event forward solves.  The same gradient is also timed without
    autograd (Problem.misfit_and_gradient: each batch's checkpointed forward and
    adjoint back to back), where no forward solve is repeated; autograd's
    forward keeps the checkpoints of the last batch only, so its backward
    recomputes the checkpointed forward of every other batch."""
    import torch
    from dataclasses import replace
    from ses3d_amd import configs
    from ses3d_amd.autograd import Problem, misfit
    from ses3d_amd.solver import Solver
    w = configs.Workload("inversion_64x64x32", (64, 64, 32), True, 2, 0.1, nt, 200, "blobs", n_events=n_events)
    true = configs.build(w, seed=0)
    cur = configs.build(replace(w, perturbation="none"), seed=0)
    t0 = time.perf_counter()
    f0 = Solver.from_run(true, adjoint_flag=0, pml=w.pml, n_events=n_events)
    f0.set_model(true.model)
    for j in range(n_events):
        f0.load_event(j, true.events[j])
    f0.init()
    f0.run()
    data = [f0.seismograms(j) for j in range(n_events)]
    f0.close()
    t_data = time.perf_counter() - t0
    print(f"# data of {n_events} events in {t_data:.1f} s", file=sys.stderr, flush=True)
    del true
    m = cur.model
    rho0 = 1.0 / m["rhoinv"].astype(np.float64)
    vs0 = np.sqrt(m["mu"] / rho0)
    vp0 = np.sqrt((m["lambda"] + 2 * m["mu"]) / rho0)
    prob = Problem(cur, data, pml=w.pml, ckpt=k, events_per_batch=1, taper_width=40.0,
                   progress=lambda msg: print(f"# {msg}", file=sys.stderr, flush=True))
    x = [torch.tensor(a, requires_grad=True) for a in (rho0, vp0, vs0, vs0.copy())]
    t1 = time.perf_counter()
    chi0 = misfit(prob, *x)
    chi0.backward()
    t_grad = time.perf_counter() - t1
    g = [t.grad.numpy() for t in x]
    print(f"# autograd gradient of {n_events} events in {t_grad:.1f} s", file=sys.stderr, flush=True)
    # the same gradient without autograd: checkpointed forward and adjoint of each
    # batch back to back (no recomputed forward for the batches whose checkpoints
    # autograd's forward could not keep)
    t1 = time.perf_counter()
    model0 = prob.fields(rho0, vp0, vs0, vs0)
    chi_e, K = prob.misfit_and_gradient(model0)
    ge = prob.gradient_from_kernels(K, model0)
    t_eager = time.perf_counter() - t1
    same = chi_e == float(chi0.detach()) and all(np.array_equal(ge[n], g[i])
                                                 for i, n in enumerate(("rho", "vp", "vsh", "vsv")))
    scale = [0.0, step_frac * vp0.mean(), step_frac * vs0.mean(), step_frac * vs0.mean()]
    new = [x[i].detach().numpy() - scale[i] * g[i] / max(np.abs(g[i]).max(), 1e-300) for i in range(4)]
    t2 = time.perf_counter()
    chi_before = prob.misfit(prob.fields(rho0, vp0, vs0, vs0))
    chi_after = prob.misfit(prob.fields(*new))
    t_judge = time.perf_counter() - t2
    return {"workload": f"inversion iteration, {w.name}, {n_events} events, nt={nt}, dt={w.dt}, checkpoints every {k}",
            "events": n_events, "gll_points": w.n_points,
            "s_per_event_gradient": t_grad / n_events, "s_gradient_total": t_grad, "s_data": t_data,
            "s_per_event_gradient_direct": t_eager / n_events, "direct_equals_autograd": bool(same),
            "s_judge_two_misfits": t_judge, "misfit_autograd": float(chi0.detach()),
            "misfit_before": chi_before, "misfit_after": chi_after, "misfit_ratio": chi_after / chi_before}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="anelastic_96x96x48")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-elastic", action="store_true", help="skip the elastic 24x24x12 leg")
    ap.add_argument("--checkpoint-every", type=int, default=100,
                    help="restart states every k steps in the gradient legs")
    ap.add_argument("--checkpoint-keep", type=int, default=1,
                    help="trailing checkpoint segments whose sampled steps stay in HBM (no replay) in the "
                         "plain and float32 gradient legs")
    ap.add_argument("--history-degree", type=int, default=4,
                    help="polynomial degree of the stored forward fields in the gradient legs (4: unchanged)")
    ap.add_argument("--gradient-events", type=int, default=4,
                    help="events solved together in the batched gradient leg (about 57 GB of HBM each)")
    ap.add_argument("--inversion", action="store_true",
                    help="only run one inversion iteration on the 64x64x32 mesh (16 events, ~10 min)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a device copy of this many MiB (for HBM counter calibration)")
    ap.add_argument("--measure-traffic", metavar="WORKLOAD",
                    help="only measure the element kernel's HBM bytes per launch with rocprofv3 counter "
                         "passes and write them, stamped with the HIP source's sha256, under profiles/")
    ap.add_argument("--traffic-probe", metavar="WORKLOAD", help=argparse.SUPPRESS)
    ap.add_argument("--gradient-probe", type=int, metavar="NT",
                    help="only one event-gradient of the 48x48x24 gradient workload at nt=NT with checkpoints "
                         "every --checkpoint-every steps (short runs for counter passes)")
    ap.add_argument("--small-gradients", action="store_true",
                    help="only the small-mesh gradient legs: one event alone and 32 events in one context, "
                         "on 8x8x4 and 16x16x8 elements")
    ap.add_argument("--nt", type=int, default=None, help="time steps of the --small-gradients legs")
    ap.add_argument("--size-sweep", action="store_true",
                    help="plain forward step time and roofline fraction against section size, then exit")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps per captured step graph in every context (0: direct launches)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole section is resident on a single MI355X and this "
              "build runs no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.measure_traffic:
        print(json.dumps(measure_traffic(args.measure_traffic)))
        return
    if args.graph_steps is not None:
        from ses3d_amd.solver import Solver
        Solver.default_graph_steps = args.graph_steps
    if args.traffic_probe:
        traffic_probe(args.traffic_probe, args.steps)
        return
    import torch
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": 4 * n, "launches": 5}))
        return

    if args.size_sweep:
        size_sweep()
        return
    if args.inversion:
        print(json.dumps(inversion_leg()))
        return
    if args.gradient_probe:
        print(json.dumps(gradient_leg(k=args.checkpoint_every, nt=args.gradient_probe)))
        return
    if args.small_gradients:
        res = {}
        for name in ("gradient_8x8x4", "gradient_16x16x8"):
            print(f"# {name}: one event, then 32 events in one context", file=sys.stderr, flush=True)
            sp = gradient_leg(k=args.checkpoint_every, workload=name, nt=args.nt)
            sb = gradient_leg(k=args.checkpoint_every, n_events=32, workload=name, nt=args.nt)
            sb["per_event_speedup_over_plain"] = sp["s_per_event_gradient"] / sb["s_per_event_gradient"]
            res[name] = {"plain": sp, "batched": sb}
        print(json.dumps(res))
        return

    from ses3d_amd import configs
    from ses3d_amd.solver import Solver

    w = configs.WORKLOADS[args.workload]
    nt = args.warmup + args.steps
    t0 = time.perf_counter()
    run = configs.build(w, seed=0, nt=nt)
    print(f"# built {w.name} ({w.n_points} GLL points) in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    s = Solver.from_run(run, adjoint_flag=0, pml=w.pml, nt=nt)
    s.set_model(run.model)
    s.load_event(0, run.events[0])
    s.init()
    del run
    print(f"# initialised; device memory {s.memory_bytes() / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    sched = None
    if args.warmup >= 6:
        sched = s.autotune(1, args.warmup)    # bit-identical steps under trial launch schedules
    elif args.warmup:
        s.run(1, args.warmup)
    s.set_profiling(True)
    torch.cuda.synchronize()
    s.synchronize()
    t1 = time.perf_counter()
    s.run(args.warmup + 1, args.steps, sync=False)
    s.synchronize()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    tm = s.timing()
    seis = s.seismograms()
    if not np.isfinite(seis).all():
        raise RuntimeError("non-finite seismograms")
    wall = t2 - t1
    n_p = w.n_points
    value = n_p * args.steps / wall
    ab = algorithmic_bytes(w.nel, w.is_diss)
    elem_ms = tm["element_ms"] / max(tm["n_element"], 1)
    achieved = ab["element"] * n_p / (elem_ms * 1e-3)
    # counter-measured HBM bytes per element-kernel launch, reported only if
    # measured on the HIP source this library is built from
    traffic, traffic_note = None, "no counter file"
    tfile = traffic_file(w.name)
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        if tj.get("source_sha256") == source_sha256():
            traffic, traffic_note = tj.get("hbm_bytes_per_launch"), "rocprofv3 PMC of this HIP source: " + tj["command"]
        else:
            traffic_note = "counter file measured on another HIP source: not reported"
    out = {
        "metric": METRIC, "value": value, "unit": "GLL-pt updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain forward {w.name} (lpd=4, {n_p} GLL points, "
                               f"{'anelastic 3 mechanisms' if w.is_diss else 'elastic'}, pml={w.pml}, "
                               f"dt={w.dt}, 1 event, {s.nr(0)} receivers)",
                   "elements": list(w.nel), "gll_points": n_p, "timed_steps_from": args.warmup + 1},
        "roofline": {"bound": "hbm", "kernel": "k_element", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "algorithmic_bytes_per_point": ab["element"], "points_per_launch": n_p,
                     "mean_launch_ms": elem_ms,
                     "impl_bytes_per_point": implemented_bytes(w.nel, w.pml, w.is_diss),
                     "frac_impl": implemented_bytes(w.nel, w.pml, w.is_diss) * n_p / (elem_ms * 1e-3) / HBM_PEAK,
                     "traffic_frac": (traffic / (elem_ms * 1e-3) / HBM_PEAK) if traffic else None,
                     "traffic_source": traffic_note},
        "step_roofline_frac": value * ab["step"] / HBM_PEAK,
        "kernel_ms_per_step": {"element": elem_ms, "node": tm["node_ms"] / max(tm["n_element"], 1)},
        "schedule": sched,
    }
    s.close()
    copy = device_copy_gbps()
    roof = out["roofline"]
    roof["device_copy_GBps"] = copy
    if traffic:
        roof["traffic_GBps"] = traffic / (elem_ms * 1e-3) / 1e9
        roof["traffic_frac_of_copy"] = roof["traffic_GBps"] / copy
    if not args.no_elastic:
        print("# elastic 24x24x12 leg", file=sys.stderr, flush=True)
        out["elastic_24x24x12"] = small_forward_leg()
    if not args.no_gradient:
        print("# gradient leg", file=sys.stderr, flush=True)
        out["gradient"] = gradient_leg(k=args.checkpoint_every, history_degree=args.history_degree,
                                       keep=args.checkpoint_keep)
        print("# batched gradient leg", file=sys.stderr, flush=True)
        gb = gradient_leg(k=args.checkpoint_every, n_events=args.gradient_events, history_degree=args.history_degree)
        gb["per_event_speedup_over_plain"] = out["gradient"]["s_per_event_gradient"] / gb["s_per_event_gradient"]
        out["gradient_batched"] = gb
        print("# gradient leg, float32 kernel accumulators", file=sys.stderr, flush=True)
        g32 = gradient_leg(k=args.checkpoint_every, history_degree=args.history_degree, kernel_bits=32,
                           keep=args.checkpoint_keep)
        g32["speedup_over_float64"] = out["gradient"]["s_per_event_gradient"] / g32["s_per_event_gradient"]
        out["gradient_f32"] = g32
        # 8 of the 30 segments kept: a 199 GB sampled-step buffer, 700 fewer replayed steps
        print("# gradient leg, float32 accumulators, 8 kept segments", file=sys.stderr, flush=True)
        g8 = gradient_leg(k=args.checkpoint_every, history_degree=args.history_degree, kernel_bits=32, keep=8)
        g8["speedup_over_float64"] = out["gradient"]["s_per_event_gradient"] / g8["s_per_event_gradient"]
        out["gradient_f32_keep8"] = g8
        # stored fields at degree 2 (27 of 125 values per element, ses3d_store's deg=2): every
        # segment's sampled steps fit in HBM (a 161 GB buffer), so nothing is replayed
        print("# gradient leg, float32 accumulators, degree-2 stored fields, all segments kept",
              file=sys.stderr, flush=True)
        nseg = -(-configs_nt("gradient_48x48x24") // args.checkpoint_every)
        gd = gradient_leg(k=args.checkpoint_every, history_degree=2, kernel_bits=32, keep=nseg)
        gd["speedup_over_float64"] = out["gradient"]["s_per_event_gradient"] / gd["s_per_event_gradient"]
        out["gradient_f32_deg2_keepall"] = gd
        # where one event's step is launch-bound, batching pays: 32 and 64 events on 8x8x4
        # in one context against one event alone
        print("# small-mesh gradient legs, plain and 32 / 64 events batched", file=sys.stderr, flush=True)
        sp = gradient_leg(k=args.checkpoint_every, workload="gradient_8x8x4")
        out["gradient_small_plain"] = sp
        for ne, key in ((32, "gradient_small_batched"), (64, "gradient_small_batched64")):
            sb = gradient_leg(k=args.checkpoint_every, n_events=ne, workload="gradient_8x8x4")
            sb["per_event_speedup_over_plain"] = sp["s_per_event_gradient"] / sb["s_per_event_gradient"]
            out[key] = sb
    if not args.no_batched:
        print("# batched leg", file=sys.stderr, flush=True)
        out["batched"] = batched_leg()
    if not args.no_cpu_baseline:
        # the reference on 16 MPI ranks (4x4x1 boxes of its one-rank maximum), this
        # box's CPU share, and on one rank beside it
        print("# timing the reference CPU solver (16 ranks, then 1)", file=sys.stderr, flush=True)
        cb = cpu_baseline_workload(w)
        if cb is None:
            cb = cpu_baseline_reference(w.is_diss, decomp=(4, 4, 1))
        one = cpu_baseline_reference(w.is_diss)
        if cb is None:
            cb = one
        elif one is not None:
            cb["one_rank"] = {k: one[k] for k in ("value", "cores", "sample")}
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
