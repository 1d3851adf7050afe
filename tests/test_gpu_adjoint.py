"""Adjoint path on the GPU: Frechet kernels (SOURCE/ses3d_grad.f90,
ses3d_grad_Q.f90) from the stored-history path and from the HBM-checkpoint
path.  The checkpoint path replays each forward segment from a restart state,
which reproduces the stored forward fields bit for bit, so its kernels must
equal the stored-history kernels exactly.  Against the float32 oracle the bar is
relative L2 <= 1e-5 (float64 accumulation on the GPU vs float32 in the
reference), and against the float64 oracle <= 1e-4."""
import numpy as np
import pytest

from tests.cases import make_case, oracle_for

pytestmark = pytest.mark.gpu
KNAMES = ("grad_rho", "grad_cp", "grad_csh", "grad_csv", "grad_Q_mu")


def _ctx(run, af, nt=None):
    from ses3d_amd.solver import Solver
    s = Solver.from_run(run, adjoint_flag=af)
    s.set_model(run.model)
    return s


def _adjoint_sources(e, seis):
    """L2-type adjoint sources at the located receivers, rows time-reversed as
    TOOLS/adjoint_source.py:89 writes them."""
    loc = e.rec32
    adsrc = (seis.astype(np.float64) * 1e6).astype(np.float32)         # (3, nr, nt) forward time
    rows = np.ascontiguousarray(adsrc[:, :, ::-1].transpose(1, 2, 0))  # (nr, nt, 3) reversed
    return loc, rows


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("is_diss,k", [(1, 7), (0, 16)])
def test_checkpointed_gradient(tmp_path, is_diss, k):
    run = make_case(str(tmp_path), nel=(6, 6, 4), nt=60, is_diss=bool(is_diss), adjoint_flag=1, samp_ad=4,
                    seed=13 + is_diss)
    e = run.events[0]
    nt = e.nt
    print(f"case diss={is_diss} k={k}", flush=True)
    # stored-history path
    f = _ctx(run, 1)
    f.load_event(0, e)
    f.init()
    f.run()
    seis = f.seismograms()
    loc, rows = _adjoint_sources(e, seis)
    a = _ctx(run, 2)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    for st in np.flatnonzero(f.saving_vector()) + 1:
        a.put_history(int(st), f.history(int(st)))
    a.run()
    K1 = a.kernels()
    # checkpoint path
    f2 = _ctx(run, 1)
    f2.load_event(0, e)
    f2.init()
    f2.forward_checkpointed(k)
    np.testing.assert_array_equal(f2.seismograms(), seis)
    a2 = _ctx(run, 2)
    a2.set_adjoint_sources(0, loc, rows)
    a2.init()
    a2.adjoint_checkpointed(f2)
    K2 = a2.kernels()
    for g in K1:
        np.testing.assert_array_equal(K2[g], K1[g], err_msg=g)
    # oracle
    of = oracle_for(run, e, adjoint_flag=1)
    of.run(1, nt)
    np.testing.assert_array_equal(of.get("seismograms").astype(np.float32), seis)
    of64 = oracle_for(run, e, adjoint_flag=1, precision=64)
    of64.run(1, nt)
    for prec, ofw, tol in ((32, of, 1e-5), (64, of64, 1e-4)):
        from oracle import pyoracle as PO
        oa = PO.Oracle(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin,
                       ymax=run.ymax, zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global, nt=nt,
                       dt=e.dt, is_diss=run.is_diss, adjoint_flag=2, samp_ad=run.samp_ad, pml=2,
                       tau_p=run.tau_p, D_p=run.D_p, precision=prec)
        oa.set_model(run.model)
        oa.set_adjoint_sources(loc, rows)
        oa.init()
        oa.run(1, nt, fwd=ofw)
        for g in K1:
            ref = oa.get(g)
            assert np.linalg.norm(ref) > 0, g
            rel = _rel(K1[g], ref)
            print(f"  {g} vs oracle{prec}: rel L2 {rel:.2e}", flush=True)
            assert rel <= tol, (g, prec, rel)


@pytest.mark.parametrize("keep", [2, 3, 9, 50])
def test_checkpointed_gradient_kept_segments(tmp_path, keep):
    """forward_checkpointed(k, keep): the sampled steps of the last `keep`
    segments stay in HBM and the adjoint replays only the earlier segments
    (nt=60, k=7: 9 segments; keep=9 and 50 keep them all, no replay).  The
    kernels equal the stored-history kernels bit for bit, and so do those of a
    second adjoint on the same forward, which replays every segment."""
    run = make_case(str(tmp_path), nel=(6, 6, 4), nt=60, is_diss=True, adjoint_flag=1, samp_ad=4, seed=14)
    e = run.events[0]
    print(f"kept segments {keep}", flush=True)
    f = _ctx(run, 1)
    f.load_event(0, e)
    f.init()
    f.run()
    loc, rows = _adjoint_sources(e, f.seismograms())
    a = _ctx(run, 2)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    for st in np.flatnonzero(f.saving_vector()) + 1:
        a.put_history(int(st), f.history(int(st)))
    a.run()
    K1 = a.kernels()
    f2 = _ctx(run, 1)
    f2.load_event(0, e)
    f2.init()
    f2.forward_checkpointed(7, keep=keep)
    for rep in range(2):
        a2 = _ctx(run, 2)
        a2.set_adjoint_sources(0, loc, rows)
        a2.init()
        a2.adjoint_checkpointed(f2)
        K2 = a2.kernels()
        a2.close()
        for g in K1:
            assert np.linalg.norm(K1[g]) > 0, g
            np.testing.assert_array_equal(K2[g], K1[g], err_msg=f"adjoint {rep + 1}: {g}")
    for s in (f, a, f2):
        s.close()


def test_prepared_checkpoints_allocate_nothing(tmp_path):
    """prepare_checkpoints(k, keep) allocates the checkpoint storage ahead of
    forward_checkpointed, which then allocates nothing; the kernels equal those
    of an unprepared context bit for bit; a prepare between forward and adjoint
    discards the checkpointed forward, so the adjoint refuses to run until the
    forward is repeated (from a fresh init, which restarts the wave field and
    keeps the prepared storage)."""
    from ses3d_amd.solver import Ses3dError
    run = make_case(str(tmp_path), nel=(6, 6, 4), nt=60, is_diss=True, adjoint_flag=1, samp_ad=4, seed=14)
    e = run.events[0]

    def gradient(prepare, keep):
        f = _ctx(run, 1)
        f.load_event(0, e)
        f.init()
        if prepare:
            f.prepare_checkpoints(7, keep=keep)
            b0 = f.memory_bytes()
        f.forward_checkpointed(7, keep=keep)
        if prepare:
            assert f.memory_bytes() == b0
        loc, rows = _adjoint_sources(e, f.seismograms())
        a = _ctx(run, 2)
        a.set_adjoint_sources(0, loc, rows)
        a.init()
        if prepare:
            f.prepare_checkpoints(7, keep=keep)
            with pytest.raises(Ses3dError, match="forward_checkpointed"):
                a.adjoint_checkpointed(f)
            f.init()
            f.forward_checkpointed(7, keep=keep)
            assert f.memory_bytes() == b0
        a.adjoint_checkpointed(f)
        K = a.kernels()
        f.close()
        a.close()
        return K

    for keep in (1, 3):
        print(f"prepared checkpoints, keep {keep}", flush=True)
        K0 = gradient(False, keep)
        K1 = gradient(True, keep)
        for g in K0:
            assert np.linalg.norm(K0[g]) > 0, g
            np.testing.assert_array_equal(K1[g], K0[g], err_msg=f"keep {keep} {g}")


def test_kept_segments_two_events_lower_degree(tmp_path):
    """Kept segments in a two-event context whose sampled steps are stored at
    degree 2: the kernels of every event equal those of keep=1 bit for bit."""
    from ses3d_amd.solver import Solver
    run = make_case(str(tmp_path), nel=(5, 6, 4), nt=45, is_diss=True, adjoint_flag=1, samp_ad=2, seed=31,
                    n_events=2)

    def ctx(af):
        s = Solver.from_run(run, adjoint_flag=af)
        s.set_history_degree(2)
        s.set_model(run.model)
        return s

    def kernels(keep):
        f = ctx(1)
        for k, e in enumerate(run.events):
            f.load_event(k, e)
        f.init()
        f.forward_checkpointed(6, keep=keep)
        a = ctx(2)
        for k, e in enumerate(run.events):
            loc, rows = _adjoint_sources(e, f.seismograms(k))
            a.set_adjoint_sources(k, e.rec32[f.receiver_index(k)], rows)
        a.init()
        a.adjoint_checkpointed(f)
        out = [a.kernels(k) for k in range(2)]
        f.close()
        a.close()
        return out

    print("two events, degree 2, keep 1 vs 4 vs all", flush=True)
    K1 = kernels(1)
    for keep in (4, 8):
        Kk = kernels(keep)
        for ev in range(2):
            for g in K1[ev]:
                assert ev > 0 or np.linalg.norm(K1[ev][g]) > 0, g
                np.testing.assert_array_equal(Kk[ev][g], K1[ev][g], err_msg=f"keep {keep} event {ev} {g}")


def test_anisotropic_adjoint(tmp_path):
    """Adjoint run of an anisotropic anelastic model (nonzero A, B, C): the
    adjoint wavefield equals the float32 oracle's bit for bit and the kernels
    agree to relative L2 <= 1e-5 (float64 accumulation)."""
    from oracle import pyoracle as PO
    run = make_case(str(tmp_path), nel=(5, 5, 4), nt=40, is_diss=True, adjoint_flag=1, samp_ad=2, seed=17)
    rng = np.random.default_rng(3)
    for k in ("A", "B", "C"):
        run.model[k] = np.asfortranarray((run.model["mu"] * rng.uniform(-0.05, 0.05, run.model["mu"].shape))
                                         .astype(np.float32))
    e = run.events[0]
    nt = e.nt
    print("anisotropic adjoint 5x5x4", flush=True)
    f = _ctx(run, 1)
    f.load_event(0, e)
    f.init()
    f.run()
    loc, rows = _adjoint_sources(e, f.seismograms())
    a = _ctx(run, 2)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    for st in np.flatnonzero(f.saving_vector()) + 1:
        a.put_history(int(st), f.history(int(st)))
    a.run()
    K = a.kernels()
    of = oracle_for(run, e, adjoint_flag=1)
    of.run(1, nt)
    oa = PO.Oracle(nx=run.nx, ny=run.ny, nz=run.nz, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin,
                   ymax=run.ymax, zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global, nt=nt,
                   dt=e.dt, is_diss=run.is_diss, adjoint_flag=2, samp_ad=run.samp_ad, pml=2,
                   tau_p=run.tau_p, D_p=run.D_p)
    oa.set_model(run.model)
    oa.set_adjoint_sources(loc, rows)
    oa.init()
    oa.run(1, nt, fwd=of)
    for c in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(a.field(c + "_global"), oa.get(c + "g"), err_msg=c)
    for g in KNAMES:
        ref = oa.get(g)
        assert np.linalg.norm(ref) > 0, g
        rel = _rel(K[g], ref)
        print(f"  {g} vs oracle32: rel L2 {rel:.2e}", flush=True)
        assert rel <= 1e-5, (g, rel)
    f.close()
    a.close()


@pytest.mark.parametrize("retire", ["init", "run", "set_model"])
def test_adjoint_refuses_retired_checkpoints(tmp_path, retire):
    """A checkpointed forward belongs to the model, events and step state it ran
    with: after set_model, init or a run of the forward context, the adjoint
    must refuse (ESTATE) before any kernel runs, instead of replaying stale
    checkpoints or reading a slot table another run uploaded; a fresh
    forward_checkpointed restores the exact kernels."""
    from ses3d_amd.solver import Ses3dError
    run = make_case(str(tmp_path), nel=(5, 5, 4), nt=40, is_diss=True, adjoint_flag=1, samp_ad=3, seed=21)
    e = run.events[0]
    print(f"retire by {retire}", flush=True)
    f = _ctx(run, 1)
    f.load_event(0, e)
    f.init()
    f.forward_checkpointed(8, keep=2)
    loc, rows = _adjoint_sources(e, f.seismograms())
    a = _ctx(run, 2)
    a.set_adjoint_sources(0, loc, rows)
    a.init()
    a.adjoint_checkpointed(f)
    K1 = a.kernels()
    if retire == "init":
        f.init()
    elif retire == "run":
        f.run()
    else:
        f.set_model(run.model)
        f.init()
    a.init()
    with pytest.raises(Ses3dError, match="error -3"):
        a.adjoint_checkpointed(f)
    f.init()
    f.forward_checkpointed(8, keep=2)
    a.init()
    a.adjoint_checkpointed(f)
    K2 = a.kernels()
    for g in K1:
        np.testing.assert_array_equal(K2[g], K1[g], err_msg=g)
    f.close()
    a.close()
