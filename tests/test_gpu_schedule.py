"""Launch schedules, the autotuner and groups of sections on the GPU.

A schedule only changes which workgroup computes which element or node, and a
group only places several sections' workgroups in one launch; no float32
operation changes.  So every result must equal the oracle's (itself bit-exact
with the reference's float32 build) bit for bit, and group kernels must equal
single-context kernels exactly."""
import numpy as np
import pytest

from tests.cases import make_case, oracle_for
from tests.test_gpu_forward import _assert_same, _gpu

pytestmark = pytest.mark.gpu

SCHEDULES = [(e, n, r) for e in (1, 2, 4) for n in (128, 256, 512) for r in (0, 1)]


def _oracle_run(run, ev, nt, af=0):
    o = oracle_for(run, ev, adjoint_flag=af)
    o.run(1, nt)
    return o


def _check_vs_oracle(s, refs, what):
    for k, o in enumerate(refs):
        _assert_same(s.seismograms(k), o.get("seismograms"), f"seismograms ev{k} {what}")
        for c in ("vx", "vy", "vz"):
            _assert_same(s.field(c + "_global", k), o.get(c + "g"), f"{c} ev{k} {what}")
        _assert_same(s.field("sxy", k), o.get("sxy"), f"sxy ev{k} {what}")
        _assert_same(s.field("Mxz", k), o.get("Mxz"), f"Mxz ev{k} {what}")


def test_every_schedule_bitexact(tmp_path):
    run = make_case(str(tmp_path), nel=(5, 6, 4), nt=24, is_diss=True, seed=11, n_events=2)
    refs = [_oracle_run(run, ev, 24) for ev in run.events]
    s = _gpu(run)
    for (e, n, r) in SCHEDULES:
        print(f"schedule epb={e} node_lanes={n} remap={r}", flush=True)
        s.set_schedule(e, n, r, 1 - r)
        sch = s.schedule()
        assert (sch["epb"], sch["node_threads"], sch["remap_elem"], sch["remap_node"]) == (e, n, r, 1 - r)
        s.init()
        s.run()
        _check_vs_oracle(s, refs, (e, n, r))
    s.close()


def test_invalid_schedule_rejected(tmp_path):
    from ses3d_amd.solver import Ses3dError
    run = make_case(str(tmp_path), nel=(3, 3, 3), nt=8, seed=1)
    s = _gpu(run)
    for bad in ((3, 256, 1, 1), (2, 64, 1, 1), (2, 256, 2, 1)):
        with pytest.raises(Ses3dError, match="schedule"):
            s.set_schedule(*bad)
    with pytest.raises(Ses3dError, match="at least 6"):
        s.autotune(1, 5)
    s.close()


def test_autotune_matches_plain_run(tmp_path):
    run = make_case(str(tmp_path), nel=(6, 5, 4), nt=44, is_diss=True, seed=12)
    refs = [_oracle_run(run, run.events[0], 44)]
    s = _gpu(run)
    sch = s.autotune(1, 20)          # 18 steps under the 6 trial schedules, 2 under the chosen one
    print("autotuned", sch, flush=True)
    assert sch["epb"] in (1, 2, 4) and sch["node_threads"] in (128, 256, 512)
    assert sch["trial_ms"] is not None and all(t > 0 for row in sch["trial_ms"] for t in row)
    best_e = min(range(6), key=lambda k: sch["trial_ms"][k][0])
    assert sch["epb"] == (1, 2, 4)[best_e % 3] and sch["remap_elem"] == best_e // 3
    s.run(21, 24)
    _check_vs_oracle(s, refs, "autotuned")
    s.close()


def _sections(tmp_path, nt, af=0, samp_ad=3):
    """Three sections of uneven element counts; the first holds two events."""
    return [make_case(str(tmp_path / "a"), nel=(6, 5, 4), nt=nt, is_diss=True, adjoint_flag=af, samp_ad=samp_ad,
                      seed=21, n_events=2),
            make_case(str(tmp_path / "b"), nel=(4, 4, 4), nt=nt, is_diss=True, adjoint_flag=af, samp_ad=samp_ad,
                      seed=22),
            make_case(str(tmp_path / "c"), nel=(7, 3, 5), nt=nt, is_diss=True, adjoint_flag=af, samp_ad=samp_ad,
                      seed=23, src=(47.0, 19.0, 80000.0))]


def test_group_forward_bitexact(tmp_path):
    from ses3d_amd.solver import Group
    nt = 36
    runs = _sections(tmp_path, nt)
    refs = [[_oracle_run(r, ev, nt) for ev in r.events] for r in runs]
    sols = [_gpu(r) for r in runs]
    g = Group(sols)
    print("group run 1..20, then members alone 21..36", flush=True)
    g.run(1, 20)                       # one captured graph of 16 steps + 4 direct steps
    for s in sols:
        s.run(21, nt - 20)
    for s, rr in zip(sols, refs):
        _check_vs_oracle(s, rr, "group then alone")
    print("group autotune 1..24, group run 25..36", flush=True)
    for s in sols:
        s.init()
    sch = g.autotune(1, 24)
    assert sch["trial_ms"] is not None
    g.run(25, nt - 24)
    for s, rr in zip(sols, refs):
        _check_vs_oracle(s, rr, f"group autotuned {sch['epb']},{sch['node_threads']}")
    g.close()
    for s in sols:
        s.close()


def test_group_members_with_and_without_receivers(tmp_path):
    """Receivers are sampled by workgroups ahead of the elements in the
    element launch; a group concatenates its members' receivers.  A member
    whose events record nothing (and one recording event beside a silent one)
    must leave every field and seismogram as in single-context runs."""
    from ses3d_amd.solver import Group
    nt = 30
    runs = _sections(tmp_path, nt)

    def build():
        sols = []
        for i, r in enumerate(runs):
            s = _gpu(r)
            if i == 1:
                s.set_receivers(0, np.zeros((0, 3), np.float32))    # member without receivers
            if i == 0:
                s.set_receivers(1, np.zeros((0, 3), np.float32))    # second event of member 0 silent
            s.init()
            sols.append(s)
        return sols

    alone = build()
    for s in alone:
        s.run()
    grouped = build()
    print("group of 3 members, receivers on 2 of 4 events", flush=True)
    with Group(grouped) as g:
        g.run(1, 20)
        g.run(21, nt - 20)
    assert alone[1].nr(0) == 0 and alone[0].nr(1) == 0 and alone[0].nr(0) > 0
    for a, b in zip(alone, grouped):
        for k in range(a.n_events):
            if a.nr(k):
                _assert_same(b.seismograms(k), a.seismograms(k), f"seismograms ev{k}")
            for c in ("vx", "vy", "vz"):
                _assert_same(b.field(c + "_global", k), a.field(c + "_global", k), f"{c} ev{k}")
    for s in alone + grouped:
        s.close()


def test_group_rejects_mixed_members(tmp_path):
    from ses3d_amd.solver import Group, Ses3dError
    a = _gpu(make_case(str(tmp_path / "a"), nel=(3, 3, 3), nt=8, is_diss=True, seed=1))
    b = _gpu(make_case(str(tmp_path / "b"), nel=(3, 3, 3), nt=8, is_diss=False, seed=2))
    with pytest.raises(Ses3dError, match="elastic"):
        Group([a, b])
    with pytest.raises(Ses3dError, match="twice"):
        Group([a, a])
    g = Group([a])
    with pytest.raises(Ses3dError, match="outside"):
        g.run(5, 8)
    g.close()


def test_group_adjoint_kernels_match_single(tmp_path):
    """Forward (storing the sampled fields) and adjoint with kernel
    accumulation as one group each: kernels equal those of single-context
    runs bit for bit."""
    from ses3d_amd.solver import Group, Solver
    nt = 40
    runs = _sections(tmp_path, nt, af=1, samp_ad=4)

    def ctx(run, af):
        s = Solver.from_run(run, adjoint_flag=af)
        s.set_model(run.model)
        return s

    def forward(grouped):
        fs = []
        for r in runs:
            f = ctx(r, 1)
            for k, e in enumerate(r.events):
                f.load_event(k, e)
            f.init()
            fs.append(f)
        if grouped:
            with Group(fs) as g:
                g.run()
        else:
            for f in fs:
                f.run()
        return fs

    def adjoint(fs, grouped):
        ads = []
        for r, f in zip(runs, fs):
            a = ctx(r, 2)
            for k, e in enumerate(r.events):
                seis = f.seismograms(k)
                rows = np.ascontiguousarray((seis.astype(np.float64) * 1e6).astype(np.float32)[:, :, ::-1]
                                            .transpose(1, 2, 0))
                a.set_adjoint_sources(k, e.rec32[f.receiver_index(k)], rows)
            a.init()
            for k in range(len(r.events)):
                for st in np.flatnonzero(f.saving_vector()) + 1:
                    a.put_history(int(st), f.history(int(st), k), k)
            ads.append(a)
        if grouped:
            with Group(ads) as g:
                g.run()
        else:
            for a in ads:
                a.run()
        return ads

    print("forward: group vs single", flush=True)
    fg, fs = forward(True), forward(False)
    for a, b, r in zip(fg, fs, runs):
        for k in range(len(r.events)):
            _assert_same(a.seismograms(k), b.seismograms(k), "forward seismograms")
    print("adjoint: group vs single", flush=True)
    ag, as_ = adjoint(fs, True), adjoint(fs, False)
    for a, b, r in zip(ag, as_, runs):
        for k in range(len(r.events)):
            Ka, Kb = a.kernels(k), b.kernels(k)
            assert np.abs(Kb["grad_csh"]).max() > 0
            for n in Kb:
                np.testing.assert_array_equal(Ka[n], Kb[n], err_msg=n)
    for s in fg + fs + ag + as_:
        s.close()
