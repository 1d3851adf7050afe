"""The element kernel's two cache policies on the GPU.

Sections whose element-local state exceeds the 256 MiB last-level cache
stream it (displacement gradient, split stresses, memory variables,
coefficients, sampled fields) with non-temporal loads and stores; smaller
sections keep it cached.  Only the cache hints differ, so forcing either
variant with ses3d_hip_set_streaming must give bit-identical results.  The
cached variant is pinned to the oracle by the other GPU tests and the
streaming forward kernel by the first steps of the 96x96x48 workload
(test_gpu_fullsize.py).  Here every other streaming instantiation is pinned
on small meshes: forward kernels (elastic, anelastic, anisotropic), the
forward that stores sampled fields (full and lower polynomial degree), the
adjoint with kernel accumulation through HBM checkpoints (the path the
48x48x24 gradient runs non-temporally), and a group of sections.
"""
import numpy as np
import pytest

from tests.cases import make_case
from tests.test_gpu_adjoint import _adjoint_sources
from tests.test_gpu_forward import DGRAD, STRESS, _assert_same, _gpu

pytestmark = pytest.mark.gpu

MEMVARS = ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz")


def _anisotropic(run, seed=3):
    rng = np.random.default_rng(seed)
    for k in ("A", "B", "C"):
        run.model[k] = np.asfortranarray((run.model["mu"] * rng.uniform(-0.05, 0.05, run.model["mu"].shape))
                                         .astype(np.float32))
    return run


def test_streaming_mode_by_state_size(tmp_path):
    from ses3d_amd.solver import Ses3dError, Solver
    run = make_case(str(tmp_path), nel=(4, 4, 3), nt=8, is_diss=True, seed=1)
    s = _gpu(run)
    assert s.streaming() == {"mode": -1, "non_temporal": False}
    s.set_streaming(1)
    assert s.streaming() == {"mode": 1, "non_temporal": True}
    s.set_streaming(0)
    assert s.streaming() == {"mode": 0, "non_temporal": False}
    with pytest.raises(Ses3dError, match="streaming"):
        s.set_streaming(2)
    s.close()
    # 30x30x16 anelastic elements: 4 B x 1.85 M slots x 43 streamed values > 256 MiB
    big = Solver(nx=29, ny=29, nz=15, xmin=run.xmin, xmax=run.xmax, ymin=run.ymin, ymax=run.ymax,
                 zmin=run.zmin, zmax=run.zmax, zmax_global=run.zmax_global, nt=4, dt=0.2, is_diss=1,
                 tau_p=run.tau_p, D_p=run.D_p)
    assert big.streaming() == {"mode": -1, "non_temporal": True}
    big.set_streaming(0)
    assert big.streaming()["non_temporal"] is False
    big.close()


def _forward_outputs(run, af, streaming, hdeg=4):
    s = _gpu(run, adjoint_flag=af)
    if hdeg != 4:
        s.set_history_degree(hdeg)
        s.init()
    s.set_streaming(streaming)
    assert s.streaming()["non_temporal"] == bool(streaming)
    s.run(1, 20)                       # one captured 16-step graph and 4 direct steps
    s.set_graph_steps(0)
    s.run(21, run.events[0].nt - 20)   # every kernel launched directly
    out = {"seismograms": s.seismograms()}
    for c in ("vx", "vy", "vz"):
        out[c] = s.field(c + "_global")
    for f in STRESS + (DGRAD if af else ()) + (MEMVARS if run.is_diss else ()):
        out[f] = s.field(f)
    if af == 1:
        for st in np.flatnonzero(s.saving_vector()) + 1:
            out[f"history {st}"] = s.history(int(st))
    s.close()
    return out


@pytest.mark.parametrize("is_diss,aniso,af,hdeg", [(0, 0, 0, 4), (1, 0, 0, 4), (1, 1, 0, 4), (0, 1, 1, 4),
                                                   (1, 0, 1, 4), (1, 0, 1, 2)])
def test_streaming_forward_bitexact(tmp_path, is_diss, aniso, af, hdeg):
    run = make_case(str(tmp_path), nel=(6, 5, 4), nt=36, is_diss=bool(is_diss), adjoint_flag=af, samp_ad=3,
                    seed=40 + 2 * is_diss + aniso)
    if aniso:
        _anisotropic(run)
    print(f"streaming vs cached: diss={is_diss} aniso={aniso} adjoint_flag={af} degree={hdeg}", flush=True)
    cached = _forward_outputs(run, af, 0, hdeg)
    streamed = _forward_outputs(run, af, 1, hdeg)
    assert cached.keys() == streamed.keys()
    assert np.abs(cached["seismograms"]).max() > 0
    for k in cached:
        _assert_same(streamed[k], cached[k], k)


@pytest.mark.parametrize("is_diss,hdeg", [(1, 4), (0, 4), (1, 3)])
def test_streaming_checkpointed_gradient_bitexact(tmp_path, is_diss, hdeg):
    """Checkpointed forward and adjoint (kernel accumulation, element kernel
    modes 1 and 2) with both contexts forced non-temporal: the kernels equal
    those of the cached variant exactly."""
    from ses3d_amd.solver import Solver
    run = make_case(str(tmp_path), nel=(5, 6, 4), nt=48, is_diss=bool(is_diss), adjoint_flag=1, samp_ad=2,
                    seed=60 + is_diss + hdeg)
    e = run.events[0]

    def ctx(af, streaming):
        s = Solver.from_run(run, adjoint_flag=af)
        s.set_history_degree(hdeg)
        s.set_model(run.model)
        s.set_streaming(streaming)
        return s

    def gradient(streaming):
        f = ctx(1, streaming)
        f.load_event(0, e)
        f.init()
        f.forward_checkpointed(10)
        loc, rows = _adjoint_sources(e, f.seismograms())
        a = ctx(2, streaming)
        a.set_adjoint_sources(0, loc, rows)
        a.init()
        a.adjoint_checkpointed(f)
        out = dict(a.kernels())
        out["adjoint vx"] = a.field("vx_global")
        out["adjoint Mxy" if is_diss else "adjoint sxy"] = a.field("Mxy" if is_diss else "sxy")
        f.close()
        a.close()
        return out

    print(f"checkpointed gradient, streaming vs cached: diss={is_diss} degree={hdeg}", flush=True)
    K0, K1 = gradient(0), gradient(1)
    assert K0.keys() == K1.keys()
    for g in K0:
        assert np.abs(K0[g]).max() > 0, g
        np.testing.assert_array_equal(K1[g], K0[g], err_msg=g)


def test_streaming_group_bitexact(tmp_path):
    from ses3d_amd.solver import Group
    nt = 30
    runs = [make_case(str(tmp_path / "a"), nel=(6, 5, 4), nt=nt, is_diss=True, seed=71, n_events=2),
            make_case(str(tmp_path / "b"), nel=(4, 4, 5), nt=nt, is_diss=True, seed=72)]

    def group_run(streaming):
        sols = [_gpu(r) for r in runs]
        with Group(sols) as g:
            assert g.streaming() == {"mode": -1, "non_temporal": False}
            g.set_streaming(streaming)
            assert g.streaming()["non_temporal"] == bool(streaming)
            g.run()
        out = [[s.seismograms(k) for k in range(s.n_events)] + [s.field("Mxz")] for s in sols]
        for s in sols:
            s.close()
        return out

    print("group of 2 sections, streaming vs cached", flush=True)
    a, b = group_run(0), group_run(1)
    for i, (xa, xb) in enumerate(zip(a, b)):
        for j, (u, v) in enumerate(zip(xa, xb)):
            _assert_same(v, u, f"section {i} output {j}")
