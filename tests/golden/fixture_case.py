"""Compact, self-contained SES3D test cases for the golden fixtures.

A fixture stores every input value it needs (geometry, event, receivers, the
source time function, the moment tensor and the model: a 1D model as
per-(k,n) vectors, or, for the cases in MODEL_3D, whole laterally varying
float32 arrays in reference layout), so the run tree can be rebuilt bit for bit
on any machine without regenerating random numbers or calling transcendental
functions.
"""
from __future__ import annotations

import os

import numpy as np

from ses3d_amd import formats as F, models as M

CASES = {
    # name: (nel, nt, dt, is_diss, source_type, src (colat, lon, depth), samp_ad, adjoint)
    "elastic_mt": ((5, 6, 3), 60, 0.2, 0, 10, (50.2, 20.7, 25000.0), 3, False),
    "anelastic_adjoint": ((5, 5, 4), 50, 0.2, 1, 10, (49.3, 19.1, 60000.0), 4, True),
    "force_r": ((4, 5, 3), 40, 0.2, 0, 3, (51.0, 21.5, 15000.0), 3, False),
    "aniso_mt": ((6, 5, 4), 40, 0.2, 1, 10, (50.5, 19.7, 40000.0), 3, False),
    # laterally varying (3D) models; point forces along theta (type 1) and phi
    # (type 2); an anelastic adjoint whose last step (47) is not a multiple of
    # samp_ad, with the shear-Q kernel of SOURCE/ses3d_grad_Q.f90
    "force_theta_3d": ((5, 4, 3), 40, 0.2, 0, 1, (50.6, 19.4, 20000.0), 3, False),
    "force_phi_3d": ((4, 5, 3), 40, 0.2, 1, 2, (49.5, 20.6, 35000.0), 3, False),
    "anelastic_3d_adjoint": ((4, 4, 3), 47, 0.2, 1, 10, (50.3, 20.2, 45000.0), 3, True),
    # split into 2 x 1 x 2 rank boxes by tests/golden/make_golden_multibox.py (elastic: the
    # reference's multi-rank anelastic runs leave sum(D_p) at 0 on ranks > 0, since
    # SOURCE/ses3d_input.f90:353 forms it on rank 0 only and :398-399 broadcast D_p but not the sum)
    "multibox_adjoint": ((6, 5, 4), 45, 0.2, 0, 10, (48.3, 19.1, 60000.0), 3, True),
}
# cases with a laterally varying model: seeded Gaussian-correlated perturbation
# (200 km correlation, 3% rms) of vp, vs and rho, and of Q (x3) when anelastic
MODEL_3D = ("force_theta_3d", "force_phi_3d", "anelastic_3d_adjoint", "multibox_adjoint")
# cases whose fixture comes from a multi-rank reference run (not make_golden.py)
MULTIBOX = {"multibox_adjoint": (2, 1, 2)}
MODEL_FIELDS = ("rhoinv", "mu", "lambda", "Q")


def compact_inputs(name: str, seed: int = 11) -> dict:
    nel, nt, dt, is_diss, st, src, samp_ad, adj = CASES[name]
    rng = np.random.default_rng(seed)
    setup = M.section_setup(*nel, is_diss=bool(is_diss), adjoint_flag=1 if adj else 0, samp_ad=samp_ad)
    box = F.make_boxes(setup)[0]
    _, _, z = M.gll_coords(box)
    rho, vp, vs = M.prem_like(z)
    mu = rho * vs * vs
    out = {
        "nel": np.array(nel), "nt": nt, "dt": dt, "is_diss": is_diss, "source_type": st,
        "src": np.array(src), "samp_ad": samp_ad, "adjoint": int(adj),
        "moment": np.round(M.random_moment_tensor(rng) / 1e15) * 1e15,
        "stf": M.ricker(nt + 8, dt, 8.0, t0=6.0),
        "rhoinv_z": (1.0 / rho).astype(np.float32),
        "mu_z": mu.astype(np.float32),
        "lambda_z": (rho * vp * vp - 2.0 * mu).astype(np.float32),
        "Q_z": M.q_mu_by_depth(z).astype(np.float32),
    }
    if name.startswith("aniso"):
        # depth-dependent anisotropy parameters of a few percent of mu
        out["A_z"] = (0.04 * mu * np.cos(z / 9e4)).astype(np.float32)
        out["B_z"] = (-0.03 * mu * np.sin(z / 7e4)).astype(np.float32)
        out["C_z"] = (0.05 * mu * np.cos(z / 5e4)).astype(np.float32)
    if name in MODEL_3D:
        pert = M.gaussian_random_field(box, seed=seed + 100, corr_len=200e3, rms=0.03)
        m3 = M.make_model(box, is_diss=bool(is_diss), perturbation=pert)
        if is_diss:
            m3["Q"] = np.asfortranarray((m3["Q"] * (1.0 + 3.0 * pert)).astype(np.float32))
        for k in MODEL_FIELDS:
            if k in m3:
                out[f"model3d_{k}"] = np.asarray(m3[k], np.float32)
    rec = M.surface_receiver_grid(setup, 2, 2)
    out["rec"] = np.stack([rec.colat, rec.lon, rec.depth], 1)
    return out


def write_tree(fx: dict, root: str, adjoint_flag: int = None, decomp=(1, 1, 1)) -> None:
    """Write the run tree for a (loaded or fresh) compact fixture; ``decomp``
    (px, py, pz) splits the section into that many rank boxes, each with its
    own model records, as MODELS/SOURCE/generate_models.f90 lays them out."""
    nel = tuple(int(v) for v in fx["nel"])
    nt, dt = int(fx["nt"]), float(fx["dt"])
    is_diss = int(fx["is_diss"])
    adj = int(fx["adjoint"])
    af = (1 if adj else 0) if adjoint_flag is None else adjoint_flag
    setup = M.section_setup(*nel, is_diss=bool(is_diss), adjoint_flag=af, samp_ad=int(fx["samp_ad"]))
    box = F.make_boxes(setup)[0]
    shape = (box.nx + 1, box.ny + 1, box.nz + 1, 5, 5, 5)
    if tuple(decomp) != (1, 1, 1):
        setup = M.decompose(setup, *decomp)

    def bz(v):
        return np.asfortranarray(np.broadcast_to(np.asarray(v, np.float32)[None, None, :, None, None, :], shape))

    model = {"rhoinv": bz(fx["rhoinv_z"]), "mu": bz(fx["mu_z"]), "lambda": bz(fx["lambda_z"]),
             "A": bz(fx["A_z"]) if "A_z" in fx else np.zeros(shape, np.float32, order="F"),
             "B": bz(fx["B_z"]) if "B_z" in fx else np.zeros(shape, np.float32, order="F"),
             "C": bz(fx["C_z"]) if "C_z" in fx else np.zeros(shape, np.float32, order="F")}
    if is_diss:
        model["Q"] = bz(fx["Q_z"])
    for k in MODEL_FIELDS:
        if f"model3d_{k}" in fx:
            a = np.asarray(fx[f"model3d_{k}"], np.float32)
            assert a.shape == shape, (k, a.shape, shape)
            model[k] = np.asfortranarray(a)
    src = [float(v) for v in fx["src"]]
    ev = F.Event(nt=nt, dt=dt, xxs=src[0], yys=src[1], zzs=src[2], source_type=int(fx["source_type"]),
                 moment=tuple(float(v) for v in fx["moment"]), ofd="../DATA/OUTPUT/1/", ssamp=nt,
                 output_displacement=1)
    r = np.asarray(fx["rec"], np.float64)
    rec = F.Receivers([f"R{k:03d}.GG" for k in range(len(r))], r[:, 0], r[:, 1], r[:, 2])
    case = M.RunCase(setup, [ev], [rec], np.asarray(fx["stf"], np.float32), model)
    M.write_run_tree(root, case)


def adjoint_inputs(fx: dict, seis: np.ndarray) -> tuple:
    """Adjoint sources at the receivers: an L2-type residual built from the
    forward seismograms (3, nr, nt), written in forward time, scaled."""
    r = np.asarray(fx["rec"], np.float64)
    adsrc = np.stack([seis[:, k, :].astype(np.float64) * 1e6 for k in range(len(r))])
    return r, adsrc


def write_adjoint(fx: dict, root: str, seis: np.ndarray) -> None:
    locs, adsrc = adjoint_inputs(fx, seis)
    d = os.path.join(root, "ADJOINT", "1")
    os.makedirs(d, exist_ok=True)
    F.write_ad_srcfile(os.path.join(d, "ad_srcfile"), locs)
    for k in range(len(locs)):
        F.write_ad_src(os.path.join(d, f"ad_src_{k + 1}"), locs[k], adsrc[k])
    s = F.Setup.read(os.path.join(root, "INPUT", "setup"))
    s.adjoint_flag = 2
    s.write(os.path.join(root, "INPUT", "setup"))


# ---------------------------------------------------------------------------
# the elastic workload at the reference build's one-rank maximum (23x28x8)
# ---------------------------------------------------------------------------
def _elastic_large():
    from ses3d_amd import configs
    return configs.Workload("elastic_23x28x8", (23, 28, 8), False, 2, 0.2, 2000, 20, "none")


ELASTIC_LARGE = _elastic_large()
ELASTIC_LARGE_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elastic_23x28x8.npz")


def input_digest(run) -> str:
    """sha256 over the float32 inputs of event 1 and the model of a RunInput."""
    import hashlib
    h = hashlib.sha256()
    e = run.events[0]
    for a in (e.stf, e.moment, e.rec32, np.array([e.xxs, e.yys, e.zzs], np.float32)):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    for k in sorted(run.model):
        h.update(np.asarray(run.model[k], np.float32).tobytes(order="F"))
    return h.hexdigest()


def field_digest(a) -> str:
    """sha256 of a float32 field in the reference's (Fortran) layout."""
    import hashlib
    return hashlib.sha256(np.asarray(a, np.float32).tobytes(order="F")).hexdigest()
