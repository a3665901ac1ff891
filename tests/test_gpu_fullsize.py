"""Parity at the benchmark sizes and step counts, one GPU.

* 196,608 tets (32^3 hexes x 6, +-10 % seeded jitter, rho ~ U[2, 3],
  vp ~ U[4, 6], vs = vp / sqrt 3), order 5:
  - one FP64 step of the whole mesh against the FP64 oracle on its first
    4,096 elements (the oracle's local step over those elements and their
    face neighbours, its neighbour step over those elements), <= 1e-12;
  - FP32 single run against FP64 single run, P plane wave, 200 steps, <= 1e-5;
  - 16 fused FP32 runs, each with its own Ricker moment-tensor point source
    (peak frequency U[1, 5]), 64 receivers, 500 steps, against the FP64
    single-run update of each run: seismograms and final DOFs of every run
    <= 1e-5.
* the viscoelastic ensemble of 24 meshes (1,767,744 tets, 3 mechanisms,
  Q_p ~ U[50, 200] per mesh, Q_s = Q_p / 2): 8 fused FP32 runs against the
  FP64 single run, 300 graph-captured steps, DOFs <= 1e-5 per run and the
  anelastic memory variables <= 3e-5 (the runs' initial states are scaled
  copies, so the single FP64 run is every run's reference).
* a union of three small viscoelastic meshes stepped against the oracle on the
  union: FP64 1e-12, FP32 1e-5.
The plane waves are P waves of each element's own medium (the wave vector
2 pi (1, 1, 1), the element's P polarisation): a wave polarised for another
material splits into waves whose DOF norm swings by a factor of about 3, and
relative errors taken at its minima overstate the drift.
Relative L2 norms are taken on the GPU in FP64 (edge_el_l2_norms for whole
arrays; torch reductions for run slices)."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.points import PointSources, Receivers, moment_tensor_scalings, ricker
from edge_amd.solver import KERNEL_MFMA, KERNEL_MFMA64, ElasticSolver, ViscoSolver

pytestmark = pytest.mark.gpu

ORDER = 5


@pytest.fixture(scope="module")
def big():
    m = M.box(32, jitter=0.1, seed=3343, block=4)
    rho, lam, mu = S.random_materials(m.n_el, seed=3343)
    dt = S.dt_cfl(m, rho, lam, mu, ORDER)
    return m, rho, lam, mu, dt


def _rel_runs(a, b):
    """Per-run relative L2 of device arrays a [..][R] against b [..][R or 1] (FP64 on the GPU)."""
    import torch
    out = []
    for r in range(a.shape[-1]):
        x = a[..., r].double()
        y = b[..., r if b.shape[-1] > 1 else 0].double()
        out.append(float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(y)))
    return out


def test_fp64_step_of_the_whole_mesh_matches_oracle(gpu, big):
    import torch
    m, rho, lam, mu, dt = big
    n_s = 4096
    q0 = S.project(m, ORDER, S.PlaneWave(k=(2 * np.pi,) * 3), chunk=8192, materials=(rho, lam, mu))
    s = ElasticSolver(m, ORDER, 1, 64, rho, lam, mu, kernel=KERNEL_MFMA64)
    gd = torch.tensor(q0, device="cuda")
    gt = torch.zeros_like(gd)
    s.step(gd, gt, dt, 1)
    got = gd[:n_s].cpu().numpy()
    # the oracle: star matrices of the sample and its neighbours, flux solvers of all faces
    nb = np.unique(np.concatenate([np.arange(n_s), m.el_fa_el[:n_s].ravel()]))
    nb = nb[nb >= 0]
    ve = m.el_crds()
    star = np.zeros((m.n_el, 3, 81))
    for e in nb:
        star[e] = orc.star(rho[e], lam[e], mu[e], orc.jac(ve[e])[2])
    jdet = np.linalg.det(ve[:, 1:, :] - ve[:, :1, :])
    fs = np.zeros((2, m.n_el, 4, 81))
    import ctypes as C
    orc.lib().orc_init_fs(C.c_int64(m.n_fa), orc._p(np.ascontiguousarray(m.fa_el)),
                          orc._p(np.ascontiguousarray(m.el_fa)), orc._p(np.ascontiguousarray(m.fa_type)),
                          orc._p(orc._c(m.normals)), orc._p(orc._c(m.tangents)), orc._p(orc._c(m.areas)),
                          orc._p(orc._c(jdet)), orc._p(orc._c(rho)), orc._p(orc._c(lam)), orc._p(orc._c(mu)),
                          orc._p(fs[0]), orc._p(fs[1]))
    mo = orc.MeshOracle(m, ORDER, 1, star=star, fs=fs)
    d = np.ascontiguousarray(q0.copy())
    t = np.zeros_like(d)
    for e in nb:
        mo.local(d, t, dt, int(e), 1)
    mo.neigh(d, t, 0, n_s)
    err = np.linalg.norm(got - d[:n_s]) / np.linalg.norm(d[:n_s])
    print(f"full-mesh FP64 step, first {n_s} elements vs oracle: {err:.2e}")
    assert err <= 1e-12
    s.close()


def test_fp32_single_run_matches_fp64_over_200_steps(gpu, big):
    import torch
    m, rho, lam, mu, dt = big
    q0 = S.project(m, ORDER, S.PlaneWave(k=(2 * np.pi,) * 3), chunk=8192, materials=(rho, lam, mu))
    out = {}
    for prec in (64, 32):
        s = ElasticSolver(m, ORDER, 1, prec, rho, lam, mu)
        d = torch.tensor(q0, dtype=s.dtype, device="cuda")
        t = torch.zeros_like(d)
        s.step(d, t, dt, 200)
        torch.cuda.synchronize()
        out[prec] = (s, d)
    s64, d64 = out[64]
    diff, ref = s64.l2_norms(out[32][1], d64)
    err = float(np.sqrt(diff.sum() / ref.sum()))
    print(f"FP32 vs FP64, one run, 200 steps: {err:.2e} (kernel {out[32][0].kernel_info()[0]})")
    assert err <= 1e-5
    for s, _ in out.values():
        s.close()


def _config3_sources(m, runs, seed=3343):
    rng = np.random.default_rng(seed + 1)
    lo, hi = m.ve_crds.min(axis=0), m.ve_crds.max(axis=0)
    pts = lo + (hi - lo) * rng.uniform(0.05, 0.95, size=(runs, 3))
    f0 = rng.uniform(1.0, 5.0, runs)
    dts = 1e-3 / f0
    t0 = -3.0 / f0
    series = [ricker(f0[r], 0.0 - t0[r], dts[r], int(6.0 / f0[r] / dts[r]) + 1) for r in range(runs)]
    scas = []
    for r in range(runs):
        a = rng.normal(size=(3, 3))
        scas.append(moment_tensor_scalings(np.triu(a) + np.triu(a, 1).T))
    # the receivers: 64 seeded points, half of them near a source so that the waves reach them
    rp = np.concatenate([lo + (hi - lo) * rng.uniform(0.05, 0.95, size=(32, 3)),
                         np.clip(pts[rng.integers(0, runs, 32)] + rng.uniform(-0.03, 0.03, (32, 3)), lo, hi)])
    return pts, t0, dts, series, np.array(scas), rp


def _config3_run(m, rho, lam, mu, dt, prec, which, n_steps, src_data):
    import torch
    pts, t0, dts, series, scas, rp = src_data
    s = ElasticSolver(m, ORDER, len(which), prec, rho, lam, mu, kernel=KERNEL_MFMA if prec == 32 else KERNEL_MFMA64)
    src = PointSources(s, pts[which], np.arange(len(which)), scas[which], t0[which], dts[which],
                       [series[r] for r in which])
    recv = Receivers(s, rp, freq=dt)
    d, t = s.alloc(), s.alloc()
    for i in range(n_steps):
        tm = i * dt
        recv.sample(d, tm, dt)
        s.local(d, t, dt)
        src.apply(tm, tm + dt, d)
        s.neigh(d, t)
    torch.cuda.synchronize()
    seis = np.array([recv.seismogram(i)[1] for i in range(recv.n)])  # [recv][sample][q][run]
    src.close()
    recv.close()
    return s, d, seis


@pytest.mark.timeout(600)
def test_config3_fused16_matches_fp64_per_run_over_500_steps(gpu, big):
    m, rho, lam, mu, dt = big
    runs, n_steps = 16, 500
    src_data = _config3_sources(m, runs)
    s32, d32, v32 = _config3_run(m, rho, lam, mu, dt, 32, np.arange(runs), n_steps, src_data)
    assert v32.shape == (64, n_steps, 9, runs)
    errs_d, errs_s = [], []
    for r in range(runs):
        s64, d64, v64 = _config3_run(m, rho, lam, mu, dt, 64, np.array([r]), n_steps, src_data)
        errs_d.append(_rel_runs(d32[..., r:r + 1], d64)[0])
        ref = v64[..., 0]
        errs_s.append(float(np.linalg.norm(v32[..., r] - ref) / np.linalg.norm(ref)))
        s64.close()
        del d64
    print("config 3, 500 steps, per run: final DOFs", " ".join(f"{e:.1e}" for e in errs_d))
    print("config 3, 500 steps, per run: seismograms", " ".join(f"{e:.1e}" for e in errs_s))
    assert max(errs_d) <= 1e-5 and max(errs_s) <= 1e-5
    s32.close()


def _visco_union(sizes, seed=3343):
    meshes, mats = [], []
    for i, n in enumerate(sizes):
        m = M.box(n, jitter=0.1, seed=seed + i)
        r = np.random.default_rng(10_000 + i)
        rho = r.uniform(2.0, 3.0, m.n_el)
        vp = r.uniform(4.0, 6.0, m.n_el)
        mu = rho * (vp / np.sqrt(3.0)) ** 2
        qp = np.full(m.n_el, r.uniform(50, 200))
        meshes.append(m)
        mats.append((rho, rho * vp ** 2 - 2 * mu, mu, qp))
    u, _ = M.union(meshes)
    rho, lam, mu, qp = [np.concatenate([x[k] for x in mats]) for k in range(4)]
    return u, rho, lam, mu, qp


@pytest.mark.parametrize("prec,runs,tol", [(64, 1, 1e-12), (32, 8, 1e-5)])
def test_union_of_meshes_matches_oracle(gpu, prec, runs, tol):
    import torch
    order, nr, n_steps = 4, 3, 4
    u, rho, lam, mu, qp = _visco_union([3, 4, 3], seed=11)
    s = ViscoSolver(u, order, runs, prec, rho, lam, mu, qp, qp / 2, nr, 1.0, 10.0)
    visco = orc.setup_mesh_visco(u, rho, lam, mu, qp, qp / 2, nr, 1.0, 10.0)
    dt = S.dt_cfl(u, rho, visco["lam_e"], visco["mu_e"], order)
    mo = orc.MeshOracle(u, order, runs, visco=visco)
    rng = np.random.default_rng(2)
    n = S.n_modes(order)
    d0 = rng.normal(size=(u.n_el, 9, n, runs))
    a0 = 0.1 * rng.normal(size=(u.n_el, nr, 6, n, runs))
    d, t, a = d0.copy(), np.zeros_like(d0), a0.copy()
    mo.step(d, t, dt, n_steps, dofs_a=a)
    dt_t = torch.float64 if prec == 64 else torch.float32
    gd, gt, ga = (torch.tensor(x, dtype=dt_t, device="cuda") for x in (d0, np.zeros_like(d0), a0))
    s.step(gd, gt, dt, n_steps, ga, graph=True)
    for got, ref in ((gd, d), (ga, a)):
        g = got.cpu().double().numpy()
        assert np.linalg.norm(g - ref) / np.linalg.norm(ref) <= tol
    s.close()


@pytest.mark.timeout(900)
def test_ensemble_fused8_matches_fp64_over_300_steps(gpu):
    import torch
    order, n_steps = 5, 300
    sizes = [int(x) for x in np.random.default_rng(3343).integers(8, 33, 24)]
    u, rho, lam, mu, qp = _visco_union(sizes)
    scale = torch.linspace(0.5, 2.0, 8, dtype=torch.float64)
    res = {}
    for prec, runs in ((64, 1), (32, 8)):
        s = ViscoSolver(u, order, runs, prec, rho, lam, mu, qp, qp / 2, 3, 1.0, 10.0)
        if prec == 64:
            setup = s.setup_a()
            dt = S.dt_cfl(u, rho, setup["lam_e"], setup["mu_e"], order)
            q0 = torch.from_numpy(S.project(u, order, S.PlaneWave(k=(2 * np.pi,) * 3), chunk=16384,
                                            materials=(rho, lam, mu)))
        # FP32 run r starts from scale[r] q0, the FP64 reference from q0
        d = (q0.cuda() * (scale[:runs].cuda() if prec == 32 else 1.0)).to(s.dtype).contiguous()
        t, a = torch.zeros_like(d), s.alloc_a()
        s.step(d, t, dt, n_steps, a, graph=True)
        torch.cuda.synchronize()
        res[prec] = (s, d, a)
        del t
    _, d64, a64 = res[64]
    _, d32, a32 = res[32]
    errs, errs_a = [], []
    for r in range(8):
        errs.append(_rel_runs((d32[..., r:r + 1].double() / scale[r]), d64)[0])
        errs_a.append(_rel_runs((a32[..., r:r + 1].double() / scale[r]), a64)[0])
    print(f"ensemble {u.n_el} tets, 300 steps, FP32 8 runs vs FP64: DOFs", " ".join(f"{e:.1e}" for e in errs),
          "| anelastic DOFs", " ".join(f"{e:.1e}" for e in errs_a))
    # the wavefield (elastic DOFs) meets the FP32 bar; the anelastic memory
    # variables carry ~1e-5 from the first steps on, with the generic FP32
    # kernels as with the MFMA ones (FP32 rounding of the relaxation updates)
    assert max(errs) <= 1e-5
    assert max(errs_a) <= 3e-5
    for s, _, _ in res.values():
        s.close()
