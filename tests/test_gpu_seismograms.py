"""Seismograms of the fused-run FP32 path against FP64: 16 fused runs at
order 5, each with its own Ricker moment-tensor point source, 64 receivers
recording every run at every step (the reference's step: receivers at the
step start, local, sources over the step, neigh), advanced 200 steps.  The
FP64 reference advances each run on its own through the FP64 single-run
kernels, which the parity tests pin to the FP64 oracle at 1e-12.  Bar:
relative L2 of every run's seismograms <= 1e-5."""
import numpy as np
import pytest

from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.points import PointSources, Receivers, moment_tensor_scalings, ricker
from edge_amd.solver import KERNEL_MFMA, KERNEL_MFMA64, ElasticSolver

pytestmark = pytest.mark.gpu


RUNS = 16


def _sources(seed=3343):
    """Per run: point, Ricker peak frequency and series, moment-tensor scalings;
    the 64 receiver points."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.1, 0.9, size=(RUNS, 3))
    f0 = rng.uniform(20.0, 40.0, RUNS)
    dts = 1e-3 / f0
    series = [ricker(f0[r], 1.2 / f0[r], dts[r], int(2.4 / f0[r] / dts[r]) + 1) for r in range(RUNS)]
    scas = []
    for r in range(RUNS):
        a = rng.normal(size=(3, 3))
        scas.append(moment_tensor_scalings(np.triu(a) + np.triu(a, 1).T))
    return pts, dts, series, np.array(scas), rng.uniform(0.05, 0.95, size=(64, 3))


def _run(m, rho, lam, mu, prec, kernel, dt, n_steps, which):
    """Seismograms of the runs `which` fused in one solver."""
    import torch
    pts, dts, series, scas, rpts = _sources()
    s = ElasticSolver(m, 5, len(which), prec, rho, lam, mu, kernel=kernel)
    src = PointSources(s, pts[which], np.arange(len(which)), scas[which], np.zeros(len(which)), dts[which],
                       [series[r] for r in which])
    recv = Receivers(s, rpts, freq=dt)
    dofs, t_int = s.alloc(), s.alloc()
    for i in range(n_steps):
        t = i * dt
        recv.sample(dofs, t, dt)
        s.local(dofs, t_int, dt)
        src.apply(t, t + dt, dofs)
        s.neigh(dofs, t_int)
    torch.cuda.synchronize()
    out = [recv.seismogram(i) for i in range(recv.n)]
    src.close()
    recv.close()
    s.close()
    return out


def test_fp32_fused_seismograms_match_fp64(gpu):
    m = M.box(4, jitter=0.1, seed=7)
    rho, lam, mu = S.random_materials(m.n_el, seed=7)
    dt = S.dt_cfl(m, rho, lam, mu, 5)
    n_steps = 200
    s32 = _run(m, rho, lam, mu, 32, KERNEL_MFMA, dt, n_steps, np.arange(RUNS))
    v32 = np.array([v for _, v in s32])  # [receiver][sample][q][run]
    assert v32.shape == (64, n_steps, 9, RUNS)
    errs = []
    for r in range(RUNS):
        s64 = _run(m, rho, lam, mu, 64, KERNEL_MFMA64, dt, n_steps, np.array([r]))
        for (t32, _), (t64, _) in zip(s32, s64):
            np.testing.assert_allclose(t32, t64, rtol=1e-12, atol=1e-15)
        ref = np.array([v for _, v in s64])[..., 0]
        # the waves reach most receivers within the 200 steps
        peak = np.abs(ref).max(axis=(1, 2))
        assert (peak > 1e-3 * peak.max()).mean() >= 0.5, r
        err = np.linalg.norm(v32[..., r] - ref) / np.linalg.norm(ref)
        errs.append(err)
    print("relative L2 of the FP32 seismograms per run:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-5, errs
