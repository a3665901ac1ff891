"""Parity of the HIP element update with the FP64 oracle, through the C ABI.

Bars: FP64 relative L2 <= 1e-12 (set-up and steps), FP32 relative L2 <= 1e-5
against the FP64 oracle after the stated number of steps.
"""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import (KERNEL_EC64, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, KERNEL_MFMA_TILES,
                             PATH_FUSED, PATH_PROJECTIONS, PATH_RANK3_FS, PATH_SPARSE_STAR, ElasticSolver)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _problem(n=4, bc=(FA_PERIODIC,) * 3, jitter=0.1, seed=21):
    m = M.box(n, bc=bc, jitter=jitter, seed=seed)
    rho, lam, mu = S.random_materials(m.n_el, seed=seed)
    return m, rho, lam, mu


def _ic(m, order, n_runs, seed=3):
    """Plane wave per run with a different amplitude and direction per run."""
    rng = np.random.default_rng(seed)
    out = np.zeros((m.n_el, 9, S.n_modes(order), n_runs))
    for r in range(n_runs):
        k = 2 * np.pi * rng.integers(1, 3, size=3)
        pw = S.PlaneWave(1.0, 2.0, 1.0, k=k, amp=rng.uniform(0.5, 2.0))
        out[..., r] = S.project(m, order, pw)[..., 0]
    return out


def _to_dev(a, prec):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


@pytest.mark.parametrize("bc", [(FA_PERIODIC,) * 3, (FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC)])
def test_setup_matches_oracle(gpu, bc):
    m, rho, lam, mu = _problem(bc=bc)
    s = ElasticSolver(m, 4, 1, 64, rho, lam, mu, kernel=KERNEL_GENERIC)
    star, fs = s.setup()
    ostar, ofs, _ = orc.setup_mesh(m, rho, lam, mu)
    assert rel_l2(star, ostar) <= 1e-14
    assert np.all(np.isfinite(fs))
    assert rel_l2(fs, ofs) <= 1e-13


@pytest.mark.parametrize("kernel", [KERNEL_GENERIC, KERNEL_MFMA64, KERNEL_EC64])
@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_fp64_local_neigh_match_oracle(gpu, order, kernel):
    m, rho, lam, mu = _problem()
    s = ElasticSolver(m, order, 1, 64, rho, lam, mu, kernel=kernel)
    assert s.kernel_info()[0] == kernel
    mo = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 1)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.local(d, t, dt)
    gd, gt = _to_dev(dofs0, 64), _to_dev(np.zeros_like(dofs0), 64)
    s.local(gd, gt, dt)
    assert rel_l2(gt.cpu().numpy(), t) <= FP64_TOL
    assert rel_l2(gd.cpu().numpy(), d) <= FP64_TOL
    mo.neigh(d, t)
    s.neigh(gd, gt)
    assert rel_l2(gd.cpu().numpy(), d) <= FP64_TOL


@pytest.mark.parametrize("order,bc", [(4, (FA_PERIODIC,) * 3), (5, (FA_PERIODIC,) * 3),
                                      (4, (FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC))])
def test_fp64_steps_match_oracle(gpu, order, bc):
    m, rho, lam, mu = _problem(bc=bc)
    s = ElasticSolver(m, order, 1, 64, rho, lam, mu)
    assert s.kernel_info()[0] == KERNEL_EC64  # the single-run default
    mo = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 1)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    n = 20
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.step(d, t, dt, n)
    gd, gt = _to_dev(dofs0, 64), _to_dev(np.zeros_like(dofs0), 64)
    s.step(gd, gt, dt, n)
    assert rel_l2(gd.cpu().numpy(), d) <= FP64_TOL


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("bc", [(FA_PERIODIC,) * 3, (FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC)])
def test_ec64_matches_oracle(gpu, order, bc, prec):
    """The element-batched single-run kernels (16 elements per wave on FP64 /
    FP32 MFMA, the single-run default): 162 elements, so the last wave is
    partly filled; free-surface, outflow and periodic faces (every face
    orientation); tInt and DOFs of the local step, the neighbour step from
    the face projections, then steps.  FP64 1e-12, FP32 1e-5."""
    m, rho, lam, mu = _problem(n=3, bc=bc, seed=30 + order)
    s = ElasticSolver(m, order, 1, prec, rho, lam, mu)
    assert s.kernel_info() == (KERNEL_EC64, PATH_SPARSE_STAR | PATH_RANK3_FS | PATH_PROJECTIONS)
    mo = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 1, seed=order)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    tol = FP64_TOL if prec == 64 else FP32_TOL
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.local(d, t, dt)
    gd, gt = _to_dev(dofs0, prec), _to_dev(np.zeros_like(dofs0), prec)
    s.local(gd, gt, dt)
    assert rel_l2(gt.cpu().numpy(), t) <= tol
    assert rel_l2(gd.cpu().numpy(), d) <= tol
    mo.neigh(d, t)
    s.neigh(gd, gt)
    assert rel_l2(gd.cpu().numpy(), d) <= tol
    mo.step(d, t, dt, 5)
    s.step(gd, gt, dt, 5)
    assert rel_l2(gd.cpu().numpy(), d) <= tol


def test_random_star_and_flux_solvers_match_oracle(gpu):
    """Dense random star matrices / flux solvers (no physical structure)
    catch transposed or mis-indexed 9x9 products.  The FP64 MFMA kernels
    run at orders 3 and 5 too, where the modes do not fill whole k-steps
    (10 -> 12, 35 -> 36) and the operands' zero tails matter."""
    m = M.box(3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC))
    rng = np.random.default_rng(5)
    star = rng.normal(size=(m.n_el, 3, 81)) * 0.2
    fs = rng.normal(size=(2, m.n_el, 4, 81)) * 0.2
    for prec, kern, order in [(64, KERNEL_GENERIC, 4), (64, KERNEL_MFMA64, 3), (64, KERNEL_MFMA64, 4),
                              (64, KERNEL_MFMA64, 5), (32, KERNEL_GENERIC, 4)]:
        s = ElasticSolver(m, order, 1, prec, star=star, fs=fs, kernel=kern)
        assert s.kernel_info()[0] == kern
        mo = orc.MeshOracle(m, order, star=star, fs=fs)
        dofs0 = rng.normal(size=(m.n_el, 9, S.n_modes(order), 1))
        d, t = dofs0.copy(), np.zeros_like(dofs0)
        mo.step(d, t, 0.05, 2)
        gd, gt = _to_dev(dofs0, prec), _to_dev(np.zeros_like(dofs0), prec)
        s.step(gd, gt, 0.05, 2)
        tol = FP64_TOL if prec == 64 else FP32_TOL
        assert rel_l2(gd.cpu().numpy(), d) <= tol


@pytest.mark.parametrize("order", [3, 4, 5])
def test_fp32_steps_match_fp64_oracle(gpu, order):
    m, rho, lam, mu = _problem()
    s = ElasticSolver(m, order, 1, 32, rho, lam, mu)
    mo = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 1)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    n = 50
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.step(d, t, dt, n)
    gd, gt = _to_dev(dofs0, 32), _to_dev(np.zeros_like(dofs0), 32)
    s.step(gd, gt, dt, n)
    assert rel_l2(gd.cpu().numpy(), d) <= FP32_TOL


@pytest.mark.parametrize("prec,order", [(32, 4), (32, 5), (64, 3)])
def test_fused_runs_match_oracle(gpu, prec, order):
    """16 fused runs with different initial conditions, against the FP64 oracle."""
    m, rho, lam, mu = _problem(n=3)
    try:
        s = ElasticSolver(m, order, 16, prec, rho, lam, mu)
    except Exception as e:  # configuration without a kernel is reported, not faked
        pytest.skip(str(e))
    mo = orc.MeshOracle(m, order, 16, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 16)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    n = 10
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.step(d, t, dt, n)
    gd, gt = _to_dev(dofs0, prec), _to_dev(np.zeros_like(dofs0), prec)
    s.step(gd, gt, dt, n)
    got = gd.cpu().numpy()
    tol = FP64_TOL if prec == 64 else FP32_TOL
    for r in range(16):
        assert rel_l2(got[..., r], d[..., r]) <= tol


def test_graph_replays_match_eager(gpu):
    """The captured step replayed several times, with eager work and
    synchronisation in between, equals the eager steps every time."""
    import torch
    m, rho, lam, mu = _problem()
    order = 4
    s = ElasticSolver(m, order, 1, 64, rho, lam, mu)
    dofs0 = _ic(m, order, 1)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    ga, ta = _to_dev(dofs0, 64), _to_dev(np.zeros_like(dofs0), 64)
    gb, tb = _to_dev(dofs0, 64), _to_dev(np.zeros_like(dofs0), 64)
    stream = torch.cuda.Stream()
    for rep in range(4):
        with torch.cuda.stream(stream):
            s.step(ga, ta, dt, 2, graph=True, stream=stream)
        stream.synchronize()
        s.step(gb, tb, dt, 2)
        torch.cuda.synchronize()
        assert torch.equal(ga, gb), rep


def test_l2_norms(gpu):
    m, rho, lam, mu = _problem()
    order = 4
    s = ElasticSolver(m, order, 1, 64, rho, lam, mu)
    a = _ic(m, order, 1, seed=1)
    b = _ic(m, order, 1, seed=2)
    d, n = s.l2_norms(_to_dev(a, 64), _to_dev(b, 64))
    ve = m.el_crds()
    jd = np.abs(np.linalg.det(ve[:, 1:, :] - ve[:, :1, :]))
    mass = S.dg_matrices(order)["mass"]
    ref_d = np.einsum("e,m,eqmr->qr", jd, mass, (a - b) ** 2)
    ref_n = np.einsum("e,m,eqmr->qr", jd, mass, b ** 2)
    np.testing.assert_allclose(d, ref_d, rtol=1e-12)
    np.testing.assert_allclose(n, ref_n, rtol=1e-12)


@pytest.mark.parametrize("order", [3, 4, 5])
def test_gpu_plane_wave_convergence(gpu, order):
    """Measured order of the FP64 GPU path on the plane wave is p+1."""
    pw = S.PlaneWave(1.0, 2.0, 1.0, k=(2 * np.pi,) * 3)
    errs = []
    for n in (4, 8):
        m = M.box(n)
        ne = m.n_el
        rho, lam, mu = np.ones(ne), 2 * np.ones(ne), np.ones(ne)
        s = ElasticSolver(m, order, 1, 64, rho, lam, mu)
        g = _to_dev(S.project(m, order, pw), 64)
        t = _to_dev(np.zeros(s.dof_shape()), 64)
        T = 0.1
        dt = S.dt_cfl(m, rho, lam, mu, order)
        ns = int(np.ceil(T / dt))
        s.step(g, t, T / ns, ns)
        _, l2, _ = S.error_norms(m, order, g.cpu().numpy(), lambda x: pw(x, T), n1d=order + 2)
        errs.append(np.sqrt(l2.sum()))
    rate = np.log2(errs[0] / errs[1])
    print(f"order {order}: L2 errors {errs[0]:.3e} (4^3 hexes), {errs[1]:.3e} (8^3), rate {rate:.2f}")
    assert rate > order - 0.3, (errs, rate)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_mfma_random_matrices_match_oracle(gpu, order):
    """The MFMA-tiled FP32 fused-16 kernels with dense random star matrices
    and flux solvers (free-surface, outflow and periodic faces), one local and
    one neighbour pass, then three steps, against the FP64 oracle."""
    m = M.box(3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=4)
    rng = np.random.default_rng(order)
    star = rng.normal(size=(m.n_el, 3, 81)) * 0.2
    fs = rng.normal(size=(2, m.n_el, 4, 81)) * 0.2
    s = ElasticSolver(m, order, 16, 32, star=star, fs=fs, kernel=KERNEL_MFMA)
    mo = orc.MeshOracle(m, order, 16, star=star, fs=fs)
    dofs0 = rng.normal(size=(m.n_el, 9, S.n_modes(order), 16))
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.local(d, t, 0.05)
    gd, gt = _to_dev(dofs0, 32), _to_dev(np.zeros_like(dofs0), 32)
    s.local(gd, gt, 0.05)
    assert rel_l2(gt.cpu().numpy(), t) <= FP32_TOL
    assert rel_l2(gd.cpu().numpy(), d) <= FP32_TOL
    mo.neigh(d, t)
    s.neigh(gd, gt)
    assert rel_l2(gd.cpu().numpy(), d) <= FP32_TOL
    mo.step(d, t, 0.05, 3)
    s.step(gd, gt, 0.05, 3)
    got = gd.cpu().numpy()
    for r in range(16):
        assert rel_l2(got[..., r], d[..., r]) <= FP32_TOL


@pytest.mark.parametrize("kernel", [KERNEL_MFMA, KERNEL_MFMA_TILES])
@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_mfma_physical_solvers_match_oracle(gpu, order, kernel):
    """The MFMA-tiled FP32 fused-16 kernels with physical set-up: sparse star
    matrices and rank-3 factored flux solvers; the neighbour step from the
    face projections the local step wrote (selector 2) or from tInt tiles
    (selector 3).  Free-surface, outflow and periodic faces; local and
    neighbour passes separately, then steps, against the FP64 oracle."""
    m, rho, lam, mu = _problem(n=3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), seed=10 + order)
    s = ElasticSolver(m, order, 16, 32, rho, lam, mu, kernel=kernel)
    # exchanged projections: multi-step calls fuse neighbour step k - 1 with local step k
    want = PATH_SPARSE_STAR | PATH_RANK3_FS | (PATH_PROJECTIONS | PATH_FUSED if kernel == KERNEL_MFMA else 0)
    assert s.kernel_info() == (2, want)
    mo = orc.MeshOracle(m, order, 16, rho=rho, lam=lam, mu=mu)
    dofs0 = _ic(m, order, 16, seed=order)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.local(d, t, dt)
    gd, gt = _to_dev(dofs0, 32), _to_dev(np.zeros_like(dofs0), 32)
    s.local(gd, gt, dt)
    assert rel_l2(gt.cpu().numpy(), t) <= FP32_TOL
    assert rel_l2(gd.cpu().numpy(), d) <= FP32_TOL
    mo.neigh(d, t)
    s.neigh(gd, gt)
    assert rel_l2(gd.cpu().numpy(), d) <= FP32_TOL
    mo.step(d, t, dt, 5)
    s.step(gd, gt, dt, 5)
    got = gd.cpu().numpy()
    for r in range(16):
        assert rel_l2(got[..., r], d[..., r]) <= FP32_TOL


def test_mfma_raw_physical_solvers_are_factored(gpu):
    """Caller-given FP64 star matrices and flux solvers with physical
    structure take the sparse-star / rank-3 / projection path (factored on
    the device from FP64) and match the oracle; dense random ones do not."""
    m, rho, lam, mu = _problem(n=3, bc=(FA_PERIODIC, FA_FREE_SURFACE, FA_OUTFLOW), seed=8)
    star, fs, _ = orc.setup_mesh(m, rho, lam, mu)
    order = 5
    s = ElasticSolver(m, order, 16, 32, star=star, fs=fs)
    assert s.kernel_info() == (2, PATH_SPARSE_STAR | PATH_RANK3_FS | PATH_PROJECTIONS | PATH_FUSED)
    mo = orc.MeshOracle(m, order, 16, star=star, fs=fs)
    dofs0 = _ic(m, order, 16, seed=1)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    d, t = dofs0.copy(), np.zeros_like(dofs0)
    mo.step(d, t, dt, 5)
    gd, gt = _to_dev(dofs0, 32), _to_dev(np.zeros_like(dofs0), 32)
    s.step(gd, gt, dt, 5)
    got = gd.cpu().numpy()
    for r in range(16):
        assert rel_l2(got[..., r], d[..., r]) <= FP32_TOL
    rng = np.random.default_rng(2)
    s2 = ElasticSolver(m, order, 16, 32, star=rng.normal(size=star.shape), fs=rng.normal(size=fs.shape))
    assert s2.kernel_info() == (2, 0)
