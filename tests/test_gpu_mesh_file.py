"""A mesh read from a Gmsh file drives the GPU solver: two FP64 steps of a
jittered box with free-surface and outflow sides, written to MSH 4.1 and
read back, against the FP64 oracle on the same mesh (relative L2 <= 1e-12)."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import gmsh
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.solver import ElasticSolver

pytestmark = pytest.mark.gpu


def test_solver_on_a_gmsh_mesh(gpu, tmp_path):
    import torch
    box = M.box((4, 3, 3), bc=(M.FA_FREE_SURFACE, M.FA_OUTFLOW, M.FA_FREE_SURFACE), jitter=0.1, seed=21)
    p = tmp_path / "box.msh"
    gmsh.write(str(p), box)
    m = gmsh.read(str(p))
    rho, lam, mu = S.random_materials(m.n_el, seed=21)
    s = ElasticSolver(m, 4, 1, 64, rho, lam, mu)
    q0 = S.project(m, 4, S.PlaneWave(k=(2 * np.pi, 0.0, 2 * np.pi)))
    dt = S.dt_cfl(m, rho, lam, mu, 4)
    d = torch.tensor(q0, device="cuda")
    t = torch.zeros_like(d)
    s.step(d, t, dt, 2)
    o = orc.MeshOracle(m, 4, rho=rho, lam=lam, mu=mu)
    ref, tr = q0.copy(), np.zeros_like(q0)
    o.step(ref, tr, dt, 2)
    err = np.linalg.norm(d.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= 1e-12, err
    s.close()
