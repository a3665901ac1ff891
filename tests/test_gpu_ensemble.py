"""An ensemble of meshes advanced as one: the disjoint union of several
viscoelastic meshes (uneven sizes, own materials and quality factors) in one
solver, stepped through a captured HIP graph, against each member stepped on
its own.  Element updates read only the element and its face neighbours, so
the union's results equal the members' bit for bit; each graph replay is
checked after eager work of the member solvers in between."""
import numpy as np
import pytest

from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd._lib import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import KERNEL_MFMA, ViscoSolver

pytestmark = pytest.mark.gpu


def _members():
    specs = [(3, (FA_PERIODIC,) * 3, 0.1), ((2, 3, 2), (FA_OUTFLOW, FA_PERIODIC, FA_FREE_SURFACE), 0.0),
             (4, (FA_PERIODIC,) * 3, 0.05)]
    out = []
    for i, (n, bc, jit) in enumerate(specs):
        m = M.box(n, bc=bc, jitter=jit, seed=10 + i)
        rho, lam, mu = S.random_materials(m.n_el, seed=20 + i)
        qp = np.random.default_rng(30 + i).uniform(50, 200, m.n_el)
        out.append((m, rho, lam, mu, qp))
    return out


@pytest.mark.parametrize("order", [4, 5])
def test_union_graph_steps_equal_members(gpu, order):
    import torch
    runs, nr = 8, 3
    members = _members()
    u, ranges = M.union([m for m, *_ in members])
    cat = [np.concatenate([x[k] for x in members]) for k in range(1, 5)]
    su = ViscoSolver(u, order, runs, 32, cat[0], cat[1], cat[2], cat[3], cat[3] / 2, nr, 1.0, 10.0,
                     kernel=KERNEL_MFMA)
    ss = [ViscoSolver(m, order, runs, 32, rho, lam, mu, qp, qp / 2, nr, 1.0, 10.0, kernel=KERNEL_MFMA)
          for m, rho, lam, mu, qp in members]
    setup_u = su.setup_a()
    dt = min(S.dt_cfl(m, rho, s.setup_a()["lam_e"], s.setup_a()["mu_e"], order)
             for (m, rho, *_), s in zip(members, ss))
    gen = torch.Generator(device="cuda").manual_seed(5)
    du = torch.randn(su.dof_shape(), generator=gen, device="cuda")
    au = 0.1 * torch.randn(su.dof_a_shape(), generator=gen, device="cuda")
    tu = torch.zeros_like(du)
    dm = [du[a:a + n].clone() for a, n in ranges]
    am = [au[a:a + n].clone() for a, n in ranges]
    tm = [torch.zeros_like(d) for d in dm]
    for (a, n), s in zip(ranges, ss):  # the members' set-up is the union's, element by element
        np.testing.assert_array_equal(s.setup_a()["srcA"], setup_u["srcA"][a:a + n])
    for rep in range(3):
        su.step(du, tu, dt, 1, au, graph=True)
        for s, d, t, a in zip(ss, dm, tm, am):
            s.step(d, t, dt, 1, a)
        torch.cuda.synchronize()
        for (a, n), d, x in zip(ranges, dm, am):
            assert torch.equal(du[a:a + n], d), (rep, a)
            assert torch.equal(au[a:a + n], x), (rep, a)
        assert torch.isfinite(du).all()
    su.close()
    for s in ss:
        s.close()
