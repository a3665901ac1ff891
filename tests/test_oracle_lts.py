"""Pins the oracle's rate-2 local time stepping (the tDofs[1] / tDofs[2]
buffers of AderDg::local, AderDg.hpp:384-417, and the per-face neighbour
data of AderDg::neigh, :620-674):

* with every element in one time group the LTS loops are the GTS loops;
* a P plane wave advanced with a fine half and a coarse half of the mesh
  (twice the step) stays as close to the exact solution as GTS at the fine
  step, and the error falls at the scheme's order when the mesh is refined.
"""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S


def _rate2(o, dofs, bufs, dt0, groups, n_cycles):
    """Advances every group by n_cycles steps of the coarsest group:
    local(g), two steps of group g - 1, neigh(g) (recursively)."""
    counts = [0] * len(groups)

    def step(g):
        ft = counts[g] % 2 == 0
        first, n = groups[g]
        o.local_lts(dofs, *bufs, dt0 * 2 ** g, ft, first, n)
        if g > 0:
            step(g - 1)
            step(g - 1)
        o.neigh_lts(dofs, *bufs, ft, first, n)
        counts[g] += 1

    for _ in range(n_cycles):
        step(len(groups) - 1)


def _problem(n, order):
    m = M.box(n)
    ne = m.n_el
    rho, lam, mu = np.ones(ne), 2 * np.ones(ne), np.ones(ne)
    o = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
    pw = S.PlaneWave(k=(2 * np.pi, 2 * np.pi, 0.0))
    return m, o, pw, S.project(m, order, pw), S.dt_cfl(m, rho, lam, mu, order)


def test_one_group_is_gts():
    m, o, pw, q0, dt = _problem(3, 3)
    o.set_lts(*orc.lts_groups(m, np.zeros(m.n_el, dtype=int)))
    a, b = q0.copy(), q0.copy()
    ta = np.zeros_like(a)
    bufs = [np.zeros_like(a) for _ in range(3)]
    for s in range(3):
        o.step(a, ta, dt)
        o.local_lts(b, *bufs, dt, s % 2 == 0)
        o.neigh_lts(b, *bufs, s % 2 == 0)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("order", [3, 4])
def test_two_groups_follow_the_plane_wave(order):
    errs = []
    for n in (4, 8):
        m, o, pw, q0, dt = _problem(n, order)
        half = m.n_el // 2
        group = np.r_[np.zeros(half, dtype=int), np.ones(m.n_el - half, dtype=int)]
        el_lts, fa_lts = orc.lts_groups(m, group)
        assert (el_lts & 1).any() and (el_lts & 2).any()
        o.set_lts(el_lts, fa_lts)
        dt = dt / 2  # the coarse group runs at the stable step
        T = 0.05
        cycles = int(round(T / (2 * dt)))
        lts = q0.copy()
        bufs = [np.zeros_like(q0) for _ in range(3)]
        _rate2(o, lts, bufs, dt, [(0, half), (half, m.n_el - half)], cycles)
        gts, t = q0.copy(), np.zeros_like(q0)
        o.step(gts, t, dt, 2 * cycles)
        ex = S.project(m, order, lambda x: pw(x, 2 * cycles * dt))
        e_lts = np.linalg.norm(lts - ex) / np.linalg.norm(ex)
        e_gts = np.linalg.norm(gts - ex) / np.linalg.norm(ex)
        assert e_lts < 1.2 * e_gts, (n, e_lts, e_gts)
        errs.append(e_lts)
    assert np.log2(errs[0] / errs[1]) > order - 0.6, errs


def test_product_characteristics_match_the_oracle():
    from edge_amd.lts import characteristics
    for bc in ((106, 106, 106), (101, 105, 106)):
        m = M.box(4, bc=bc)
        g = np.random.default_rng(1).integers(0, 3, m.n_el)
        for a, b in zip(characteristics(m, g), orc.lts_groups(m, g)):
            np.testing.assert_array_equal(a, b)
