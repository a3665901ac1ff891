"""Rate-2 local time stepping driver over the solver's LTS entry points: the
time groups of src/time/TimeGroupStatic.cpp (group g steps with 2^g times
the fundamental step) and an order of their local / neighbour updates that
meets the data dependencies of the reference's scheduler
(src/impl/seismic/inc/time/man_sched.inc, time::Manager::getTimePredAvailable /
getTimePredConsumed, src/time/Manager.cpp): a group's time prediction, then
two steps of the next finer group, then the group's neighbour update.

Elements of one group form a contiguous range, as in the reference's
element ordering by time group.
"""
import numpy as np

FA_FREE_SURFACE, FA_OUTFLOW = 101, 105


def characteristics(mesh, group):
    """Per-element and per-face LTS characteristics from a time group per
    element: el_lts [nEl] (bit 0: an adjacent element has a larger step, bit 1:
    a smaller one), fa_lts [nEl][4] (0 equal, 1 the element's step is the
    greater, 2 the smaller); boundary faces count as equal."""
    group = np.asarray(group)
    ne = np.asarray(mesh.el_fa_el)
    ty = np.asarray(mesh.fa_type)[np.asarray(mesh.el_fa)]
    own = (ty == FA_FREE_SURFACE) | (ty == FA_OUTFLOW) | (ne < 0)
    gn = np.where(own, group[:, None], group[np.where(ne < 0, 0, ne)])
    g = group[:, None]
    fa = np.where(gn == g, 0, np.where(g > gn, 1, 2)).astype(np.uint8)
    el = ((fa == 2).any(axis=1).astype(np.uint8) | ((fa == 1).any(axis=1).astype(np.uint8) << 1)).astype(np.uint8)
    return el, fa


def groups_from_cfl(dt_el, n_groups):
    """Time group per element from its stable step (tools/edge_v/src/time/
    Groups.cpp): group g holds steps in [dt0 2^g, dt0 2^(g+1)), the last
    group everything above; dt0 is the smallest step."""
    dt_el = np.asarray(dt_el, dtype=np.float64)
    dt0 = float(dt_el.min())
    g = np.floor(np.log2(dt_el / dt0) + 1e-12).astype(int)
    return np.minimum(g, n_groups - 1), dt0


class Rate2Lts:
    """Advances a solver's DOFs with rate-2 local time stepping.

    ranges: (first, n) per time group, finest first, contiguous and ascending;
    dt0: the finest group's step."""

    def __init__(self, solver, ranges, dt0):
        import torch
        self.solver = solver
        self.ranges = [(int(a), int(b)) for a, b in ranges]
        pos = 0
        for a, b in self.ranges:
            if a != pos or b <= 0:
                raise ValueError("time groups must be contiguous, ascending element ranges covering the mesh")
            pos = a + b
        if pos != solver.mesh.n_el:
            raise ValueError("time groups must cover the mesh")
        group = np.concatenate([np.full(b, g) for g, (a, b) in enumerate(self.ranges)])
        self.el_lts, self.fa_lts = characteristics(solver.mesh, group)
        solver.set_lts(self.el_lts, self.fa_lts)
        self.dt0 = float(dt0)
        self.t = [torch.zeros(solver.dof_shape(), dtype=solver.dtype, device="cuda") for _ in range(3)]
        self.counts = [0] * len(self.ranges)

    def _step(self, dofs, g, stream, dofs_a):
        ft = self.counts[g] % 2 == 0
        first, n = self.ranges[g]
        self.solver.local_lts(dofs, *self.t, self.dt0 * 2 ** g, ft, first, n, stream=stream, dofs_a=dofs_a)
        if g > 0:
            self._step(dofs, g - 1, stream, dofs_a)
            self._step(dofs, g - 1, stream, dofs_a)
        self.solver.neigh_lts(dofs, *self.t, ft, first, n, stream=stream, dofs_a=dofs_a)
        self.counts[g] += 1

    def cycle(self, dofs, n_cycles=1, stream=None, dofs_a=None):
        """n_cycles steps of the coarsest group (2^(G-1) steps of the finest);
        dofs_a: the anelastic DOFs of a viscoelastic solver."""
        for _ in range(n_cycles):
            self._step(dofs, len(self.ranges) - 1, stream, dofs_a)
