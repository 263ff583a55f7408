"""Batched parameter sweeps: many SWEEP paths advanced side by side.

The reference's SWEEP command (src/pscf.fp.f90:557-698) moves one system
along a line in parameter space: s goes from 0 to s_max, every parameter
changes by step * (its increment) (increment_parameters,
src/sweep_mod.f90), the next omega (and, with domain = T, the cell) is
predicted by polynomial continuation through the last three converged points
(continuation, :541-597; update_history), and the SCF equations are solved
from there.  The step schedule is 0.1, then 0.9 (or whatever reaches s = 1
after a failure), then 1; a failed solve backtracks and halves the step
unit, down to 1/16.

Here each path keeps exactly that state machine, while the solves of all
paths at one round go into one batched Anderson solve (pscf_batch_*: per
instance the reference's iterate_AM; instances on one grid share every
contour-step launch).  Paths finish (s >= s_max) or fail independently;
the others go on.
"""
import copy
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import batch
from .iterate import unit_cell

N_HIST = 3    # src/sweep_mod.f90:97

# The reference writes the step-schedule constants as default-real
# (single-precision) literals (src/pscf.fp.f90:590-697): the steps are
# 0.1f and 0.9f times the step unit, so after the first two points s is
# 0.99999997765..., not 1, and every swept parameter carries that offset.
# The same values here keep the path's parameters equal to the reference's.
_F01 = float(np.float32(0.1))
_F09 = float(np.float32(0.9))
_F0001 = float(np.float32(0.001))
_BACKTRACK_MIN = float(np.float32(np.float32(1.0 / 16.0) + np.float32(0.001)))


@dataclass
class Increments:
    """Per unit of s (src/sweep_mod.f90:100-300); None: unchanged.
    d_chi is the full symmetric matrix; d_block_length one list per chain."""
    d_kuhn: Optional[list] = None
    d_chi: Optional[np.ndarray] = None
    d_block_length: Optional[list] = None
    d_phi_chain: Optional[list] = None
    d_phi_solvent: Optional[list] = None
    d_mu_chain: Optional[list] = None
    d_mu_solvent: Optional[list] = None
    d_solvent_size: Optional[list] = None
    d_cell_param: Optional[list] = None


@dataclass
class SweepPoint:
    s: float
    itr: int
    f_helmholtz: float
    pressure: float
    cell_param: np.ndarray
    stress: np.ndarray
    omega: np.ndarray
    error: float


@dataclass
class SweepPath:
    """A converged starting state (s = 0: system, its star tables and omega
    stars) and the line to follow."""
    system: object
    tables: dict
    omega: np.ndarray
    incr: Increments
    s_max: float
    points: List[SweepPoint] = field(default_factory=list)
    status: str = "running"        # "done", or "failed" (no convergence at step unit 1/16)


class _State:
    def __init__(self, path, domain):
        self.p = path
        self.sys = copy.deepcopy(path.system)
        self.sys.chi = np.array(self.sys.chi, dtype=float)
        self.cell = np.array(self.sys.cell_param, dtype=float)
        self.domain = domain
        self.s = 0.0
        self.i = 0
        self.step_unit = 1.0
        self.step = 0.0
        self.converge = True
        self.s_hist, self.om_hist, self.cell_hist = [], [], []
        self.update_history(0.0, np.array(path.omega, dtype=float), self.cell)
        self.scf = None

    def update_history(self, s, omega, cell):
        self.s_hist.insert(0, s)
        self.om_hist.insert(0, omega.copy())
        self.cell_hist.insert(0, cell.copy())
        del self.s_hist[N_HIST:], self.om_hist[N_HIST:], self.cell_hist[N_HIST:]

    def next_step(self):
        """src/pscf.fp.f90:582-593."""
        if self.i == 0:
            return _F01 * self.step_unit
        if self.i == 1:
            return _F09 * self.step_unit if self.converge else self.step_unit - self.s
        return self.step_unit

    def increment(self, step):
        """increment_parameters (src/sweep_mod.f90): every parameter += step * d."""
        d, y = self.p.incr, self.sys
        if d.d_kuhn is not None:
            y.kuhn = [k + step * dk for k, dk in zip(y.kuhn, d.d_kuhn)]
        if d.d_chi is not None:
            y.chi = y.chi + step * np.asarray(d.d_chi, dtype=float)
        if d.d_block_length is not None:
            y.chains = [[(m, L + step * dl) for (m, L), dl in zip(c, dc)]
                        for c, dc in zip(y.chains, d.d_block_length)]
        if y.ensemble == 0:
            if d.d_phi_chain is not None and y.N_chain > 0:
                y.phi_chain = [p + step * dp for p, dp in zip(y.phi_chain, d.d_phi_chain)]
            if d.d_phi_solvent is not None and len(y.solvents) > 0:
                y.phi_solvent = [p + step * dp for p, dp in zip(y.phi_solvent, d.d_phi_solvent)]
        else:
            if d.d_mu_chain is not None and y.N_chain > 0:
                y.mu_chain = [m + step * dm for m, dm in zip(y.mu_chain, d.d_mu_chain)]
        if d.d_solvent_size is not None and len(y.solvents) > 0:
            y.solvents = [(m, z + step * dz) for (m, z), dz in zip(y.solvents, d.d_solvent_size)]
        if not self.domain and d.d_cell_param is not None:
            self.cell = self.cell + step * np.asarray(d.d_cell_param, dtype=float)

    def continuation(self, step):
        """Zeroth, first or second order prediction from the history
        (src/sweep_mod.f90:541-597): omega, and the cell when domain."""
        n = len(self.s_hist)
        h, w = self.om_hist, self.cell_hist
        if n > 2:
            ds2 = self.s_hist[1] - self.s_hist[0]
            ds3 = self.s_hist[2] - self.s_hist[0]
            A2 = step * (ds3 / ds2) / (ds3 - ds2)
            A3 = step * (ds2 / ds3) / (ds2 - ds3)
            B2 = (step ** 2) / (ds2 * (ds2 - ds3))
            B3 = (step ** 2) / (ds3 * (ds3 - ds2))
            C2 = A2 + B2
            C3 = A3 + B3
            C1 = 1.0 - C2 - C3
            om = C1 * h[0] + C2 * h[1] + C3 * h[2]
            cell = C1 * w[0] + C2 * w[1] + C3 * w[2]
        elif n == 2:
            ds2 = self.s_hist[1] - self.s_hist[0]
            C2 = step / ds2
            C1 = 1.0 - C2
            om = C1 * h[0] + C2 * h[1]
            cell = C1 * w[0] + C2 * w[1]
        else:
            om = h[0].copy()
            cell = w[0].copy()
        if self.domain:
            self.cell = cell
        return om


def run(paths, domain=False, n_hist=20, error_max=1e-8, max_itr=500, lib=None, log=None):
    """Advance every path to its s_max.  Fills path.points (every converged
    point) and path.status; returns the number of batched solves."""
    from ._lib import lib as _default_lib
    lib = lib or _default_lib()
    states = [_State(p, domain) for p in paths]
    rounds = 0
    while True:
        live = [st for st in states if st.p.status == "running"]
        if not live:
            return rounds
        solver = batch.BatchSolver(n_hist=n_hist, domain=domain, error_max=error_max,
                                   max_itr=max_itr, lib=lib)
        for st in live:
            st.step = st.next_step()
            st.s = st.s + st.step
            st.increment(st.step)
            om = st.continuation(st.step)
            st.sys.cell_param = [float(c) for c in st.cell]
            if st.scf is None:
                st.scf = batch.make_system(st.sys, st.p.tables, lib=lib)
            else:
                _, G, _ = unit_cell(lib, st.sys.dim, st.sys.crystal_system, st.sys.cell_param)
                st.scf.system = st.sys
                st.scf.set_cell(G)
                st.scf.set_chemistry(st.sys)
                st.scf.startup(st.sys.chain_step)
            solver.add(st.scf, om)
        solver.solve()
        rounds += 1
        for k, st in enumerate(live):
            r = solver.result(k)
            st.converge = r.converge
            if r.converge:
                st.i += 1
                if domain:
                    st.cell = np.array(r.cell_param, dtype=float)
                st.update_history(st.s, r.omega, st.cell)
                st.p.points.append(SweepPoint(s=st.s, itr=r.itr, f_helmholtz=r.f_helmholtz,
                                              pressure=r.pressure, cell_param=st.cell.copy(),
                                              stress=r.stress.copy(), omega=r.omega.copy(),
                                              error=r.error))
                if log:
                    log("path %d s=%.4f itr=%d f=%.10f" % (states.index(st), st.s, r.itr, r.f_helmholtz))
                if st.s + _F0001 >= st.p.s_max:
                    st.p.status = "done"
            elif st.step_unit > _BACKTRACK_MIN:
                st.s = st.s - st.step
                st.increment(-st.step)
                st.step_unit = 0.5 * st.step_unit
            else:
                st.p.status = "failed"
