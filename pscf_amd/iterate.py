"""SCFT iteration over the device hot path: the control flow of the
reference's iterate_AM (src/iterate_mod.fp.f90:1042-1490) driving
SCFSystem.density / scf_stress, with the Anderson update, the unit-cell
geometry (make_unit_cell, src/unit_cell_mod.f90:309-664) and make_dGsq
(src/basis_mod.f90:1993-2014) from libpscf_hip's host-side FP64 code.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from ._lib import dptr, iptr


def unit_cell(lib, dim, crystal_system, cell_param):
    """(R_basis, G_basis, dGG_basis) with numpy index (i, j[, k]) = Fortran (i+1, j+1[, k+1])."""
    cp = np.ascontiguousarray(cell_param, dtype=np.float64)
    R = np.zeros(dim * dim)
    G = np.zeros(dim * dim)
    dGG = np.zeros(dim * dim * 6)
    lib.call("pscf_unit_cell", int(dim), crystal_system.encode(), len(cp), dptr(cp), dptr(R),
             dptr(G), dptr(dGG))
    return (R.reshape(dim, dim, order="F"), G.reshape(dim, dim, order="F"),
            dGG.reshape(dim, dim, 6, order="F"))


def make_dgsq(lib, basis, dGG, ncp):
    dim = basis.wave_of_star.shape[1]
    out = np.zeros((basis.N_star, ncp))
    for k in range(ncp):
        d = np.ascontiguousarray(dGG[:, :, k].T)      # Fortran (dim, dim) layout
        col = np.zeros(basis.N_star)
        lib.call("pscf_make_dgsq", int(dim), basis.N_star, iptr(basis.wave_of_star), dptr(d),
                 dptr(col))
        out[:, k] = col
    return out


@dataclass
class IterResult:
    omega: np.ndarray
    rho: np.ndarray
    itr: int
    converge: bool
    error: float
    f_helmholtz: float
    pressure: float
    stress: np.ndarray
    cell_param: np.ndarray
    q: np.ndarray
    history: List[tuple] = field(default_factory=list)   # (itr, error, f_helmholtz)


class AndersonMixing:
    """iterate_AM_startup + iterate_AM over an SCFSystem."""

    def __init__(self, scfsys, n_hist, domain):
        self.S = scfsys
        self.lib = scfsys.lib
        s = scfsys.system
        self.ncp = len(s.cell_param)
        self.domain = bool(domain)
        h = ctypes.c_int64(0)
        self.lib.call("pscf_am_create", scfsys.N_star, s.N_monomer, self.ncp, int(n_hist),
                      int(self.domain), ctypes.byref(h))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", 0):
            self.lib.dll.pscf_am_destroy(self.h)
            self.h = 0

    def run(self, omega, error_max, max_itr, log=None):
        S, lib, sysm = self.S, self.lib, self.S.system
        cell = np.array(sysm.cell_param, dtype=np.float64)
        chi = np.ascontiguousarray(np.asarray(sysm.chi, dtype=np.float64).T)
        omega = np.array(omega, dtype=np.float64)
        # initial density and mu (:1123-1128)
        rho, q, qs = S.density(omega)
        S.mu_phi_chain(q)
        S.mu_phi_solvent(qs)
        _, _, dGG = unit_cell(lib, sysm.dim, sysm.crystal_system, cell)
        dGsq = make_dgsq(lib, S.basis, dGG, self.ncp)
        if sysm.ensemble == 0:
            omega = S.set_omega_uniform(omega)
        om_f = np.ascontiguousarray(omega.T)                 # Fortran (N_monomer, N_star)
        lib.call("pscf_am_start", self.h, dptr(om_f), dptr(cell))
        stress = np.zeros(self.ncp)
        err = ctypes.c_double(0.0)
        status = ctypes.c_int(0)
        hist = []
        itr = 0
        while True:
            itr += 1
            rho, q, qs = S.density(om_f.T)
            if self.domain:
                stress = S.scf_stress(dGsq)
            S.mu_phi_chain(q)
            S.mu_phi_solvent(qs)
            f, p = S.free_energy(rho, om_f.T)
            rho_f = np.ascontiguousarray(rho.T)
            lib.call("pscf_am_iterate", self.h, dptr(om_f), dptr(rho_f), dptr(chi), dptr(stress),
                     dptr(cell), float(error_max), int(max_itr), ctypes.byref(err),
                     ctypes.byref(status))
            hist.append((itr, err.value, f))
            if log:
                log(itr, err.value, f, stress, cell)
            if status.value != 0:
                break
            if self.domain:
                _, G, dGG = unit_cell(lib, sysm.dim, sysm.crystal_system, cell)
                S.set_cell(G)
                dGsq = make_dgsq(lib, S.basis, dGG, self.ncp)
        if not self.domain:
            stress = S.scf_stress(dGsq)
        return IterResult(omega=om_f.T.copy(), rho=rho, itr=itr, converge=status.value == 1,
                          error=err.value, f_helmholtz=f, pressure=p, stress=stress,
                          cell_param=cell, q=q, history=hist)
