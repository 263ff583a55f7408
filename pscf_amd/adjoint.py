"""Reverse-mode derivatives of the SCFT density map as a torch.autograd.Function.

`density(S, omega_grid, cell_param, block_lengths)` evaluates the reference's
density() (src/scf_mod.fp.f90:144-229) on real-space grids held in device
memory -- omega_grid (N_monomer, N3, N2, N1) -> rho_grid (same shape), Q per
chain -- and differentiates a scalar loss of (rho_grid, Q) with respect to
omega(r), the unit-cell parameters and the block lengths, through the
library's adjoint sweeps (pscf_density_adjoint_dev).  The derivatives are
those of the discrete map the forward computes (Richardson-extrapolated
split-step contour steps, Simpson contour weights, src/step_mod.f90:146-176,
src/scf_mod.fp.f90:362-424), so they agree with central finite differences of
the forward to rounding error.  Block-length derivatives hold the block's
step count fixed (chain_mod.f90:141: iblk is an integer function of the
length); the chain length in the density prefactor is the sum of the block
lengths (src/chemistry_mod.f90:125-129).

`free_energy` restates free_energy (src/scf_mod.fp.f90:834-886) on grid
fields with torch operations, so gradients of F flow through the same
Function (free_energy_stars sums over star coefficients as the reference
does).  Star-basis helpers map star coefficients to grids and back; their
backward passes use the symmetric-field adjoint (M times the star
coefficient), which is exact for stars without Nyquist-plane waves.

The library keeps only the contour slabs of the last forward evaluation of a
system: call backward before the next forward on the same system.
"""
import ctypes

import numpy as np
import torch

from ._lib import dptr
from .iterate import unit_cell


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _sync(t):
    if t.device.type == "cuda":
        torch.cuda.synchronize(t.device)


def block_lengths(system):
    """Block lengths of every chain, chain-major, as a float64 tensor."""
    return torch.tensor([L for c in system.chains for _, L in c], dtype=torch.float64)


def set_state(S, cell_param=None, lengths=None):
    """Move system S to the given cell parameters and block lengths (host
    values); the chain grids are rebuilt (density_startup) only when a block
    length changes."""
    sysm = S.system
    if cell_param is not None:
        cp = np.asarray(cell_param, dtype=np.float64).ravel()
        if not np.array_equal(cp, np.asarray(sysm.cell_param, dtype=np.float64)):
            sysm.cell_param = cp.copy()
            _, G, _ = unit_cell(S.lib, sysm.dim, sysm.crystal_system, cp)
            S.set_cell(G)
    if lengths is not None:
        Ls = np.asarray(lengths, dtype=np.float64).ravel()
        cur = np.array([L for c in sysm.chains for _, L in c])
        if not np.array_equal(Ls, cur):
            k = 0
            chains = []
            for c in sysm.chains:
                nc = []
                for m, _ in c:
                    nc.append((m, float(Ls[k])))
                    k += 1
                chains.append(nc)
            sysm.chains = chains
            phi, mu = S.phi_chain.copy(), S.mu_chain.copy()
            S.set_chemistry(sysm)
            S.phi_chain[:], S.mu_chain[:] = phi, mu
            S.startup(sysm.chain_step)


class DensityFunction(torch.autograd.Function):
    """(omega_grid, cell_param, block_lengths, phi_chain) -> (rho_grid, Q)."""

    @staticmethod
    def forward(ctx, omega_grid, cell_param, lengths, phi, S):
        set_state(S, cell_param.detach().cpu().numpy(), lengths.detach().cpu().numpy())
        S.phi_chain[: len(phi)] = phi.detach().cpu().numpy()
        s = S.system
        om = omega_grid.detach().to(torch.float64).contiguous()
        assert om.shape == (s.N_monomer,) + S.shape, (om.shape, S.shape)
        rho = torch.empty_like(om)
        q = np.zeros(max(s.N_chain, 1))
        qs = np.zeros(max(len(s.solvents), 1))
        _sync(om)
        S.lib.call("pscf_density_grid_dev", S.h, _vp(om), _vp(rho), dptr(q), dptr(qs),
                   dptr(S.phi_chain), dptr(S.mu_chain), dptr(S.phi_solvent), dptr(S.mu_solvent))
        S._density_gen = getattr(S, "_density_gen", 0) + 1
        ctx.S = S
        ctx.gen = S._density_gen
        ctx.ncp = len(s.cell_param)
        return rho, torch.tensor(q[: s.N_chain], dtype=torch.float64, device=om.device)

    @staticmethod
    def backward(ctx, g_rho, g_Q):
        S = ctx.S
        if getattr(S, "_density_gen", 0) != ctx.gen:
            raise RuntimeError("the system evaluated density() again after this forward; "
                               "its contour slabs are gone (call backward first)")
        s = S.system
        dev = torch.device(S.lib.tensor_device)
        lam = (torch.zeros((s.N_monomer,) + S.shape, dtype=torch.float64, device=dev)
               if g_rho is None else g_rho.detach().to(torch.float64).contiguous())
        dQ = (np.zeros(max(s.N_chain, 1)) if g_Q is None
              else np.ascontiguousarray(g_Q.detach().cpu().numpy(), dtype=np.float64))
        grad = torch.empty_like(lam)
        want_cell, want_len, want_phi = ctx.needs_input_grad[1:4]
        ncp = ctx.ncp
        dcell = np.zeros(max(ncp, 1))
        dlen = np.zeros(sum(len(c) for c in s.chains))
        dphi = np.zeros(max(s.N_chain, 1))
        dgg_arg = None
        if want_cell:
            _, _, dGG = unit_cell(S.lib, s.dim, s.crystal_system, s.cell_param)
            dgg_arg = np.ascontiguousarray(np.asarray(dGG[:, :, :ncp]).transpose(2, 1, 0))
        _sync(lam)
        S.lib.call("pscf_density_adjoint_dev", S.h, _vp(lam), dptr(dQ), _vp(grad), ncp,
                   dptr(dgg_arg) if want_cell else None, dptr(dcell) if want_cell else None,
                   dptr(dlen) if want_len else None, dptr(dphi) if want_phi else None)
        return (grad,
                torch.tensor(dcell[:ncp], dtype=torch.float64) if want_cell else None,
                torch.tensor(dlen, dtype=torch.float64) if want_len else None,
                torch.tensor(dphi[: s.N_chain], dtype=torch.float64) if want_phi else None,
                None)


def density(S, omega_grid, cell_param=None, lengths=None, phi=None):
    """rho_grid, Q of system S (an SCFSystem) at omega_grid; cell_param,
    lengths (block lengths, chain-major) and phi (chain volume fractions),
    float64 tensors, optional, make those differentiable inputs."""
    if cell_param is None:
        cell_param = torch.tensor(np.asarray(S.system.cell_param, dtype=np.float64))
    if lengths is None:
        lengths = block_lengths(S.system)
    if phi is None:
        phi = torch.tensor(np.asarray(S.phi_chain[: S.system.N_chain], dtype=np.float64))
    return DensityFunction.apply(omega_grid, cell_param, lengths, phi, S)


def free_energy(S, omega_grid, rho_grid, Q, lengths=None, phi=None):
    """Helmholtz free energy per monomer (src/scf_mod.fp.f90:834-886), canonical
    ensemble, with the star sums written as grid averages.  The two agree
    except for stars with waves on a Nyquist plane, whose grid functions are
    not unit-normalised (about 1e-7 relative at 8^3); free_energy_stars is
    the reference's sum term for term."""
    s = S.system
    if s.ensemble != 0:
        raise ValueError("free_energy: canonical ensemble only")
    if lengths is None:
        lengths = block_lengths(s)
    dev = rho_grid.device
    f = torch.zeros((), dtype=torch.float64, device=dev)
    k = 0
    for c, ch in enumerate(s.chains):
        N = lengths[k:k + len(ch)].sum().to(dev)
        k += len(ch)
        ph = phi[c].to(dev) if phi is not None else float(s.phi_chain[c])
        if float(ph) > 1.0e-8:
            mu = torch.log(ph / Q[c])
            f = f + ph * (mu - 1.0) / N
    if len(s.solvents):
        raise ValueError("free_energy: solvent species are not supported here")
    nm = s.N_monomer
    chi = np.asarray(s.chi, dtype=np.float64)
    inter = torch.zeros_like(rho_grid[0])
    for a in range(nm):
        for b in range(a + 1, nm):
            inter = inter + chi[a, b] * rho_grid[a] * rho_grid[b]
        inter = inter - omega_grid[a] * rho_grid[a]
    return f + inter.mean()


def free_energy_stars(S, omega_stars, rho_stars, Q, lengths=None, phi=None):
    """free_energy (src/scf_mod.fp.f90:834-886) in star coefficients, term for
    term as the reference sums it: (N_monomer, N_star) omega and rho."""
    s = S.system
    if s.ensemble != 0 or len(s.solvents):
        raise ValueError("free_energy_stars: canonical ensemble without solvents only")
    if lengths is None:
        lengths = block_lengths(s)
    f = torch.zeros((), dtype=torch.float64, device=rho_stars.device)
    k = 0
    for c, ch in enumerate(s.chains):
        N = lengths[k:k + len(ch)].sum().to(rho_stars.device)
        k += len(ch)
        ph = phi[c].to(rho_stars.device) if phi is not None else float(s.phi_chain[c])
        if float(ph) > 1.0e-8:
            f = f + ph * (torch.log(ph / Q[c]) - 1.0) / N
    chi = np.asarray(s.chi, dtype=np.float64)
    nm = s.N_monomer
    tot = torch.zeros_like(rho_stars[0])
    for a in range(nm):
        for b in range(a + 1, nm):
            tot = tot + rho_stars[a] * chi[a, b] * rho_stars[b]
        tot = tot - omega_stars[a] * rho_stars[a]
    return f + tot.sum()


class _StarsToGrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stars, S):
        nf = stars.shape[0]
        b = stars.detach().to(torch.float64).t().contiguous()       # (N_star, nf): Fortran (nf, N_star)
        out = torch.empty((nf,) + S.shape, dtype=torch.float64, device=b.device)
        _sync(b)
        S.lib.call("pscf_basis_to_grid_dev", S.h, _vp(b), nf, _vp(out))
        ctx.S = S
        return out

    @staticmethod
    def backward(ctx, g):
        # d/dc_star of sum_r g(r) f_star(r) = M * (star coefficient of g) for symmetric g
        return grid_to_stars(ctx.S, g.detach(), scale=1.0, differentiable=False), None


def stars_to_grid(S, stars):
    """Star coefficients (nf, N_star) -> grids (nf, N3, N2, N1) (basis_to_kgrid + ifft,
    src/scf_mod.fp.f90:182-185)."""
    return _StarsToGrid.apply(stars, S)


def grid_to_stars(S, grid, scale=None, differentiable=True):
    """Grids (nf, N3, N2, N1) -> star coefficients (nf, N_star), times `scale`
    (default 1/M: fft + kgrid_to_basis, src/scf_mod.fp.f90:221-225)."""
    M = int(np.prod(S.shape))
    if scale is None:
        scale = 1.0 / M
    if differentiable:
        return _GridToStars.apply(grid, S, scale)
    nf = grid.shape[0]
    g = grid.to(torch.float64).contiguous()
    out = torch.empty((S.N_star, nf), dtype=torch.float64, device=g.device)
    _sync(g)
    S.lib.call("pscf_grid_to_basis_dev", S.h, _vp(g), nf, float(scale), _vp(out))
    return out.t().contiguous()


class _GridToStars(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grid, S, scale):
        ctx.S, ctx.scale = S, scale
        return grid_to_stars(S, grid.detach(), scale, differentiable=False)

    @staticmethod
    def backward(ctx, g):
        # the coefficient map restricted to symmetric fields is (1/M) <f_star, .>,
        # whose adjoint is (scale) * sum_star g_star f_star
        return _StarsToGrid.apply(g.detach(), ctx.S) * ctx.scale, None, None
