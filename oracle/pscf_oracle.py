"""Oracle for the SCFT inner loop -- TEST INFRASTRUCTURE ONLY.

A plain CPU (numpy, FP64) restatement of the reference PSCF inner loop.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker: the product path
(``pscf_amd``) never calls it and fails loudly when its HIP library is absent.

Every function cites the reference routine it restates (paths relative to the
reference repository root).  Arrays on the spatial grid are numpy arrays of
shape ``(N3, N2, N1)`` in C order, which is the same memory layout as the
reference's Fortran ``(0:N1-1, 0:N2-1, 0:N3-1)`` column-major arrays (x index
fastest).  The half spectrum has shape ``(N3, N2, N1//2+1)``, matching
``qk(0:N1/2, 0:N2-1, 0:N3-1)`` (src/step_mod.f90:89).

How each function is pinned (tests/test_oracle.py):
  * chain_grid            -- integer rule, hand-worked step counts, and the
                             reference's own discretisation via pscf_dump dumps.
  * make_propg/step_exp   -- closed forms (uniform field: e^{-w ds}; single
                             plane wave at w=0: e^{-ds b^2 |G|^2/6}) and the
                             reference step_exp dumped by oracle/_ref/pscf_dump.
  * fft/ifft              -- brute-force DFT on a tiny grid and the reference
                             (FFTW3 API through MKL) dumps.
  * chain_density/density -- homogeneous-field closed forms (Q = e^{-sum w L},
                             rho = phi L/N) and reference density() dumps.
  * scf_stress            -- zero for a homogeneous field, reference dumps,
                             and central finite differences of the free energy.
  * free_energy & co      -- Flory-Huggins closed form and reference dumps.
  * basis_to_kgrid /
    kgrid_to_basis        -- reference dumps and the round-trip identity.
  * make_ksq / G_to_bz    -- reference ksq and G_to_bz dumps (bit-exact).
"""
import numpy as np

# ----------------------------------------------------------------------------
# chain discretisation                                (src/chain_mod.f90:97-175)
# ----------------------------------------------------------------------------

def chain_grid(blk_length, ds0, order=None):
    """Return (block_bgn, block_ds) exactly as make_chain_grid computes them.

    block_bgn is 1-based, length nblk+1 (src/chain_mod.f90:137-156)."""
    block_bgn = [1]
    block_ds = []
    bgns = 1
    for L in blk_length:
        iblk = int(L / ds0 / 2.0 + 0.5)           # chain_mod.f90:141
        if iblk == 0:                              # chain_mod.f90:143-145
            iblk = 1
            bds = L / 2.0
        else:
            bds = L / float(iblk) / 2.0            # chain_mod.f90:147
        if order is not None:                      # chain_mod.f90:150-153
            bds = bds / 2.0 ** order
            iblk = iblk * 2 ** order
        bgns = bgns + iblk * 2                     # chain_mod.f90:155
        block_ds.append(bds)
        block_bgn.append(bgns)
    return block_bgn, block_ds


# ----------------------------------------------------------------------------
# FFT conventions                                     (src/fft3_mod.f90:85-122)
# ----------------------------------------------------------------------------

def fft(r):
    """Unnormalised r2c, sign -1, x (last numpy axis) halved (fft3_mod.f90:85-94)."""
    return np.fft.rfftn(r)


def ifft(k, shape):
    """Unnormalised c2r, sign +1 (fft3_mod.f90:113-122)."""
    return np.fft.irfftn(k, s=shape, axes=tuple(range(len(shape)))) * float(np.prod(shape))


# ----------------------------------------------------------------------------
# wavevector geometry                                 (src/grid_mod.f90:141-413)
# ----------------------------------------------------------------------------

def norm(G, G_basis):
    """|G . G_basis|^2 with the reference's summation order (grid_mod.f90:402-413,
    group_mod.f90:768-780, 659-667)."""
    dim = G_basis.shape[0]
    gvec = [0.0] * dim
    for j in range(dim):
        s = 0.0
        for i in range(dim):
            s = s + G[i] * G_basis[i, j]
        gvec[j] = s
    out = 0.0
    for j in range(dim):
        out = out + gvec[j] * gvec[j]
    return out


def G_to_bz(G, ngrid, G_basis):
    """Minimum-image wavevector with the reference's tie rule: trial shifts in
    order +N, 0, -N per dimension, accepted only if Gsq < Gsq_min - 1e-8
    (grid_mod.f90:225-283)."""
    dim = G_basis.shape[0]
    G_min = [0, 0, 0]
    Gsq_min = 1.0e10
    shifts = (1, 0, -1)
    import itertools
    for t in itertools.product(shifts, repeat=dim):
        G_try = [G[i] + t[i] * ngrid[i] for i in range(dim)]
        Gsq = norm(G_try, G_basis)
        if Gsq < Gsq_min - 1.0e-8:
            G_min[:dim] = G_try
            Gsq_min = Gsq
    return G_min


def g_to_bz_grid(ngrid, G_basis):
    """Vectorised G_to_bz + norm over every half-grid point (grid_mod.f90:141-160,
    225-283).  Returns (Gbz, ksq): Gbz has shape (N3, N2, N1//2+1, 3) (unused
    dimensions 0), ksq has shape (N3, N2, N1//2+1)."""
    import itertools
    N1, N2, N3 = ngrid
    dim = G_basis.shape[0]
    i3, i2, i1 = np.meshgrid(np.arange(N3), np.arange(N2), np.arange(N1 // 2 + 1),
                             indexing="ij")
    G = [i1, i2, i3][:dim]
    best = np.full(i1.shape, 1.0e10)
    Gbz = np.zeros(i1.shape + (3,), dtype=np.int64)
    for t in itertools.product((1, 0, -1), repeat=dim):
        Gt = [G[i] + t[i] * ngrid[i] for i in range(dim)]
        gvec = []
        for j in range(dim):
            acc = np.zeros(i1.shape)
            for i in range(dim):
                acc = acc + Gt[i] * G_basis[i, j]
            gvec.append(acc)
        gsq = np.zeros(i1.shape)
        for j in range(dim):
            gsq = gsq + gvec[j] * gvec[j]
        take = gsq < best - 1.0e-8
        best = np.where(take, gsq, best)
        for i in range(dim):
            Gbz[..., i] = np.where(take, Gt[i], Gbz[..., i])
    return Gbz, best


def make_ksq(ngrid, G_basis):
    """ksq_grid on the half grid, shape (N3, N2, N1//2+1) (grid_mod.f90:141-160)."""
    return g_to_bz_grid(ngrid, G_basis)[1]


# ----------------------------------------------------------------------------
# star basis <-> k grid                           (src/grid_basis_mod.f90:47-200)
# ----------------------------------------------------------------------------

class Tables:
    """Star-basis tables produced by basis_mod (src/basis_mod.f90:60-84),
    as dumped by oracle/pscf_dump (DUMP_TABLES)."""

    def __init__(self, d, G_basis=None):
        self.dim = int(d["dim"])
        self.ngrid = [int(x) for x in d["ngrid"]]
        self.wave = np.asarray(d["wave"])                # (N_wave, dim)
        self.coeff = np.asarray(d["coeff"])              # (N_wave,)
        self.star_of_wave = np.asarray(d["star_of_wave"])  # 1-based
        self.star_invert = np.asarray(d["star_invert"])
        self.wave_of_star = np.asarray(d["wave_of_star"])  # (N_star, dim)
        self.N_star = int(self.star_invert.shape[0])
        self.N_wave = int(self.coeff.shape[0])
        self._which = {tuple(int(v) for v in w): i + 1 for i, w in enumerate(self.wave)}
        if G_basis is None:
            self.kmap = np.asarray(d["kmap"])            # (N3,N2,N1/2+1,4)
        else:
            self.kmap = self.make_kmap(G_basis)

    def make_kmap(self, G_basis):
        """For every half-grid k: (G_to_bz(k), which_wave(G_to_bz(k))), as the
        loop of basis_to_kgrid evaluates them (grid_basis_mod.f90:76-78)."""
        Gbz, _ = g_to_bz_grid(self.ngrid, G_basis)
        km = np.zeros(Gbz.shape[:3] + (4,), dtype=np.int64)
        km[..., :3] = Gbz
        flat = Gbz.reshape(-1, 3)
        km.reshape(-1, 4)[:, 3] = [self.which_wave(g) for g in flat]
        return km

    def which_wave(self, kbz):
        return self._which.get(tuple(int(v) for v in kbz[:self.dim]), 0)


def basis_to_kgrid(basis, T):
    """Star coefficients -> half k grid (grid_basis_mod.f90:47-129)."""
    N1, N2, N3 = T.ngrid
    sq2 = np.sqrt(2.0)
    out = np.zeros((N3, N2, N1 // 2 + 1), dtype=complex)
    iw = T.kmap[..., 3]
    for i3 in range(N3):
        for i2 in range(N2):
            for i1 in range(N1 // 2 + 1):
                w = int(iw[i3, i2, i1])
                if w < 1 or w > T.N_wave:
                    continue
                c = T.coeff[w - 1]
                s = int(T.star_of_wave[w - 1])          # 1-based
                inv = int(T.star_invert[s - 1])
                if inv == 0:
                    z = basis[s - 1] * c
                elif inv == 1:
                    z = complex(basis[s - 1], -basis[s]) * c / sq2
                elif inv == -1:
                    z = complex(basis[s - 2], basis[s - 1]) * c / sq2
                else:
                    raise ValueError("Illegal star_invert value in basis_to_kgrid.")
                out[i3, i2, i1] = z
    return out


def kgrid_to_basis(kgrid, T):
    """Half k grid -> star coefficients, reading one representative wave per
    star (grid_basis_mod.f90:146-200).  No 1/M: the caller divides."""
    N1, N2, N3 = T.ngrid
    sq2 = np.sqrt(2.0)
    basis = np.zeros(T.N_star)
    for s in range(T.N_star):
        kbz = [0, 0, 0]
        kbz[:T.dim] = [int(v) for v in T.wave_of_star[s]]
        c = T.coeff[T.which_wave(kbz) - 1]
        kfft = [kbz[i] % T.ngrid[i] if i < T.dim else 0 for i in range(3)]
        if kfft[0] > N1 // 2:
            kneg = [(-kbz[i]) % T.ngrid[i] if i < T.dim else 0 for i in range(3)]
            z = np.conj(kgrid[kneg[2], kneg[1], kneg[0]])
        else:
            z = kgrid[kfft[2], kfft[1], kfft[0]]
        inv = int(T.star_invert[s])
        if inv == 0:
            basis[s] = (z / c).real
        elif inv == 1:
            basis[s] = (z / c).real * sq2
        elif inv == -1:
            basis[s] = (z / c).imag * sq2
        else:
            raise ValueError("Illegal star_invert value in grid_to_basis.")
    return basis


# ----------------------------------------------------------------------------
# propagator                                         (src/step_mod.f90:109-176)
# ----------------------------------------------------------------------------

class Propagator:
    """make_propg + step_exp (step_mod.f90:109-129, 146-176)."""

    def __init__(self, ngrid):
        self.shape = (ngrid[2], ngrid[1], ngrid[0])
        self.M = float(ngrid[0] * ngrid[1] * ngrid[2])

    def make_propg(self, ds, b, ksq, omega):
        lap_coeff = ds * b * b / 6.0                     # step_mod.f90:120
        pot_coeff = ds / 2.0                             # step_mod.f90:121
        self.exp_ksq1 = np.exp(-lap_coeff * ksq)         # step_mod.f90:123
        self.exp_ksq2 = np.exp(-lap_coeff / 2.0 * ksq)   # step_mod.f90:124
        self.exp_omega1 = np.exp(-pot_coeff * omega)     # step_mod.f90:126
        self.exp_omega2 = np.exp(-pot_coeff / 2.0 * omega)  # step_mod.f90:127

    def step_exp(self, q_in):
        M, s = self.M, self.shape
        qr = self.exp_omega1 * q_in                      # step_mod.f90:156-160
        qk = fft(qr) * self.exp_ksq1
        qr = ifft(qk, s)
        q1 = self.exp_omega1 * qr / M
        qr = self.exp_omega2 * q_in                      # step_mod.f90:162-166
        qk = fft(qr) * self.exp_ksq2
        qr = ifft(qk, s)
        qr = self.exp_omega1 * qr / M                    # step_mod.f90:167-171
        qk = fft(qr) * self.exp_ksq2
        qr = ifft(qk, s)
        q2 = self.exp_omega2 * qr / M
        return (4.0 * q2 - q1) / 3.0                     # step_mod.f90:174


# ----------------------------------------------------------------------------
# chain / density / stress / thermodynamics     (src/scf_mod.fp.f90:144-941)
# ----------------------------------------------------------------------------

class Chemistry:
    """The subset of chemistry_mod globals the hot path reads
    (src/chemistry_mod.f90:48-84): per chain a list of (monomer, length)
    blocks (1-based monomer ids), phi_chain, kuhn lengths, chi matrix."""

    def __init__(self, kuhn, chains, phi_chain, chi, ensemble=0, mu_chain=None,
                 solvents=(), phi_solvent=(), mu_solvent=None):
        self.solvents = [tuple(s) for s in solvents]     # (monomer, size)
        self.phi_solvent = list(phi_solvent)
        self.mu_solvent = list(mu_solvent) if mu_solvent is not None else [0.0] * len(solvents)
        self.N_solvent = len(self.solvents)
        self.kuhn = list(kuhn)
        self.chains = [list(c) for c in chains]
        self.phi_chain = list(phi_chain)
        self.chi = np.asarray(chi, dtype=float)
        self.ensemble = ensemble
        self.mu_chain = list(mu_chain) if mu_chain is not None else [0.0] * len(chains)
        self.N_monomer = len(kuhn)
        self.N_chain = len(chains)

    def chain_length(self, i):
        s = 0.0                                          # chemistry_mod.f90:560-566
        for _, L in self.chains[i]:
            s = s + L
        return s


class ChainState:
    def __init__(self, block_bgn, block_ds):
        self.block_bgn = block_bgn
        self.block_ds = block_ds
        self.qf = None
        self.qr = None
        self.rho = None
        self.bigQ = None


def chain_density(i_chain, chain, chem, ksq, omega_grid, prop):
    """Forward/backward contour sweeps, Q and Simpson rho per block
    (scf_mod.fp.f90:315-426).  omega_grid[alpha-1] is the grid field of
    monomer alpha."""
    blocks = chem.chains[i_chain]
    nblk = len(blocks)
    bgn = chain.block_bgn
    ns1 = bgn[-1]                               # number of stored points
    shape = omega_grid.shape[1:]
    qf = np.empty((ns1,) + shape)
    qr = np.empty((ns1,) + shape)
    qf[0] = 1.0                                          # scf_mod.fp.f90:331
    for ib in range(nblk):                               # :332-342
        mon, _ = blocks[ib]
        prop.make_propg(chain.block_ds[ib], chem.kuhn[mon - 1], ksq, omega_grid[mon - 1])
        for istep in range(bgn[ib], bgn[ib + 1]):
            qf[istep] = prop.step_exp(qf[istep - 1])
    qr[ns1 - 1] = 1.0                                    # :345-346
    for ib in range(nblk - 1, -1, -1):                   # :347-356
        mon, _ = blocks[ib]
        prop.make_propg(chain.block_ds[ib], chem.kuhn[mon - 1], ksq, omega_grid[mon - 1])
        for istep in range(bgn[ib + 1], bgn[ib], -1):
            qr[istep - 2] = prop.step_exp(qr[istep - 1])
    chain.bigQ = qf[ns1 - 1].sum() / float(qf[ns1 - 1].size)   # :359-360
    rho = np.zeros((nblk,) + shape)
    for ib in range(nblk):                               # :364-421
        ibgn, iend = bgn[ib], bgn[ib + 1]
        r = qf[ibgn - 1] * qr[ibgn - 1]
        r = r + qf[iend - 1] * qr[iend - 1]
        for j in range(ibgn + 1, iend, 2):
            r = r + qf[j - 1] * qr[j - 1] * 4.0
        for j in range(ibgn + 2, iend - 1, 2):
            r = r + qf[j - 1] * qr[j - 1] * 2.0
        rho[ib] = r * chain.block_ds[ib] / 3.0
    rho = rho / chem.chain_length(i_chain) / chain.bigQ   # :424
    chain.qf, chain.qr, chain.rho = qf, qr, rho
    return chain


def mu_phi_solvent(chem, qs):
    """scf_mod.fp.f90:803-820."""
    if chem.ensemble == 0:
        chem.mu_solvent = [float(np.log(chem.phi_solvent[i] / qs[i])) for i in range(chem.N_solvent)]
    else:
        chem.phi_solvent = [float(qs[i] * np.exp(chem.mu_solvent[i])) for i in range(chem.N_solvent)]


def solvent_density(chem, i_solv, omega_grid, rho_grid):
    """Boltzmann-weight solvent density (scf_mod.fp.f90:251-295)."""
    mon, size = chem.solvents[i_solv]
    w = np.exp((-size) * omega_grid[mon - 1])
    Qs = w.sum() / float(w.size)
    if chem.ensemble == 1:
        chem.phi_solvent[i_solv] = Qs * np.exp(chem.mu_solvent[i_solv])
    rho_grid[mon - 1] = rho_grid[mon - 1] + chem.phi_solvent[i_solv] * w / Qs
    return Qs


def mu_phi_chain(chem, q):
    """scf_mod.fp.f90:768-785."""
    if chem.ensemble == 0:
        chem.mu_chain = [float(np.log(chem.phi_chain[i] / q[i])) for i in range(chem.N_chain)]
    else:
        chem.phi_chain = [float(q[i] * np.exp(chem.mu_chain[i])) for i in range(chem.N_chain)]


class SCF:
    """density / scf_stress / free_energy over one system
    (scf_mod.fp.f90:82-229, 451-550, 834-886)."""

    def __init__(self, chem, T, ksq, chain_step):
        self.chem, self.T, self.ksq = chem, T, ksq
        self.ngrid = T.ngrid
        self.shape = (T.ngrid[2], T.ngrid[1], T.ngrid[0])
        self.prop = Propagator(T.ngrid)
        self.chains = []
        for i in range(chem.N_chain):                    # density_startup :99-106
            bgn, bds = chain_grid([L for _, L in chem.chains[i]], chain_step)
            self.chains.append(ChainState(bgn, bds))

    def density(self, omega):
        """omega: (N_monomer, N_star).  Returns (rho (N_monomer,N_star), q)."""
        chem, T = self.chem, self.T
        M = float(np.prod(self.shape))
        omega_grid = np.empty((chem.N_monomer,) + self.shape)
        for a in range(chem.N_monomer):                  # :182-185 (no 1/M)
            omega_grid[a] = ifft(basis_to_kgrid(omega[a], T), self.shape)
        q = []
        for i in range(chem.N_chain):                    # :188-191
            chain_density(i, self.chains[i], chem, self.ksq, omega_grid, self.prop)
            q.append(self.chains[i].bigQ)
        rho_grid = np.zeros((chem.N_monomer,) + self.shape)
        qs = []
        for i in range(chem.N_solvent):                  # :193-200
            qs.append(solvent_density(chem, i, omega_grid, rho_grid))
        self.q_solvent = np.array(qs)
        if chem.ensemble == 1:                           # :202-209
            mu_phi_chain(chem, q)
            mu_phi_solvent(chem, qs)
        for i in range(chem.N_chain):                    # :212-218
            for ib, (mon, _) in enumerate(chem.chains[i]):
                rho_grid[mon - 1] = rho_grid[mon - 1] + chem.phi_chain[i] * self.chains[i].rho[ib]
        rho = np.empty((chem.N_monomer, T.N_star))
        for a in range(chem.N_monomer):                  # :221-225
            rho[a] = kgrid_to_basis(fft(rho_grid[a]), T) / M
        self.omega_grid, self.rho_grid = omega_grid, rho_grid
        return rho, np.array(q)

    def scf_stress(self, dGsq):
        """dGsq: (N_star, N_cell_param).  Requires a previous density()
        (scf_mod.fp.f90:451-550)."""
        chem, T = self.chem, self.T
        ncp = dGsq.shape[1]
        M = float(np.prod(self.shape))
        normal = M * M * 3.0 * 6.0                       # :485-488
        stress = np.zeros(ncp)
        for sp in range(chem.N_chain):
            ch = self.chains[sp]
            dQ = np.zeros(ncp)
            for ib, (mon, _) in enumerate(chem.chains[sp]):
                b = chem.kuhn[mon - 1]
                ds0 = ch.block_ds[ib]
                ibgn, iend = ch.block_bgn[ib], ch.block_bgn[ib + 1]
                for i in range(ibgn, iend + 1):          # :505-530
                    qf_b = kgrid_to_basis(fft(ch.qf[i - 1]), T)
                    qr_b = kgrid_to_basis(fft(ch.qr[i - 1]), T)
                    ds = ds0
                    if i != ibgn and i != iend:
                        ds = 4.0 * ds if i % 2 == 0 else 2.0 * ds
                    for beta in range(ncp):
                        inc = float(np.dot(qr_b * dGsq[:, beta], qf_b))
                        inc = inc * b ** 2 * ds / normal
                        dQ[beta] = dQ[beta] - inc
            if chem.ensemble == 0:                       # :536-544
                stress = stress - (dQ / ch.bigQ) * chem.phi_chain[sp] / chem.chain_length(sp)
            else:
                stress = stress - (dQ / ch.bigQ) * np.exp(chem.mu_chain[sp]) * ch.bigQ \
                    / chem.chain_length(sp)
        return stress


def free_energy(chem, rho, omega):
    """Helmholtz free energy per monomer and pressure (scf_mod.fp.f90:834-886)."""
    f = 0.0
    for i in range(chem.N_chain):
        if chem.phi_chain[i] > 1.0e-8:
            f = f + chem.phi_chain[i] * (chem.mu_chain[i] - 1.0) / chem.chain_length(i)
    for i in range(chem.N_solvent):
        if chem.phi_solvent[i] > 1.0e-8:
            f = f + chem.phi_solvent[i] * (chem.mu_solvent[i] - 1.0) / chem.solvents[i][1]
    N = rho.shape[1]
    nm = chem.N_monomer
    for i in range(N):
        for a in range(nm):
            for b in range(a + 1, nm):
                f = f + rho[a, i] * chem.chi[a, b] * rho[b, i]
            f = f - omega[a, i] * rho[a, i]
    p = -f
    for i in range(chem.N_chain):
        if chem.phi_chain[i] > 1.0e-8:
            p = p + chem.mu_chain[i] * chem.phi_chain[i] / chem.chain_length(i)
    for i in range(chem.N_solvent):
        if chem.phi_solvent[i] > 1.0e-8:
            p = p + chem.mu_solvent[i] * chem.phi_solvent[i] / chem.solvents[i][1]
    return f, p


def free_energy_FH(chem):
    """Flory-Huggins free energy of the homogeneous mixture (scf_mod.fp.f90:903-941)."""
    f = 0.0
    rho = np.zeros(chem.N_monomer)
    for i in range(chem.N_chain):
        L = chem.chain_length(i)
        if chem.phi_chain[i] > 1.0e-8:
            f = f + (chem.phi_chain[i] / L) * (np.log(chem.phi_chain[i]) - 1)
        for mon, bl in chem.chains[i]:
            rho[mon - 1] = rho[mon - 1] + chem.phi_chain[i] * bl / L
    for i in range(chem.N_solvent):
        mon, size = chem.solvents[i]
        if chem.phi_solvent[i] > 1.0e-8:
            f = f + (chem.phi_solvent[i] / size) * (np.log(chem.phi_solvent[i]) - 1)
        rho[mon - 1] = rho[mon - 1] + chem.phi_solvent[i]
    for i in range(chem.N_monomer - 1):
        for j in range(i + 1, chem.N_monomer):
            f = f + chem.chi[i, j] * rho[i] * rho[j]
    return f


def set_omega_uniform(chem, omega):
    """omega(:,1) = chi . phi_mon (scf_mod.fp.f90:724-750)."""
    phi_mon = np.zeros(chem.N_monomer)
    for i in range(chem.N_chain):
        L = chem.chain_length(i)
        for mon, bl in chem.chains[i]:
            phi_mon[mon - 1] = phi_mon[mon - 1] + chem.phi_chain[i] * bl / L
    for i in range(chem.N_solvent):
        phi_mon[chem.solvents[i][0] - 1] += chem.phi_solvent[i]
    for a in range(chem.N_monomer):
        s = 0.0
        for b in range(chem.N_monomer):
            s = s + chem.chi[a, b] * phi_mon[b]
        omega[a, 0] = s
    return omega


def make_dGsq(T, dGG):
    """dGsq(star) = sum_jk G_j G_k dGG(j,k) over wave_of_star (basis_mod.f90:1993-2014)."""
    out = np.zeros(T.N_star)
    for s in range(T.N_star):
        G = [float(v) for v in T.wave_of_star[s]]
        acc = 0.0
        for j in range(T.dim):
            for k in range(T.dim):
                acc = acc + G[j] * G[k] * dGG[j, k]
        out[s] = acc
    return out
