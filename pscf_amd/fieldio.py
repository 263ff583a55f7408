"""Star-basis field files in the reference's format (src/field_io_mod.f90:55-135).

Layout written by output_field (:101-135): a `format 1 0` line, the unit cell
block (dim, crystal_system, N_cell_param, cell_param), group_name, N_monomer,
N_star, then one line per star: N_monomer coefficients, the star's
representative wave (dim integers) and star_count.  input_field (:55-85)
skips the 12 header lines after the version line, reads N_star and maps each
line's wave to its star, rescaling by sqrt(N_arm / star_count).
"""
import numpy as np


def _label(lines, name, value):
    lines.append("%-20s" % name)
    lines.append(value)


def write_field(path, field, system, basis, star_count, group_name=None, n_basis_out=None):
    """field: (N_monomer, N_star)."""
    nm, ns = field.shape
    n_out = ns if n_basis_out is None else min(n_basis_out, ns)
    L = ["format  1  0"]
    _label(L, "dim", "%20d" % system.dim)
    _label(L, "crystal_system", "%20s" % ("'" + system.crystal_system + "'"))
    _label(L, "N_cell_param", "%20d" % len(system.cell_param))
    _label(L, "cell_param", "".join("%20.10E" % c for c in system.cell_param))
    _label(L, "group_name", "%20s" % ("'" + (group_name or system.group_name) + "'"))
    _label(L, "N_monomer", "%20d" % nm)
    _label(L, "N_star", "%20d" % n_out)
    dim = basis.wave_of_star.shape[1]
    for i in range(n_out):
        vals = "".join("%25.16E" % field[a, i] for a in range(nm))
        wv = "".join("%4d" % basis.wave_of_star[i, d] for d in range(dim))
        L.append(vals + "    " + wv + "%6d" % star_count[i])
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def read_field(path, n_monomer, basis, star_count):
    """Read a field file onto this basis (input_field semantics); waves not in
    the basis are skipped (valid_wave), missing stars stay 0."""
    with open(path) as f:
        lines = f.read().splitlines()
    # version line + 12 header lines, then the N_star label and value
    n = int(lines[14].split()[0])
    which = {tuple(int(v) for v in w): i for i, w in enumerate(basis.wave)}
    field = np.zeros((n_monomer, basis.N_star))
    dim = basis.wave.shape[1]
    for line in lines[15:15 + n]:
        tok = line.split()
        vals = [float(t) for t in tok[:n_monomer]]
        G = tuple(int(t) for t in tok[n_monomer:n_monomer + dim])
        n_arm = int(tok[n_monomer + dim])
        w = which.get(G)
        if w is None:
            continue
        s = int(basis.star_of_wave[w]) - 1
        field[:, s] = np.array(vals) * np.sqrt(float(n_arm) / float(star_count[s]))
    return field
