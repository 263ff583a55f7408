"""Starting fields of the lamellar known-answer case (tests/golden/lam32.npz
tables; f = 0.5, chi N = 15, 32 points), written by this library's own
Anderson solve on the CPU emulator (inputs for the drop-in tests, not
reference outputs):

  lam32_start.npz  the solution to residual 1e-6, every star times
                   (1 + 0.05 N(0,1)) (seed 7) except the first, rounded to
                   6 decimals: a start from which both programs take the same
                   non-chaotic Anderson path (the committed file came from an
                   earlier build of the library; rerun, the recipe moves its
                   6th decimals, so it is rewritten only with --start);
  lam32_near.npz   the solution to residual 1e-10, stars past the first times
                   (1 + 0.01 N(0,1)) (seed 12): a start for well-conditioned
                   Newton steps.

    python tests/golden/make_starts.py [--start]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

from conftest import get_lib, load_golden  # noqa: E402
from make_golden import cases  # noqa: E402
from pscf_amd import batch, iterate, scf  # noqa: E402


def main():
    lib = get_lib("emu")
    g = load_golden("lam32")
    sysm = cases()["lam32"][0]
    S = scf.SCFSystem(sysm, scf.Basis.from_tables(g), g["G_basis"], lib=lib)
    om = np.zeros((2, S.N_star))
    om[0, 1], om[1, 1] = 3.0, -3.0
    r = iterate.AndersonMixing(S, n_hist=10, domain=False).run(om, error_max=1e-6, max_itr=300)
    rng = np.random.default_rng(7)
    start = r.omega * (1 + 0.05 * rng.normal(size=r.omega.shape))
    start[:, 0] = r.omega[:, 0]
    start = np.round(start, 6)
    if "--start" in sys.argv:
        np.savez_compressed(os.path.join(HERE, "lam32_start.npz"), omega=start)
    else:
        start = load_golden("lam32_start")["omega"]
    B = batch.BatchSolver(n_hist=10, domain=False, error_max=1e-10, max_itr=300, lib=lib)
    B.add(batch.make_system(sysm, g, lib=lib), start)
    B.solve()
    near = B.result(0).omega
    rng = np.random.default_rng(12)
    near[:, 1:] *= 1.0 + 0.01 * rng.standard_normal(near[:, 1:].shape)
    np.savez_compressed(os.path.join(HERE, "lam32_near.npz"), omega=near)


if __name__ == "__main__":
    main()
