"""Batched parameter sweeps (pscf_amd/sweep.py) against the reference
program's own SWEEP command (oracle/_ref/pscf, src/pscf.fp.f90:557-698) on
the lamellar known-answer case (tools/python/tests/outfile/in: f = 0.5,
chi N = 15, 32 points), and batched paths against the same paths run alone."""
import copy
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, get_lib, load_golden

sys.path.insert(0, GOLDEN)
from make_golden import cases  # noqa: E402

from pscf_amd import batch, sweep  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "pscf")


def _converged_start(lib, error_max=1e-10):
    g = load_golden("lam32")
    s = cases()["lam32"][0]
    B = batch.BatchSolver(n_hist=10, domain=False, error_max=error_max, max_itr=300, lib=lib)
    B.add(batch.make_system(s, g, lib=lib), load_golden("lam32_start")["omega"])
    B.solve()
    r = B.result(0)
    assert r.converge
    return s, g, r.omega


def _path(s, g, om, d_chi, s_max):
    dc = np.array([[0.0, d_chi], [d_chi, 0.0]])
    return sweep.SweepPath(system=copy.deepcopy(s), tables=g, omega=om,
                           incr=sweep.Increments(d_chi=dc), s_max=s_max)


@pytest.mark.gpu
def test_sweep_matches_reference_program_gpu(tmp_path):
    """chi N 15 -> 17 (d_chi = 1, s_max = 2): the reference's schedule
    (s = 0.1, 1, 2 -- as the reference computes them, from single-precision
    step literals) and its free energies and pressures at s = 1, 2, all
    printed digits.  Both programs iterate to residual 5e-12."""
    if not os.path.exists(REF):
        pytest.skip("reference program not built (make -C oracle)")
    import test_dropin as T
    T._write_case(tmp_path, max_itr=300)
    p = tmp_path / "param"
    p.write_text(p.read_text().replace("1.0E-10", "5.0E-12").replace(
        "FINISH\n", "SWEEP\ns_max\n2.0\nd_chi\n1.0\nend_increments\n\nFINISH\n"))
    T._run(REF, str(tmp_path))
    lib = get_lib("gpu")
    s, g, om = _converged_start(lib, 5e-12)
    path = _path(s, g, om, 1.0, 2.0)
    sweep.run([path], domain=False, n_hist=10, error_max=5e-12, max_itr=300, lib=lib)
    assert path.status == "done"
    f01, f09 = float(np.float32(0.1)), float(np.float32(0.9))
    assert [q.s for q in path.points] == [f01, f01 + f09, f01 + f09 + 1.0]
    for k in (1, 2):
        ref = T._summary(str(tmp_path / "out" / ("%d.out" % k)))
        mine = [q for q in path.points if abs(q.s - round(q.s)) < 1e-3 and round(q.s) == k][0]
        # the summary prints 10 decimals
        assert abs(mine.f_helmholtz - ref["f_Helmholtz"]) < 2e-10, (k, mine.f_helmholtz, ref)
        assert abs(mine.pressure - ref["pressure"]) < 2e-10, (k, mine.pressure, ref)


@pytest.mark.gpu
def test_batched_paths_equal_single_paths_gpu():
    """Three lamellar paths (d_chi = 1, 0.5, -1; s_max 2, 3, 2) in one batch
    against each path alone: the same schedule and the same points."""
    lib = get_lib("gpu")
    s, g, om = _converged_start(lib)
    spec = [(1.0, 2.0), (0.5, 3.0), (-1.0, 2.0)]
    together = [_path(s, g, om, d, m) for d, m in spec]
    sweep.run(together, domain=False, n_hist=10, error_max=1e-10, max_itr=300, lib=lib)
    for (d, m), pt in zip(spec, together):
        alone = _path(s, g, om, d, m)
        sweep.run([alone], domain=False, n_hist=10, error_max=1e-10, max_itr=300, lib=lib)
        assert pt.status == alone.status == "done"
        assert [q.s for q in pt.points] == [q.s for q in alone.points]
        for a, b in zip(pt.points, alone.points):
            assert a.itr == b.itr
            assert abs(a.f_helmholtz - b.f_helmholtz) <= 1e-12 * abs(b.f_helmholtz)
            np.testing.assert_allclose(a.omega, b.omega, rtol=0, atol=1e-11 * np.abs(b.omega).max())
