"""Inverse design of ABA + A blends toward a target lamellar period
(pscf_amd/design.py): the 1D lamellar star tables, the implicit gradient
dsigma/d(f_A, phi_h) from grouped adjoint passes against central finite
differences of converged solves, and the descent driving the stress at the
target period to zero.  (No reference counterpart: the reference has no
derivatives; the finite differences pin the gradient.)"""
import numpy as np
import pytest

from conftest import get_lib, load_golden

from pscf_amd import design


def test_lamellar_tables_match_reference_basis():
    """The constructed tables equal the reference's make_basis output for the
    1D lamellar group -1 on 32 points (tests/golden/lam32.npz); the
    reference's 1/sqrt(2) coefficients carry single-precision rounding."""
    g = load_golden("lam32")
    t = design.lamellar_tables(32)
    for k in ("wave", "star_of_wave", "star_invert", "wave_of_star", "star_count"):
        assert np.array_equal(np.asarray(g[k]).ravel(), np.asarray(t[k]).ravel()), k
    assert np.abs(g["coeff"] - t["coeff"]).max() < 5e-8


def _start(chi, nstar):
    om = np.zeros((2, nstar))
    om[0, 1] = chi
    om[1, 1] = -chi
    return om


@pytest.mark.gpu
def test_stress_gradient_matches_fd_gpu():
    """dsigma/dtheta from the grouped adjoint (checkpoint stride 7) against
    central differences (step 1e-3) of sigma from converged solves, 32-point
    lamellae, N = 100, ds = 0.5."""
    lib = get_lib("gpu")
    t = design.lamellar_tables(32)
    b = design.Blend(f_A=0.45, phi_h=0.1, chi_N=20.0)
    om = _start(0.2, 17)
    sig, dsig, omn = design.stress_and_gradient([b], [15.0], [(om, om)], t, 32, 0.5, stride=7,
                                                error_max=1e-10, lib=lib)
    for j in range(2):
        vals = []
        for sgn in (1.0, -1.0):
            bb = design.Blend(f_A=0.45 + (sgn * 1e-3 if j == 0 else 0.0),
                              phi_h=0.1 + (sgn * 1e-3 if j == 1 else 0.0), chi_N=20.0)
            s2, _, _ = design.stress_and_gradient([bb], [15.0], omn, t, 32, 0.5, stride=7,
                                                  error_max=1e-10, lib=lib)
            vals.append(s2[0])
        fd = (vals[0] - vals[1]) / 2e-3
        assert abs(dsig[0, j] - fd) <= 2e-3 * abs(fd), (j, dsig[0, j], fd)


@pytest.mark.gpu
def test_fit_spacing_drives_stress_to_zero_gpu():
    """Four blends, target periods 12.5-15.5: three descent steps cut |sigma(d_t)| by
    more than 20x for every instance."""
    lib = get_lib("gpu")
    t = design.lamellar_tables(32)
    rng = np.random.default_rng(4)
    blends = [design.Blend(f_A=float(rng.uniform(0.35, 0.5)), phi_h=float(rng.uniform(0.05, 0.2)),
                           chi_N=float(rng.uniform(15.0, 25.0))) for _ in range(4)]
    target = [12.5, 13.5, 14.5, 15.5]
    om = [(_start(0.2, 17), _start(0.2, 17)) for _ in blends]
    final, hist = design.fit_spacing(blends, target, om, t, 32, 0.5, iters=3, lib=lib)
    s0, s1 = np.abs(hist[0]["sigma"]), np.abs(hist[-1]["sigma"])
    assert np.all(s1 < s0 / 20.0), (s0, s1)
