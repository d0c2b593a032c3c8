"""Device initial conditions and power spectrum against the oracle (which the golden vectors pin:
2LPT at 1e-12 through /ics, ranlxd1 through the runtests.jl known answers)."""
import numpy as np
import pytest

import golden as G
import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")

COSMO = (0.2621, 0.049, 0.6766, 0.8102, 0.9665)  # Omega_c, Omega_b, h, sigma8, n_s of the workloads


def _oracle_ics(n, L, white, a, order=2):
    oc = O.Cosmology(*COSMO)
    k, pk = O.linear_power_spectrum(oc)
    fphi = O.white_noise_to_fphi(O.ic_operator(n, L, k, pk), white)
    psi1, psi2 = O.compute_core2(fphi, L)
    if order == 1:
        psi2 = np.zeros_like(psi2)
    return O.nbody_ics_lpt2(oc, psi1, psi2, a), (k, pk), oc.Dplus(a)


@pytest.mark.parametrize("n", [16, 32, 64])
def test_ics_from_white_noise_fp64(n):
    L, a = 2.0 * n, 0.05
    white = O.ngenic_wnoise_real(42, n)
    (psi_ref, pi_ref), (k, pk), Dg = _oracle_ics(n, L, white, a)
    with D.PMSolverHIP(D.PMConfig(res_pm=max(2 * n, 32), n_part=n, boxsize=L), precision=64) as S:
        S.ics_lpt(k, pk, Dg, order=2, white=white)
        psi, pi = S.get_state()
    assert G.rel(psi, psi_ref) < 1e-12
    assert G.rel(pi, pi_ref) < 1e-12


def test_ics_from_seed_and_order1():
    n, L, a = 32, 100.0, 0.05
    white = O.ngenic_wnoise_real(42, n)
    (psi_ref, pi_ref), (k, pk), Dg = _oracle_ics(n, L, white, a)
    with D.PMSolverHIP(D.PMConfig(res_pm=64, n_part=n, boxsize=L), precision=64) as S:
        S.ics_lpt(k, pk, Dg, order=2, seed=42)
        psi, pi = S.get_state()
        assert G.rel(psi, psi_ref) < 1e-12 and G.rel(pi, pi_ref) < 1e-12
        (p1_ref, q1_ref), _, _ = _oracle_ics(n, L, white, a, order=1)
        S.ics_lpt(k, pk, Dg, order=1, seed=42)
        psi1, pi1 = S.get_state()
        assert G.rel(psi1, p1_ref) < 1e-12 and G.rel(pi1, q1_ref) < 1e-12
    with D.PMSolverHIP(D.PMConfig(res_pm=64, n_part=n, boxsize=L), precision=32) as S:
        S.ics_lpt(k, pk, Dg, order=2, seed=42)
        psi32, _ = S.get_state()
    assert psi32.dtype == np.float32 and G.rel(psi32, psi_ref) < 1e-6


@pytest.mark.parametrize("prec", [32, 64])
def test_power_spectrum_vs_oracle(prec):
    n, N, L = 32, 64, 200.0
    rng = np.random.default_rng(3)
    psi = rng.normal(scale=3.0, size=(3, n, n, n))
    with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=n, boxsize=L), precision=prec) as S:
        k, pk, nm = S.power_spectrum(psi.astype(S.dtype), W=2)
    mesh = O.pm_scatter(O.lagrangian_positions(psi.astype(np.float32 if prec == 32 else np.float64), L)
                        .astype(np.float64), N, L, 2)
    delta = mesh * (N**3 / n**3) - 1
    ko, pko, nmo = O.power_spectrum_lin(delta, L, W=2, nbins=N // 4 - 1)
    assert np.array_equal(nm, nmo)
    assert np.allclose(k, ko, rtol=1e-14)
    tol = 1e-10 if prec == 64 else 2e-4
    assert np.max(np.abs(pk / pko - 1)) < tol
