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


# ---------------------------------------------------- field, cross and binned spectra (power_spectrum.jl)
def _fields(N, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(N, N, N))
    # a correlated second field: f smoothed along x plus noise, so the cross spectrum has both signs
    g = 0.6 * np.roll(f, 1, axis=2) - 0.3 * f + 0.5 * rng.normal(size=(N, N, N))
    return f, g


@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("log", [True, False])
@pytest.mark.parametrize("prec", [32, 64])
def test_field_power_spectrum_vs_oracle(N, log, prec):
    """evaluate_power_spectrum of a field: the same modes in every bin (bit-exact |k| and bins), P and the
    mean |k| of each bin against the oracle (FP64 1e-10; FP32: the field and its transform in FP32)."""
    L = 3.0 * N
    f, _ = _fields(N, N + int(log))
    dt = np.float32 if prec == 32 else np.float64
    with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=16, boxsize=L), precision=prec) as S:
        k, pk, nm = S.field_power_spectrum(f.astype(dt), bins=20, logarithmic=log)
    ko, pko, nmo = O.power_spectrum(f.astype(dt).astype(np.float64), L, bins=20, logarithmic=log)
    assert np.array_equal(nm, nmo)
    assert np.allclose(k, ko, rtol=1e-12, atol=0)
    tol = 1e-10 if prec == 64 else 2e-5
    ok = nmo > 0
    assert np.max(np.abs(pk[ok] / pko[ok] - 1)) < tol
    assert np.all(pk[~ok] == 0)


@pytest.mark.parametrize("log", [True, False])
@pytest.mark.parametrize("prec", [32, 64])
def test_field_cross_power_spectrum_vs_oracle(log, prec):
    N, L = 48, 150.0
    f, g = _fields(N, 7)
    dt = np.float32 if prec == 32 else np.float64
    with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=16, boxsize=L), precision=prec) as S:
        k, pc = S.field_cross_power_spectrum(f.astype(dt), g.astype(dt), bins=16, logarithmic=log)
        _, pa = S.field_cross_power_spectrum(f.astype(dt), f.astype(dt), bins=16, logarithmic=log)
    ko, pco = O.cross_power_spectrum(f.astype(dt).astype(np.float64), g.astype(dt).astype(np.float64), L, bins=16,
                                     logarithmic=log)
    assert np.allclose(k, ko, rtol=1e-12, atol=0)
    scale = np.max(np.abs(pco))
    tol = 1e-10 if prec == 64 else 2e-5
    assert np.max(np.abs(pc - pco)) < tol * scale
    assert np.any(pco < 0) and np.any(pco > 0)
    _, pao = O.cross_power_spectrum(f.astype(dt).astype(np.float64), f.astype(dt).astype(np.float64), L, bins=16,
                                    logarithmic=log)
    assert np.max(np.abs(pa - pao)) < tol * np.max(pao)


@pytest.mark.parametrize("n, N", [(32, 64), (32, 32), (24, 48)])
@pytest.mark.parametrize("prec", [32, 64])
def test_power_spectra_binned_vs_oracle(n, N, prec):
    """pk_full of the study estimator (measure.jl:403-447): P_delta of the CIC density of q + psi on the
    mesh, window-deconvolved, and P_Psi of the displacement on the particle grid, over logarithmic edges
    from 0.999 kf to k_Nyq of the mesh, against the oracle."""
    L = 2.0 * n
    rng = np.random.default_rng(n + N)
    psi = rng.normal(scale=1.5, size=(3, n, n, n))
    dt = np.float32 if prec == 32 else np.float64
    psi_t = psi.astype(dt)
    edges = np.exp(np.linspace(np.log(2 * np.pi / L * 0.999), np.log(np.pi * N / L), 21))
    with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=n, boxsize=L), precision=prec) as S:
        k, pd, pp = S.power_spectra_binned(edges, psi_t)
        S.set_state(psi_t)
        _, pd2, _ = S.power_spectra_binned(edges, None, displacement=False)
    mesh = O.pm_scatter(O.lagrangian_positions(psi_t.astype(np.float64), L), N, L, 2)
    delta = mesh * (N**3 / n**3) - 1.0
    pdo, ppo = O.pk_binned(L, edges, delta=delta, psi=psi_t.astype(np.float64), deconv_W=2)
    assert np.allclose(k, np.sqrt(edges[:-1] * edges[1:]), rtol=1e-15, atol=0)
    tol = 1e-10 if prec == 64 else 5e-5
    okd, okp = pdo > 0, ppo > 0
    assert np.max(np.abs(pd[okd] / pdo[okd] - 1)) < tol and np.all(pd[~okd] == 0)
    assert np.max(np.abs(pp[okp] / ppo[okp] - 1)) < tol and np.all(pp[~okp] == 0)
    assert np.allclose(pd, pd2, rtol=1e-12, atol=0)  # the bin sums are double atomics (order-dependent)


def test_spectra_reject_bad_arguments():
    with D.PMSolverHIP(D.PMConfig(res_pm=32, n_part=16, boxsize=100.0), precision=64) as S:
        f = np.zeros((32, 32, 32))
        with pytest.raises(D.DJError):
            S.field_power_spectrum(f, bins=257)
        with pytest.raises(D.DJError):
            S.power_spectra_binned(np.array([1.0, 0.5, 2.0]))


@pytest.mark.parametrize("seed, res", [(42, 8), (7, 16), (123, 24), (42, 64), (1000, 128)])
def test_ngenic_field_on_the_device(seed, res):
    """The N-GenIC field generated on the device against the host's (ngenic_field_gsl): the same zero
    pattern (Nyquist planes, DC, the Hermitian kz = 0 plane), values within 2 ulp (the device's sqrt, log,
    cos and sin against the host libm's; the ranlxd1 uniforms underneath are integer-exact), and the
    reference's known answer |field(42, 8)| = 13.448527 (runtests.jl:417)."""
    a = D.ngenic_field_on_device(seed, res)
    b = D.ngenic_field(seed, res)
    assert np.array_equal(a == 0, b == 0)
    assert np.max(np.abs(a - b)) <= 2 * np.finfo(np.float64).eps * np.max(np.abs(b))
    if (seed, res) == (42, 8):
        assert np.linalg.norm(a) == pytest.approx(G.KNOWN["ngenic_field_gsl_42_8_norm"], abs=G.KNOWN["ngenic_field_gsl_42_8_norm_atol"])
