"""Pins of the oracle that the golden vectors leave open: known answers, closed forms, invariants.

Each check is chosen so that a plausible slip (dropped term, wrong sign or index, transposed axis)
fails it; none restates the formula under test.
"""
import numpy as np
import pytest

import golden as G
import oracle as O


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 30, 48, 64, 96, 7 * 11])
def test_fft1_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    m = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(m, m) / n) @ x
    assert np.max(np.abs(O.fft1(x, -1) - dft)) < 1e-12 * max(1, np.max(np.abs(dft)))
    idft = np.exp(2j * np.pi * np.outer(m, m) / n) @ x
    assert np.max(np.abs(O.fft1(x, +1) - idft)) < 1e-12 * max(1, np.max(np.abs(idft)))


@pytest.mark.parametrize("N", [4, 8, 12, 16, 24, 32])
def test_rfft3_vs_numpy_layout(N):
    # column-major (N,N,N) [ix,iy,iz] halved along z == numpy C-order [iz,iy,ix] rfftn over the
    # first axis last: numpy.fft.rfftn(a.transpose(2,1,0)) is halved along the C-order last axis.
    rng = np.random.default_rng(N)
    f = rng.normal(size=(N, N, N))                  # [iz, iy, ix]
    ours = O.rfft3(f)                                # [kz, ky, kx]
    ref = np.fft.rfftn(f.transpose(2, 1, 0))         # [ix,iy,iz] -> [kx, ky, kz]
    assert np.max(np.abs(ours - ref.transpose(2, 1, 0))) < 1e-12 * np.max(np.abs(ref))
    back = O.brfft3(ours) / N**3
    assert np.max(np.abs(back - f)) < 1e-13 * np.max(np.abs(f))


def test_brfft3_takes_real_part_of_dc_and_nyquist_planes():
    # c2r of a non-Hermitian half spectrum: numpy's irfftn (c2c on x, y, then c2r on z) is the
    # semantics of the reference's FFTW/JAX path
    N = 8
    rng = np.random.default_rng(3)
    s = rng.normal(size=(N // 2 + 1, N, N)) + 1j * rng.normal(size=(N // 2 + 1, N, N))
    ref = np.fft.irfftn(s.transpose(2, 1, 0), s=(N, N, N), axes=(0, 1, 2)) * N**3
    assert np.max(np.abs(O.brfft3(s) - ref.transpose(2, 1, 0))) < 1e-12 * np.max(np.abs(ref))


def test_rfft3_float32_close_to_float64():
    rng = np.random.default_rng(1)
    f = rng.normal(size=(32, 32, 32))
    a, b = O.rfft3(f), O.rfft3(f.astype(np.float32))
    assert b.dtype == np.complex64
    assert np.max(np.abs(a - b)) < 2e-5 * np.max(np.abs(a))


# ---------------------------------------------------------------- ranlxd1 / N-GenIC
def test_ranlxd1_known_answers():
    # runtests.jl:409-414
    k = G.KNOWN
    u = O.ranlxd1_uniforms(42, 5)
    assert np.max(np.abs(u - np.array(k["ranlxd1_seed42_first5"]))) <= k["ranlxd1_atol"]


def test_ngenic_field_known_answer():
    # runtests.jl:415-420
    k = G.KNOWN
    f = O.ngenic_field_gsl(42, 8)
    assert f.shape[::-1] == tuple(k["ngenic_field_gsl_42_8_shape"])
    assert abs(np.sqrt(np.sum(np.abs(f) ** 2)) - k["ngenic_field_gsl_42_8_norm"]) < k["ngenic_field_gsl_42_8_norm_atol"]
    w = O.ngenic_wnoise_real(42, 8)
    assert 0.5 < np.sqrt(np.mean(w**2)) < 1.5


def test_ngenic_noise_unit_variance_and_hermitian():
    w = O.ngenic_wnoise_real(7, 32)
    assert abs(np.std(w) - 1) < 0.05 and abs(np.mean(w)) < 0.05
    f = O.ngenic_field_gsl(7, 16)
    # kz = 0 plane Hermitian: F(-kx, -ky) = conj F(kx, ky) away from the Nyquist lines
    p = f[0]
    for ky in range(1, 8):
        for kx in range(1, 8):
            assert p[ky, kx] == np.conj(p[(16 - ky) % 16, (16 - kx) % 16])


# ---------------------------------------------------------------- transfer function and P(k)
def test_eisenstein_hu_limits():
    c = O.Cosmology(0.2621, 0.049, 0.6766, 0.8102, 0.9665)
    T = O.eisenstein_hu(c, np.array([1e-6, 1e-5, 1e-4, 1.0, 10.0, 100.0]))
    assert abs(T[0] - 1) < 1e-4 and abs(T[1] - 1) < 1e-3  # T -> 1 as k -> 0
    assert np.all(np.diff(T) < 0) and T[-1] > 0            # monotone decreasing, positive


def test_eisenstein_hu_zero_baryon_limit():
    """As Omega_b -> 0 the EH98 fit (arXiv:astro-ph/9709112) reduces to its zero-baryon form T0(q) of
    eqs. 28-29: T0 = L0 / (L0 + C0 q^2), L0 = ln(e + 1.8 q), C0 = 14.2 + 386 / (1 + 69.9 q^1.08),
    q = k / (13.41 k_eq) with k_eq = 0.0746 Omega_m h^2 / Theta^2 Mpc^-1 (eq. 3).  Pins the CDM branch,
    the shape parameter q and k_eq's units (h/Mpc)."""
    h, Om, Tcmb = 0.6766, 0.3111, 2.72548
    c = O.Cosmology(Om - 1e-9, 1e-9, h, 0.8, 0.96)
    k = np.logspace(-4, 1.5, 200)
    T = O.eisenstein_hu(c, k, Tcmb)
    theta = Tcmb / 2.7
    k_eq = 0.0746 * Om * h * h / theta**2 / h          # h/Mpc
    q = k / (13.41 * k_eq)
    L0 = np.log(np.e + 1.8 * q)
    C0 = 14.2 + 386.0 / (1.0 + 69.9 * q**1.08)
    assert np.max(np.abs(T - L0 / (L0 + C0 * q * q))) < 1e-6
    # with baryons the transfer function is suppressed below the zero-baryon form at k > k_eq
    Tb = O.eisenstein_hu(O.Cosmology(0.2621, 0.049, h, 0.8, 0.96), k, Tcmb)
    hi = k > 0.1
    assert np.all(Tb[hi] < L0[hi] / (L0[hi] + C0[hi] * q[hi] ** 2))


def test_eisenstein_hu_scales_against_their_definitions():
    """k_eq = (2 Omega_m H0^2 z_eq)^(1/2) / c (EH98 eq. 3 is its numerical value, 0.0746 Omega_m h^2 Theta^-2
    Mpc^-1 with z_eq of eq. 2), and the sound horizon of eq. 6 against the paper's approximation eq. 26,
    s = 44.5 ln(9.83 / Omega_m h^2) / sqrt(1 + 10 (Omega_b h^2)^(3/4)) Mpc, good to 2 %: a wrong baryon
    term in R_d, R_eq or z_d moves s by more."""
    for (Oc, Ob, h) in ((0.2621, 0.049, 0.6766), (0.25, 0.03, 0.7), (0.3, 0.06, 0.65)):
        c = O.Cosmology(Oc, Ob, h, 0.8, 0.96)
        k_eq, z_eq, z_d, s = O.eh98_scales(c)
        om, ob = (Oc + Ob) * h * h, Ob * h * h
        k_def = np.sqrt(2 * (Oc + Ob) * (100.0 / 299792.458) ** 2 * z_eq)  # h/Mpc, H0 = 100 h km/s/Mpc
        assert k_eq == pytest.approx(k_def, rel=2e-3)
        s_fit = 44.5 * np.log(9.83 / om) / np.sqrt(1 + 10 * ob**0.75)        # Mpc
        assert s / h == pytest.approx(s_fit, rel=0.02)
        assert 1000 < z_d < 1100 < z_eq


def test_sigma8_normalisation_recomputed():
    # runtests.jl:70-86: sigma8 recomputed from the normalised table by an independent integrator
    c = O.Cosmology(0.2621, 0.049, 0.6766, 0.8102, 0.9665)
    k, pk = O.linear_power_spectrum(c)
    x = k * 8.0
    W = 3 * (np.sin(x) - x * np.cos(x)) / x**3
    lnk = np.log(k)
    from scipy.integrate import simpson
    s8 = np.sqrt(simpson(k**3 * pk * W**2, x=lnk) / (2 * np.pi**2))
    assert s8 == pytest.approx(0.8102, rel=1e-2)


def test_power_spectrum_single_mode():
    # a single cosine mode delta = A cos(k1 x) has P = A^2 V / 4 in its bin and zero elsewhere
    N, L, A = 32, 100.0, 0.3
    x = np.arange(N) * L / N
    field = np.broadcast_to(A * np.cos(2 * np.pi * 3 * x / L)[None, None, :], (N, N, N)).copy()
    k, pk, nm = O.power_spectrum(field, L, bins=N, logarithmic=False)
    kf = 2 * np.pi / L
    edges = np.linspace(kf, np.pi * N / L * np.sqrt(3), N + 1)
    b = np.searchsorted(edges, 3 * kf, side="right") - 1
    # |F(+-k1)|^2 V / N^6 = A^2 V / 4 at the two kz = 0 entries kx = +-3 (weight 1/2 each); the bin
    # average divides by the weight sum of all half-spectrum entries in the bin (enumerated here)
    m = np.arange(N)
    kx = np.where(m <= N // 2, m, m - N) * kf
    kz = np.arange(N // 2 + 1) * kf
    kk = np.sqrt(kz[:, None, None] ** 2 + kx[None, :, None] ** 2 + kx[None, None, :] ** 2)
    w = np.ones_like(kk)
    w[0] = w[-1] = 0.5
    inbin = (kk >= edges[b]) & (kk < edges[b + 1])
    assert pk[b] == pytest.approx(A**2 * L**3 / 4 * 1.0 / w[inbin].sum(), rel=1e-10)
    assert nm[b] == inbin.sum()
    assert np.sum(pk > 1e-20) == 1


def _half_grid(N, L):
    kf = 2 * np.pi / L
    m = np.arange(N)
    kx = np.where(m <= N // 2, m, m - N) * kf
    kz = np.arange(N // 2 + 1) * kf
    return np.sqrt(kz[:, None, None] ** 2 + kx[None, :, None] ** 2 + kx[None, None, :] ** 2)


def test_cross_power_spectrum_of_two_phased_modes():
    """f = A cos(k1.x + a), g = B cos(k1.x + b): Re(conj f_k g_k) V / N^6 = A B V cos(b - a) / 4 at +-k1;
    with k1 off the kz = 0 plane only +k1 is in the half spectrum, and the bin's unweighted mean divides
    by all its modes.  (The sign of the imaginary product is pinned: re re - im im would give cos(a + b).)"""
    N, L, A, B, a, b = 16, 80.0, 0.7, 1.3, 0.4, 1.5
    kf = 2 * np.pi / L
    x = np.arange(N) * L / N
    ph = kf * (2 * x[None, None, :] + 1 * x[None, :, None] + 3 * x[:, None, None])  # k1 = (2, 1, 3) kf
    f, g = A * np.cos(ph + a), B * np.cos(ph + b)
    bins = 12
    k, pc = O.cross_power_spectrum(f, g, L, bins=bins, logarithmic=False)
    edges = [(1 - t) * kf + t * (np.pi * N / L * np.sqrt(3)) for t in np.arange(bins + 1) / bins]
    kk = _half_grid(N, L)
    k1 = kf * np.sqrt(14.0)
    bb = max(i for i in range(bins) if edges[i] <= k1)
    n_k = np.sum((kk >= edges[bb]) & (kk < edges[bb + 1]))
    assert pc[bb] == pytest.approx(A * B * L**3 * np.cos(b - a) / 4 / n_k, rel=1e-10)
    assert np.sum(np.abs(pc) > 1e-12 * abs(pc[bb])) == 1
    # symmetric, and Re(conj f f) of one field is the unweighted mean of |f_k|^2
    assert np.allclose(O.cross_power_spectrum(g, f, L, bins=bins, logarithmic=False)[1], pc, rtol=1e-12, atol=0)
    ks, ps = O.cross_power_spectrum(f, f, L, bins=bins, logarithmic=False)
    assert ps[bb] == pytest.approx(A * A * L**3 / 4 / n_k, rel=1e-10)
    # bin centres: the mean |k| of the modes in each non-empty bin (logarithmic bins)
    kl, _ = O.cross_power_spectrum(f, g, L, bins=bins, logarithmic=True)
    le = [10 ** ((1 - t) * np.log10(kf) + t * np.log10(np.pi * N / L * np.sqrt(3))) for t in np.arange(bins + 1) / bins]
    for i in range(bins):
        m = (kk >= le[i]) & (kk < le[i + 1])
        if m.any():
            assert kl[i] == pytest.approx(kk[m].mean(), rel=1e-12)


def test_binned_estimator_closed_forms():
    """pk_full over given edges.  Density part with linear edges kf (b + 1/2) equals the pinned linear-bin
    estimator (same weights and window); displacement part of psi_c = A_c cos(k1.q) is
    L^3 sum_c w1 A_c^2 / 4 / (sum_w / 3) with w1 = 2 off the kz = 0, N/2 planes."""
    N, L = 32, 100.0
    kf = 2 * np.pi / L
    delta = np.random.default_rng(3).normal(size=(N, N, N))
    nb = N // 4 - 1
    edges = kf * (np.arange(nb + 1) + 0.5)
    pd, _ = O.pk_binned(L, edges, delta=delta, deconv_W=2)
    _, plin, _ = O.power_spectrum_lin(delta, L, W=2, nbins=nb)
    assert np.allclose(pd, plin, rtol=1e-12, atol=0)
    n = 16
    q = np.arange(n) * L / n
    ph = kf * (1 * q[None, None, :] + 2 * q[None, :, None] + 2 * q[:, None, None])  # |k1| = 3 kf
    Ac = np.array([0.5, 1.0, 2.0])
    psi = Ac[:, None, None, None] * np.cos(ph)[None]
    e2 = kf * np.array([0.9, 2.5, 3.5, 6.0])
    _, pp = O.pk_binned(L, e2, psi=psi)
    kk = _half_grid(n, L)
    w = np.full_like(kk, 2.0)
    w[0] = w[-1] = 1.0
    sw = w[(kk >= e2[1]) & (kk < e2[2])].sum()
    assert pp[1] == pytest.approx(L**3 * np.sum(2.0 * Ac**2 / 4) / sw, rel=1e-10)
    assert abs(pp[0]) < 1e-20 * pp[1] and abs(pp[2]) < 1e-20 * pp[1]


def test_power_spectrum_white_noise_amplitude():
    # runtests.jl:174-186: unit white noise on N^3 cells of volume (L/N)^3 has P = (L/N)^3
    N, L = 32, 200.0
    w = np.random.default_rng(0).normal(size=(N, N, N))
    k, pk, nm = O.power_spectrum_lin(w, L, W=0)
    assert np.mean(pk) == pytest.approx((L / N) ** 3, rel=0.03)


def test_ic_operator_amplitude_and_sign():
    # grf.jl:118-144 maps white noise omega to the potential phi of a density with spectrum P(k).
    # Pinned without restating the map: (1) delta = -k^2 phi has the input P(k) on average over the
    # modes of each bin (amplitude: the square root and the (N/L)^3 factor); (2) delta_k / omega_k is
    # real and positive (the density follows the noise; the Poisson sign); (3) the first-order
    # displacement of the golden-pinned compute_core converges on those overdensities.
    n, L = 64, 400.0
    c = O.Cosmology(0.2621, 0.049, 0.6766, 0.8102, 0.9665)
    k, pk = O.linear_power_spectrum(c)
    w = O.ngenic_wnoise_real(42, n)
    fphi = O.white_noise_to_fphi(O.ic_operator(n, L, k, pk), w)
    kf = 2 * np.pi / L
    kfull = np.fft.fftfreq(n, 1.0 / n) * kf
    kz = np.arange(n // 2 + 1) * kf
    k2 = kz[:, None, None] ** 2 + kfull[None, :, None] ** 2 + kfull[None, None, :] ** 2
    delta_k = -k2 * fphi
    Wk = O.rfft3(w)
    live = (k2 > 0) & (np.abs(Wk) > 1e-9 * np.max(np.abs(Wk)))  # N-GenIC leaves some modes empty
    assert live.mean() > 0.9
    ratio = delta_k[live] / Wk[live]
    assert np.max(np.abs(ratio.imag)) <= 1e-10 * np.max(np.abs(ratio.real))
    assert np.all(ratio.real > 0)
    delta = O.brfft3(delta_k) / n**3
    kb, pb, nm = O.power_spectrum_lin(delta, L, W=0)
    ok = nm > 400
    pin = np.exp(np.interp(np.log(kb[ok]), np.log(k), np.log(pk)))
    r = pb[ok] / pin
    assert ok.sum() >= 8
    assert abs(np.mean(r) - 1) < 0.03
    assert np.all(np.abs(r - 1) < 6 * np.sqrt(2.0 / nm[ok]))
    psi1, _ = O.compute_core2(fphi, L)
    div = sum(1j * kk * O.rfft3(psi1[a]) for a, kk in ((0, kfull[None, None, :]), (1, kfull[None, :, None]),
                                                       (2, kz[:, None, None])))
    conv = -O.brfft3(div) / n**3
    assert np.corrcoef(conv.ravel(), delta.ravel())[0, 1] > 0.99


# ---------------------------------------------------------------- PM force invariants
def test_lean_equals_solver():
    # pm_lean_tests.jl:5-21 on other random inputs (the reference's MersenneTwister stream is Julia-only)
    n, res, L = 24, 48, 120.0
    psi = 3 * np.random.default_rng(5).normal(size=(3, n, n, n))
    for W, g in ((2, 4), (3, 2), (4, 6)):
        cfg = O.pm_config(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=g)
        a = O.pm_acceleration(psi, cfg, "solver")
        b = O.pm_acceleration(psi, cfg, "lean")
        assert G.rel(b, a) < 1e-12
        a32 = O.pm_acceleration(psi.astype(np.float32), cfg, "solver")
        b32 = O.pm_acceleration(psi.astype(np.float32), cfg, "lean")
        assert G.rel(b32, a32) < 2e-5
    cfg = O.pm_config(res_pm=res, n_part=n, boxsize=L, worder=2, deconvolve=True)
    assert G.rel(O.pm_acceleration(psi, cfg, "lean"), O.pm_acceleration(psi, cfg, "solver")) < 2e-5
    with pytest.raises(ValueError):
        O.pm_acceleration(psi, O.pm_config(res_pm=res, n_part=n, boxsize=L, antialias=1), "lean")


def test_fd_stencil_equals_kspace_symbol():
    # circulant identity (pm.jl:309-311): real-space FD gradient == k-space FD symbol, all orders
    n, res, L = 16, 32, 100.0
    psi = 2 * np.random.default_rng(9).normal(size=(3, n, n, n))
    for g in (2, 4, 6):
        cfg = O.pm_config(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=g, antialias=1)
        assert G.rel(O.pm_acceleration(psi, cfg, "solver"), O.pm_acceleration(psi, cfg, "reference")) < 1e-12


def test_cic_mass_conservation_and_gather_adjoint():
    # runtests.jl:430; and <S x, m> = <x, G m>: scatter and gather are transposes
    rng = np.random.default_rng(2)
    X = rng.uniform(-50, 150, size=(3, 500))
    for W in (2, 3, 4):
        m = O.pm_scatter(X, 16, 100.0, W)
        assert m.sum() == pytest.approx(500, rel=1e-12)
        f = rng.normal(size=(16, 16, 16))
        g = O.pm_gather(f, X, 100.0, W)
        assert np.sum(m * f) == pytest.approx(np.sum(g), rel=1e-11)


def test_fixed_point_deposit_matches_float_and_is_order_free():
    rng = np.random.default_rng(4)
    X = rng.uniform(0, 100, size=(3, 4000)).astype(np.float32)
    a = O.pm_scatter_fixed(X, 16, np.float32(100.0), 3)
    perm = rng.permutation(4000)
    b = O.pm_scatter_fixed(np.ascontiguousarray(X[:, perm]), 16, np.float32(100.0), 3)
    assert np.array_equal(a, b)
    m = O.pm_scatter(X.astype(np.float64), 16, 100.0, 3)
    assert np.max(np.abs(a / 2.0**32 - m)) < 1e-5


def test_plane_wave_linear_growth():
    # nbody_parity_tests.jl:68-90: BullFrog with one and 16 steps, FastPM with 16, symplectic with 64
    c = O.Cosmology.named("Planck18EEBAOSN")
    L, a0 = 100.0, 0.05
    D0, D1 = c.Dplus(a0), c.Dplus(1.0)
    err = {}
    for n in (16, 32):
        q = np.arange(n) * L / n
        psi1 = np.zeros((3, n, n, n))
        psi1[0] = 1e-3 * np.sin(2 * np.pi * q / L)[None, None, :]
        for st, ns in (("bullfrog", 1), ("bullfrog", 16), ("fastpm", 16), ("symplectic", 64)):
            cfg = O.pm_config(res_pm=2 * n, n_part=n, boxsize=L)
            psi, _ = O.run_nbody(c, D0 * psi1, psi1.copy(), cfg, a0, 1.0, ns, stepper=st)
            assert np.max(np.abs(psi[1:])) < 1e-12
            err[(n, st, ns)] = abs(np.sum(psi[0] * psi1[0]) / np.sum(psi1[0] ** 2) / D1 - 1)
    for st, ns in (("bullfrog", 1), ("bullfrog", 16), ("fastpm", 16), ("symplectic", 64)):
        assert err[(32, st, ns)] < 1e-2
        assert err[(16, st, ns)] / err[(32, st, ns)] > 3


def test_callback_steps():
    # pm_lean_tests.jl:30-33 shape: the callback sees k = 0..n_steps in order
    c = O.Cosmology.named("Planck18EEBAOSN")
    n = 8
    psi = 0.05 * np.random.default_rng(2).normal(size=(3, n, n, n))
    seen = []
    O.run_nbody(c, psi, psi / 0.05, O.pm_config(res_pm=16, n_part=n, boxsize=100.0), 0.05, 0.5, 10,
                lean=True, callback=lambda k, a: seen.append(k))
    assert seen == list(range(11))


# ------------------------------------------------------------------ learned-correction oracle
def test_nn_oracle_bf16_rounding_known_answers():
    import nn_ref as R
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even is 1; 1 + 3 2^-9 rounds up; sign and
    # exact values are kept
    x = np.array([1 + 2.0**-8, 1 + 3 * 2.0**-9, -(1 + 2.0**-8), 3.0, 1 + 2.0**-7 + 2.0**-8], np.float32)
    assert list(R.bf16(x)) == [1.0, 1 + 2.0**-7, -1.0, 3.0, 1 + 2.0**-6]


def test_nn_oracle_gradient_is_the_derivative_of_its_forward():
    """With the roundings off, the oracle's parameter gradient equals central differences of its forward
    (which pins the chain rule through the residual blocks, the relu masks and the transposed taps)."""
    import nn_ref as R
    N, n, a, s = 8, 4, 0.5, 0.3
    rng = np.random.default_rng(0)
    p = np.zeros(R.N_PARAMS, np.float32)
    for l, (co, ci) in enumerate([(16, 2), (16, 16), (16, 16), (1, 16)]):
        W = np.zeros((27, 16, 16), np.float32)
        W[:, :co, :ci] = rng.normal(scale=0.3, size=(27, co, ci))
        p[l * R.PER_LAYER:l * R.PER_LAYER + 27 * 256] = W.reshape(-1)
        p[l * R.PER_LAYER + 27 * 256:l * R.PER_LAYER + 27 * 256 + co] = rng.normal(scale=0.1, size=co)
    p = p.astype(np.float64)
    dens = rng.lognormal(size=(N, N, N)) * (n / N) ** 3
    g = rng.normal(size=(N, N, N))
    out, cache = R.forward(p, dens, (N / n) ** 3, a, s, rnd=R.exact)
    grad = R.backward(cache, g, s)
    idx = [0, 27 * 256 + 3, R.PER_LAYER + 5 * 256 + 17, 2 * R.PER_LAYER + 13 * 256 + 40, 3 * R.PER_LAYER + 26 * 16 + 7,
           3 * R.PER_LAYER + 27 * 256]
    for i in idx:
        eps = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (np.sum(g * R.forward(pp, dens, (N / n) ** 3, a, s, rnd=R.exact)[0]) -
              np.sum(g * R.forward(pm, dens, (N / n) ** 3, a, s, rnd=R.exact)[0])) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
