"""Reverse mode of the PM path against central finite differences of the forward map.

The checker is the directional derivative: <ybar, (f(x + eps v) - f(x - eps v)) / 2 eps> must equal
<vjp(ybar), v>.  Each adjoint is checked against finite differences of the device forward and of the
oracle's forward (the CPU restatement, pinned by the golden vectors).  FP64 throughout; the assignment
weights are C^1 for TSC and C^2 for PCS, and CIC is piecewise linear, so its tolerance allows a rare
cell-face crossing inside the difference step.
"""
import numpy as np
import pytest

import golden as G
import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")

COSMO = (0.2621, 0.049, 0.6766, 0.8102, 0.9665)


def _dot(a, b):
    return float(np.sum(np.asarray(a, np.float64) * np.asarray(b, np.float64)))


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _psi(n, L, seed, scale=2.0):
    return np.random.default_rng(seed).normal(scale=scale * L / 100.0, size=(3, n, n, n))


@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("aa", [0, 1])
def test_vjp_acceleration_fd(W, order, aa):
    n, res, L = 16, 32, 100.0
    rng = np.random.default_rng(10 * W + order + aa)
    psi, abar, v = _psi(n, L, W + 3 * order + aa), rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=order, lap_order=order if W > 2 else 0,
                     antialias=aa)
    eps = 1e-5
    with D.PMSolverHIP(cfg, precision=64) as S:
        g = _dot(S.vjp_acceleration(psi, abar), v)
        fd = (_dot(abar, S.pm_acceleration(psi + eps * v)) - _dot(abar, S.pm_acceleration(psi - eps * v))) / (2 * eps)
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=order, lap_order=order if W > 2 else 0, antialias=aa)
    fdo = (_dot(abar, O.pm_acceleration(psi + eps * v, ocfg, "solver")) -
           _dot(abar, O.pm_acceleration(psi - eps * v, ocfg, "solver"))) / (2 * eps)
    tol = 1e-4 if W == 2 else 1e-6
    assert _rel(g, fd) < tol, (g, fd)
    assert _rel(g, fdo) < tol, (g, fdo)


@pytest.mark.parametrize("deconvolve", [False, True])
def test_vjp_acceleration_with_transfer_function_fd(deconvolve):
    """The force VJP with a fitted transfer function g(|k|) on the Poisson kernel (and window
    deconvolution): g is real and isotropic, so the kernel stays symmetric and the same reverse pass
    applies; checked against central differences of the corrected forward."""
    n, res, L = 16, 32, 100.0
    rng = np.random.default_rng(21)
    psi, abar, v = _psi(n, L, 22), rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=4, lap_order=4, antialias=1,
                     deconvolve=deconvolve)
    gain = 1.0 + 0.3 * np.sin(np.linspace(0.0, 3.0, 12))
    eps = 1e-5
    with D.PMSolverHIP(cfg, precision=64) as S:
        S.set_kgain(gain, np.pi * res / L)
        g = _dot(S.vjp_acceleration(psi, abar), v)
        fd = (_dot(abar, S.pm_acceleration(psi + eps * v)) - _dot(abar, S.pm_acceleration(psi - eps * v))) / (2 * eps)
        S.set_kgain(None)
        plain = _dot(S.vjp_acceleration(psi, abar), v)
    assert _rel(g, fd) < 1e-6, (g, fd)
    assert _rel(g, plain) > 1e-3  # the transfer function changes the gradient


@pytest.mark.parametrize("W", [2, 3, 4])
def test_vjp_deposit_fd(W):
    n, res, L = 16, 32, 100.0
    rng = np.random.default_rng(W)
    psi, mbar, v = _psi(n, L, 7 + W), rng.normal(size=(res, res, res)), rng.normal(size=(3, n, n, n))
    eps = 1e-5
    with D.PMSolverHIP(D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W), precision=64) as S:
        g = _dot(S.vjp_deposit(psi, mbar), v)
        fd = (_dot(mbar, S.deposit(psi + eps * v)) - _dot(mbar, S.deposit(psi - eps * v))) / (2 * eps)
    dep = lambda p: O.pm_scatter(O.lagrangian_positions(p, L), res, L, W)  # noqa: E731
    fdo = (_dot(mbar, dep(psi + eps * v)) - _dot(mbar, dep(psi - eps * v))) / (2 * eps)
    tol = 1e-4 if W == 2 else 1e-7
    assert _rel(g, fd) < tol, (g, fd)
    assert _rel(g, fdo) < tol, (g, fdo)


@pytest.mark.parametrize("stepper, W, aa", [("bullfrog", 3, 0), ("fastpm", 4, 1), ("symplectic", 3, 1)])
def test_vjp_run_fd(stepper, W, aa):
    """Three DKD steps: <Psibar, dPsi_f> + <Pibar, dPi_f> along (v, u) against <(Psibar_0, Pibar_0), (v, u)>."""
    n, res, L, steps = 16, 32, 100.0, 3
    rng = np.random.default_rng(5)
    psi0 = _psi(n, L, 11, 1.0)
    pi0 = rng.normal(scale=2.0, size=(3, n, n, n))
    pb, mb = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    v, u = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, stepper, 0.05, 0.5, steps, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=4, antialias=aa)
    eps = 1e-6
    with D.PMSolverHIP(cfg, precision=64) as S:
        S.set_state(psi0, pi0)
        S.run_checkpointed(*coef)
        g0, g1 = S.vjp_run(pb, mb)

        def f(e):
            S.set_state(psi0 + e * v, pi0 + e * u)
            S.run(*coef)
            p, m = S.get_state()
            return _dot(pb, p) + _dot(mb, m)

        fd = (f(eps) - f(-eps)) / (2 * eps)
    g = _dot(g0, v) + _dot(g1, u)
    assert _rel(g, fd) < 1e-6, (g, fd)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("n", [16, 32])
def test_vjp_ics_fd(order, n):
    """2LPT is quadratic in the white noise, so the central difference is exact up to rounding."""
    L, a = 2.0 * n, 0.05
    rng = np.random.default_rng(n + order)
    white = O.ngenic_wnoise_real(42, n)
    v = rng.normal(size=(n, n, n))
    pb, mb = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    oc = O.Cosmology(*COSMO)
    k, pk = O.linear_power_spectrum(oc)
    Dg = oc.Dplus(a)
    eps = 1e-3
    with D.PMSolverHIP(D.PMConfig(res_pm=2 * n, n_part=n, boxsize=L), precision=64) as S:
        wb = S.vjp_ics_lpt(k, pk, Dg, pb, mb, order=order, white=white)

        def f(e):
            S.ics_lpt(k, pk, Dg, order=order, white=white + e * v)
            p, m = S.get_state()
            return _dot(pb, p) + _dot(mb, m)

        fd = (f(eps) - f(-eps)) / (2 * eps)

    def fo(e):
        fphi = O.white_noise_to_fphi(O.ic_operator(n, L, k, pk), white + e * v)
        psi1, psi2 = O.compute_core2(fphi, L)
        if order == 1:
            psi2 = np.zeros_like(psi2)
        p, m = O.nbody_ics_lpt2(oc, psi1, psi2, a)
        return _dot(pb, p) + _dot(mb, m)

    fdo = (fo(eps) - fo(-eps)) / (2 * eps)
    g = _dot(wb, v)
    assert _rel(g, fd) < 1e-9, (g, fd)
    assert _rel(g, fdo) < 1e-9, (g, fdo)


def test_vjp_ics_reuses_forward_fields_only_for_the_same_inputs():
    """The IC reverse pass takes the nine second-order fields A_ik from the last forward when that forward
    had the same inputs (seed or white-noise values, P(k) table, D aside) and recomputes them otherwise:
    the cotangent is the same bits either way, and a forward with other noise, another seed, another
    table or a white noise changed in place is not taken for the same one."""
    n, L, a = 16, 32.0, 0.05
    rng = np.random.default_rng(7)
    white = O.ngenic_wnoise_real(42, n)
    other = white + 1e-3 * rng.normal(size=white.shape)
    pb, mb = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    oc = O.Cosmology(*COSMO)
    k, pk = O.linear_power_spectrum(oc)
    Dg = oc.Dplus(a)
    with D.PMSolverHIP(D.PMConfig(res_pm=2 * n, n_part=n, boxsize=L), precision=64) as S:
        # recomputed (no forward yet), then after the forward of the same noise (kept fields)
        ref = S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white)
        S.ics_lpt(k, pk, Dg, white=white)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white), ref)
        # a second reverse pass from the same forward
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white), ref)
        # the forward of other noise: its fields must not serve the first noise
        ref_o = S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=other)
        assert not np.array_equal(ref_o, ref)
        S.ics_lpt(k, pk, Dg, white=other)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white), ref)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=other), ref_o)
        # another P(k) table
        S.ics_lpt(k, pk, Dg, white=white)
        assert not np.array_equal(S.vjp_ics_lpt(k, 1.01 * pk, Dg, pb, mb, white=white), ref)
        # N-GenIC noise of a seed (generated on the device): the same seed hits, another seed does not
        S.ics_lpt(k, pk, Dg, seed=43)
        ref_s = S.vjp_ics_lpt(k, pk, Dg, pb, mb, seed=42)
        S.ics_lpt(k, pk, Dg, seed=42)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, seed=42), ref_s)
        S.ics_lpt(k, pk, Dg, seed=43)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, seed=42), ref_s)
        assert G.rel(ref_s, ref) < 1e-9
        # a forward from a device array: its values are the key, and the array changed in place and
        # used again is another noise
        wd = D.DeviceArray.from_numpy(white)
        S.ics_lpt_device(k, pk, Dg, wd)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white), ref)
        wd.copy_from(other)
        S.ics_lpt_device(k, pk, Dg, wd)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=white), ref)
        assert np.array_equal(S.vjp_ics_lpt(k, pk, Dg, pb, mb, white=other), ref_o)
        wd.free()


def test_density_loss_gradient_directions():
    """The differentiation workload's check at reduced size: the gradient of sum (delta - target)^2 with
    respect to the white noise through 2LPT and 10 BullFrog CIC steps (grad order 4), against central
    differences of the whole pipeline along 32 random directions (FP64), and the FP32 gradient against
    the FP64 one."""
    n, L = 16, 64.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    white = O.ngenic_wnoise_real(7, n)
    target = np.random.default_rng(3).normal(scale=0.5, size=(n, n, n))
    kw = dict(boxsize=L, res_pm=n, n_steps=10, stepper="bullfrog", order=2, worder=2, grad_kernel_order=4)
    cfg = D.PMConfig(res_pm=n, n_part=n, boxsize=L, worder=2, grad_order=4)
    with D.PMSolverHIP(cfg, precision=64) as S:
        loss, wb, _ = D.density_loss_gradient(c, white, target, precision=64, solver=S, **kw)
        rng = np.random.default_rng(11)
        eps = 1e-6
        bad = []
        for i in range(32):
            v = rng.normal(size=(n, n, n))
            lp = D.density_loss_gradient(c, white + eps * v, target, precision=64, solver=S, **kw)[0]
            lm = D.density_loss_gradient(c, white - eps * v, target, precision=64, solver=S, **kw)[0]
            fd, g = (lp - lm) / (2 * eps), _dot(wb, v)
            if _rel(g, fd) > 1e-4:
                bad.append((i, g, fd))
    assert loss > 0
    assert len(bad) <= 1, bad  # a CIC cell-face crossing inside the step may spoil one direction
    with D.PMSolverHIP(cfg, precision=32) as S32:
        _, wb32, _ = D.density_loss_gradient(c, white, target, precision=32, solver=S32, **kw)
    assert G.rel(wb32, wb) < 2e-3


@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("aa", [0, 1])
def test_vjp_acceleration_fp32_tiled_vector_deposit(W, aa):
    """The FP32 force VJP on a 64^3 mesh, where every axis holds at least two 32x16x8 FP32 tiles, so the
    cotangent's vector deposit runs tiled in LDS (the 16^3-particle tests above have one tile per axis
    in FP32 and fall back to device atomics): against the FP64 VJP of the same positions."""
    n, res, L = 32, 64, 100.0
    rng = np.random.default_rng(100 + 10 * W + aa)
    psi, abar = _psi(n, L, 30 + W + aa), rng.normal(size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=4, antialias=aa)
    with D.PMSolverHIP(cfg, precision=64) as S:
        ref = S.vjp_acceleration(psi, abar)
    with D.PMSolverHIP(cfg, precision=32) as S:
        got = S.vjp_acceleration(psi.astype(np.float32), abar.astype(np.float32))
    assert G.rel(got, ref) < 5e-4, G.rel(got, ref)


def test_density_loss_gradient_fp32_tiled_64():
    """The density-loss gradient at n = res_pm = 64 (tiled FP32 vector deposit in every step) against
    the FP64 gradient."""
    n, L = 64, 125.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    white = O.ngenic_wnoise_real(9, n)
    target = np.random.default_rng(4).normal(scale=0.5, size=(n, n, n))
    kw = dict(boxsize=L, res_pm=n, n_steps=10, stepper="bullfrog", order=2, worder=2, grad_kernel_order=4)
    _, wb64, _ = D.density_loss_gradient(c, white, target, precision=64, **kw)
    _, wb32, _ = D.density_loss_gradient(c, white, target, precision=32, **kw)
    assert G.rel(wb32, wb64) < 5e-3, G.rel(wb32, wb64)


@pytest.mark.parametrize("prec", [32, 64])
def test_density_loss_gradient_on_device_matches_host_arrays(prec):
    """The gradient with device arrays (white noise, target and white-noise cotangent in device memory, the
    loss formed on the device by dj_density_loss_device) equals the host-array gradient: the same deposit,
    the same cotangent bits into the reverse pass.  The solver is reused across calls, first with 1LPT ICs,
    so the IC pipeline's kept buffers are regrown for 2LPT."""
    n, res, L = 32, 64, 100.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    white = O.ngenic_wnoise_real(5, n)
    target = np.random.default_rng(6).normal(scale=0.5, size=(res, res, res))
    kw = dict(boxsize=L, res_pm=res, n_steps=6, stepper="fastpm", order=2, worder=3, grad_kernel_order=4)
    loss_h, wb_h, _ = D.density_loss_gradient(c, white, target, precision=prec, **kw)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=4)
    k_tab, pk_tab = D.linear_power_spectrum(c)
    w_d, t_d = D.DeviceArray.from_numpy(white), D.DeviceArray.from_numpy(target)
    with D.PMSolverHIP(cfg, precision=prec) as S:
        S.ics_lpt(k_tab, pk_tab, c.Dplus(0.05), order=1, seed=3)
        for _ in range(2):
            loss_d, wb_d, _ = D.density_loss_gradient(c, w_d, t_d, precision=prec, solver=S, on_device=True, **kw)
            wd = wb_d.numpy()
            assert _rel(loss_d, loss_h) < 1e-12, (loss_d, loss_h)
            # FP32: the vector deposit's float atomics sum in a varying order
            assert G.rel(wd, wb_h) < (1e-4 if prec == 32 else 1e-11), G.rel(wd, wb_h)
        # the loss alone, against the host's from the same state (zero target: target None)
        mesh = S.deposit().astype(np.float64) / (n**3 / res**3) - 1.0
        assert _rel(S.density_loss_device(None), float(np.sum(mesh * mesh))) < 1e-12
        with pytest.raises(TypeError):
            S.density_loss_device(target)
        with pytest.raises(ValueError):
            S.density_loss_device(D.DeviceArray((res, res, res - 1), np.float64))
    with D.PMSolverHIP(cfg, precision=prec, deposit="fixed64") as S:
        S.ics_lpt(k_tab, pk_tab, c.Dplus(0.05), order=1, seed=3)
        with pytest.raises(D.DJError):
            S.density_loss_device(t_d)


def test_resident_cotangents_are_checked():
    """dj_vjp_run and dj_vjp_ics_lpt with NULL cotangents need resident ones from dj_vjp_deposit (or a
    resident dj_vjp_run); dj_vjp_acceleration overwrites them; options without an adjoint are refused by
    dj_run_checkpointed before it runs."""
    n, L = 16, 64.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 0.5, 2, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    oc = O.Cosmology(*COSMO)
    k, pk = O.linear_power_spectrum(oc)
    white = O.ngenic_wnoise_real(3, n)
    with D.PMSolverHIP(D.PMConfig(res_pm=32, n_part=n, boxsize=L), precision=64) as S:
        S.ics_lpt(k, pk, oc.Dplus(0.05), white=white)
        S.run_checkpointed(*coef)
        with pytest.raises(D.DJError):
            S.vjp_run()
        with pytest.raises(D.DJError):
            S.vjp_ics_lpt(k, pk, oc.Dplus(0.05), white=white)
        S.vjp_deposit(None, np.ones((32, 32, 32)), resident=True)
        S.vjp_run()
        S.vjp_acceleration(None, np.ones((3, n, n, n)))
        with pytest.raises(D.DJError):
            S.vjp_ics_lpt(k, pk, oc.Dplus(0.05), white=white)
    # the conv net of the learned correction has no reverse pass through the steps
    with D.PMSolverHIP(D.PMConfig(res_pm=64, n_part=n, boxsize=L), precision=32) as S:
        S.set_state(np.zeros((3, n, n, n), np.float32))
        S.nn_set_params(D.learned.init_params(1))
        S.nn_enable(True, 1.0)
        with pytest.raises(D.DJError):
            S.run_checkpointed(*coef)


@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("aa", [0, 1])
@pytest.mark.parametrize("deconvolve", [False, True])
def test_vjp_acceleration_ik_fd(W, aa, deconvolve):
    """The force VJP with the spectral (ik) gradient (grad_order 0, the discreteness-suppressed settings use
    it with PCS and deconvolution): against central differences of the device forward and of the oracle's."""
    n, res, L = 16, 32, 100.0
    rng = np.random.default_rng(200 + 10 * W + 2 * aa + int(deconvolve))
    psi, abar, v = _psi(n, L, 40 + W + aa), rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=0, antialias=aa, deconvolve=deconvolve)
    eps = 1e-5
    with D.PMSolverHIP(cfg, precision=64) as S:
        g = _dot(S.vjp_acceleration(psi, abar), v)
        fd = (_dot(abar, S.pm_acceleration(psi + eps * v)) - _dot(abar, S.pm_acceleration(psi - eps * v))) / (2 * eps)
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=0, antialias=aa, deconvolve=deconvolve)
    fdo = (_dot(abar, O.pm_acceleration(psi + eps * v, ocfg, "solver")) -
           _dot(abar, O.pm_acceleration(psi - eps * v, ocfg, "solver"))) / (2 * eps)
    tol = 1e-4 if W == 2 else 1e-6
    assert _rel(g, fd) < tol, (g, fd)
    assert _rel(g, fdo) < tol, (g, fdo)


@pytest.mark.parametrize("prec", [64, 32])
def test_vjp_run_ik_fd(prec):
    """Three BullFrog steps with PCS, ik gradient, deconvolution and interlacing (64^3 mesh, so the FP32
    vector deposit runs tiled): the reverse pass against central differences (FP64), FP32 against FP64."""
    n, res, L, steps = 32, 64, 100.0, 3
    rng = np.random.default_rng(9)
    psi0 = _psi(n, L, 12, 1.0)
    pi0 = rng.normal(scale=2.0, size=(3, n, n, n))
    pb, mb = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    v, u = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 0.5, steps, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=4, grad_order=0, antialias=1, deconvolve=True)
    dt = np.float64 if prec == 64 else np.float32
    with D.PMSolverHIP(cfg, precision=prec) as S:
        S.set_state(psi0.astype(dt), pi0.astype(dt))
        S.run_checkpointed(*coef)
        g0, g1 = S.vjp_run(pb.astype(dt), mb.astype(dt))
        if prec == 32:
            with D.PMSolverHIP(cfg, precision=64) as S64:
                S64.set_state(psi0, pi0)
                S64.run_checkpointed(*coef)
                r0, r1 = S64.vjp_run(pb, mb)
            assert G.rel(g0, r0) < 1e-3 and G.rel(g1, r1) < 1e-3
            return
        eps = 1e-6

        def f(e):
            S.set_state(psi0 + e * v, pi0 + e * u)
            S.run(*coef)
            p, m = S.get_state()
            return _dot(pb, p) + _dot(mb, m)

        fd = (f(eps) - f(-eps)) / (2 * eps)
    g = _dot(g0, v) + _dot(g1, u)
    assert _rel(g, fd) < 1e-6, (g, fd)


@pytest.mark.parametrize("resampling", ["fourier", "linear"])
@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("grad", [4, 0])
def test_vjp_acceleration_resampling_fd(resampling, W, grad):
    """The force VJP with sheet resampling (n_resample 2, interlacing): the resampled points enter the
    density, their cotangents go back to Psi through the transposed interpolation (trilinear, or the
    conjugate Fourier phase); against central differences of the device forward and of the oracle's."""
    n, res, L = 16, 32, 100.0
    rng = np.random.default_rng(300 + 10 * W + grad + (resampling == "linear"))
    psi, abar, v = _psi(n, L, 60 + W + grad), rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    kw = dict(worder=W, grad_order=grad, antialias=1, n_resample=2, resampling=resampling)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, **kw)
    ocfg = O.pm_config(res, n, L, **kw)
    # CIC weights are piecewise linear: a point (8 per particle here) crossing a cell face inside the
    # difference step spoils that step's difference (by 1e-3 to 1e-1 at 1e-5); the best of two steps is
    # taken, which a single crossing cannot spoil both of
    steps = (1e-6, 3e-7)
    with D.PMSolverHIP(cfg, precision=64) as S:
        g = _dot(S.vjp_acceleration(psi, abar), v)
        fd = [(_dot(abar, S.pm_acceleration(psi + e * v)) - _dot(abar, S.pm_acceleration(psi - e * v))) / (2 * e)
              for e in steps]
    fdo = [(_dot(abar, O.pm_acceleration(psi + e * v, ocfg, "solver")) -
            _dot(abar, O.pm_acceleration(psi - e * v, ocfg, "solver"))) / (2 * e) for e in steps]
    tol = 1e-5 if W == 2 else 1e-6
    assert min(_rel(g, f) for f in fd) < tol, (g, fd)
    assert min(_rel(g, f) for f in fdo) < tol, (g, fdo)


@pytest.mark.parametrize("prec", [64, 32])
def test_vjp_run_resampling_fd(prec):
    """Three BullFrog steps with the discreteness-suppressed settings (PCS, ik gradient, deconvolution,
    interlacing, 2^3 Fourier sheet resampling): the run's VJP against central differences (FP64); FP32
    against the FP64 VJP."""
    n, res, L, steps = 16, 32, 100.0, 3
    rng = np.random.default_rng(7)
    psi0 = _psi(n, L, 13, 1.0)
    pi0 = rng.normal(scale=2.0, size=(3, n, n, n))
    pb, mb = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    v, u = rng.normal(size=(3, n, n, n)), rng.normal(size=(3, n, n, n))
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 0.5, steps, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=4, grad_order=0, antialias=1, deconvolve=True,
                     n_resample=2, resampling="fourier")
    eps = 1e-6
    with D.PMSolverHIP(cfg, precision=64) as S:
        S.set_state(psi0, pi0)
        S.run_checkpointed(*coef)
        g0, g1 = S.vjp_run(pb, mb)

        def f(e):
            S.set_state(psi0 + e * v, pi0 + e * u)
            S.run(*coef)
            p, m = S.get_state()
            return _dot(pb, p) + _dot(mb, m)

        fd = (f(eps) - f(-eps)) / (2 * eps)
    g = _dot(g0, v) + _dot(g1, u)
    if prec == 64:
        assert _rel(g, fd) < 1e-6, (g, fd)
        return
    with D.PMSolverHIP(cfg, precision=32) as S:
        S.set_state(psi0.astype(np.float32), pi0.astype(np.float32))
        S.run_checkpointed(*coef)
        h0, h1 = S.vjp_run(pb.astype(np.float32), mb.astype(np.float32))
    assert G.rel(h0, g0) < 1e-3 and G.rel(h1, g1) < 1e-3, (G.rel(h0, g0), G.rel(h1, g1))
