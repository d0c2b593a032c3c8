"""Parity of the HIP library (through its C ABI) with the CPU oracle and the reference's golden vectors.

Bars: integer work (cell assignment, fixed-point deposit) bit-exact; FP64 device runs at the
reference's golden tolerances (1e-12 acc, 1e-11 trajectories); FP32 runs within the stated
tolerances against the FP64 oracle.
"""
import numpy as np
import pytest

import golden as G
import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")


def _cfg(name):
    kw = G.kwargs(name)
    return D.PMConfig(res_pm=kw["res_pm"], n_part=G.N_PART, boxsize=G.BOXSIZE, worder=kw["worder"],
                      deconvolve=kw["deconvolve"], antialias=kw["antialias"], grad_order=kw["grad_kernel_order"],
                      lap_order=kw["laplace_kernel_order"], n_resample=kw["n_resample"],
                      resampling=kw["resampling_method"])


RESAMPLE = [c for c in G.CONFIGS if G.kwargs(c)["n_resample"] > 1]


# ------------------------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [16, 24, 32, 48, 64, 96, 128, 256, 512])
@pytest.mark.parametrize("prec", [32, 64])
def test_fft3d_vs_numpy(N, prec):
    rng = np.random.default_rng(N)
    x = rng.normal(size=(N, N, N)).astype(np.float32 if prec == 32 else np.float64)
    ref = np.fft.rfftn(x.astype(np.float64), axes=(0, 1, 2))
    got = D.fft3d(x, -1)
    tol = 5e-6 if prec == 32 else 1e-13
    assert np.max(np.abs(got - ref)) < tol * np.max(np.abs(ref))
    back = D.fft3d(got, +1) / N**3
    assert np.max(np.abs(back - x)) < tol * np.max(np.abs(x))


def test_fft3d_1024_fp32_vs_fp64():
    """The 1024^3 mesh (its FP32 c2r x pass runs 16 lines per workgroup) against the FP64 device transform
    of the same field, and the FP32 round trip (numpy's FFT of 1024^3 would take minutes)."""
    N = 1024
    x = np.random.default_rng(5).normal(size=(N, N, N)).astype(np.float32)
    ref = D.fft3d(x.astype(np.float64), -1)
    got = D.fft3d(x, -1)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) < 5e-6 * scale
    del ref
    back = D.fft3d(got, +1)
    assert np.max(np.abs(back / N**3 - x)) < 5e-6 * np.max(np.abs(x))


# -------------------------------------------------------------------------------- mass assignment
def _random_psi(n, L, scale, seed, dtype):
    rng = np.random.default_rng(seed)
    psi = rng.normal(scale=scale, size=(3, n, n, n))
    # edge cases: exact cell faces, exact box edges, negative positions, half-cell ties
    psi.reshape(3, -1)[:, :8] = [[-1e-7, 0.0, L / 2, -L, 3 * L, 0.5 * L / n, -0.25 * L / n, 1e-30]] * 3
    return psi.astype(dtype)


@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("prec", [32, 64])
def test_fixed_point_deposit_bit_exact(W, prec):
    n, res, L = 32, 64, 100.0
    dt = np.float32 if prec == 32 else np.float64
    psi = _random_psi(n, L, 4.0, W, dt)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W)
    with D.PMSolverHIP(cfg, precision=prec, deposit="fixed64") as S:
        for shift in (None, [0.5 * L / res] * 3):
            got = S.deposit(psi, shift=shift)
            X = O.lagrangian_positions(psi, L)
            if shift is not None:
                X = X + np.array(shift, np.float64).astype(dt)[:, None, None, None]
            ref = O.pm_scatter_fixed(X, res, dt(L), W, 32)
            assert np.array_equal(got, ref), "cell assignment or weights differ from the oracle"
            assert got.sum() == pytest.approx(n**3 * 2.0**32, rel=1e-9)


@pytest.mark.parametrize("W", [2, 3, 4])
def test_float_deposit(W):
    n, res, L = 32, 64, 100.0
    psi = _random_psi(n, L, 4.0, 10 + W, np.float32)
    with D.PMSolverHIP(D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W), precision=32) as S:
        got = S.deposit(psi)
        again = S.deposit(psi)
    ref = O.pm_scatter(O.lagrangian_positions(psi, L).astype(np.float64), res, L, W)
    # FP32 weights summed as integers of 2^-31 (order-independent, so two deposits agree bit for bit)
    # and rounded once to float; the difference to the FP64 deposit is the FP32 grid coordinate
    # g = X / L * N, whose ulp (7.6e-6 at g ~ 64) enters each weight directly
    assert np.array_equal(got, again)
    assert np.max(np.abs(got - ref)) < 1e-5 * max(1.0, np.max(ref))
    assert got.sum() == pytest.approx(n**3, rel=1e-6)


# ------------------------------------------------------------------ PM force vs golden (FP64 device)
@pytest.mark.parametrize("name", G.CONFIGS)
def test_acc_ini_fp64_golden(name):
    acc = D.pm_acceleration(G.particles("ics/psi_ini"), _cfg(name), precision=64)
    assert G.rel(acc, G.particles(f"{name}/acc_ini")) < 1e-12


@pytest.mark.parametrize("name", G.CONFIGS)
def test_run_nbody_fp64_golden(name):
    kw = G.kwargs(name)
    c = D.Cosmology.named("Planck18EEBAOSN")
    tv = kw["time_var"]
    tv = np.array(tv, np.float64) if isinstance(tv, list) else tv
    if kw["ic_method"] == "bullfrog":
        # nbody_parity_tests.jl:52-61: ICs from one BullFrog step with 2^3 Fourier sheet resampling
        p0, m0 = D.bullfrog_ics(c, G.particles("ics/psi1"), kw["a_ini"], boxsize=G.BOXSIZE, res_pm=kw["res_pm"])
    else:
        p0, m0 = G.particles("ics/psi_ini"), G.particles("ics/pi_ini")
    psi, P, _ = D.run_nbody(c, p0, m0, boxsize=G.BOXSIZE,
                            a_ini=kw["a_ini"], a_end=kw["a_end"], n_steps=kw["n_steps"], res_pm=kw["res_pm"],
                            time_var=tv, stepper=kw["stepper"], antialias=kw["antialias"],
                            grad_kernel_order=kw["grad_kernel_order"],
                            laplace_kernel_order=kw["laplace_kernel_order"], worder=kw["worder"],
                            deconvolve=kw["deconvolve"], n_resample=kw["n_resample"],
                            resampling_method=kw["resampling_method"], return_displacement=True, precision=64)
    assert G.rel(psi, G.particles(f"{name}/psi_out")) < 1e-11
    assert G.rel(P, G.particles(f"{name}/p_out")) < 1e-11


@pytest.mark.parametrize("W", [2, 3, 4])
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("prec", [32, 64])
def test_tile_pipeline_vs_oracle(W, order, prec):
    """Tiled deposit and LDS-staged FD gather on a mesh larger than a tile's staged region (64^3: the
    golden 32^3 meshes are read from global memory in FP32), interlaced, with edge-case positions,
    against the oracle's FP64 solver-form force of the same positions."""
    n, res, L = 32, 64, 100.0
    dt = np.float32 if prec == 32 else np.float64
    psi = _random_psi(n, L, 3.0, 100 + 10 * W + order, dt)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=order, lap_order=order, antialias=1)
    acc = D.pm_acceleration(psi, cfg, precision=prec)
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=order, lap_order=order, antialias=1)
    ref = O.pm_acceleration(psi.astype(np.float64), ocfg, "solver")
    # FP64: same arithmetic up to association; FP32: mesh, FFT and gather in single precision
    assert G.rel(acc, ref) < (1e-12 if prec == 64 else 2e-5)


@pytest.mark.parametrize("W, order, aa", [(3, 6, 1), (3, 6, 0), (3, 6, 3), (2, 4, 1), (4, 2, 2), (4, 6, -1)])
def test_tile_pipeline_large_mesh(W, order, aa):
    """A 128^3 mesh, where most tiles' staged regions do not wrap around the box (the gather's
    incremental staging addresses), for shift sets whose reach fits the compile-time staging geometry
    of the FP32 tiles (antialias -1, 0, 1, 2) and one that does not (antialias 3: quarter-cell shifts)."""
    n, res, L = 64, 128, 200.0
    psi = _random_psi(n, L, 4.0, 7 + 10 * W + order + aa, np.float32)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=order, lap_order=order, antialias=aa)
    acc = D.pm_acceleration(psi, cfg, precision=32)
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=order, lap_order=order, antialias=aa)
    ref = O.pm_acceleration(psi.astype(np.float64), ocfg, "solver")
    assert G.rel(acc, ref) < 2e-5


@pytest.mark.parametrize("res, lap", [(256, 0), (256, 6), (512, 4), (1024, 2)])
def test_persistent_z_pass_meshes(res, lap):
    """FP32 meshes of 256^3 to 1024^3 run the Poisson solve's z pass as a persistent kernel that keeps
    the next block's lines in flight (exact Laplacian and FD Laplacians): the force against the FP64
    device force, whose z pass runs one block per workgroup."""
    n, L = 64, 300.0
    psi = _random_psi(n, L, 4.0, res + lap, np.float32)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=4, lap_order=lap, antialias=1)
    acc = D.pm_acceleration(psi, cfg, precision=32)
    ref = D.pm_acceleration(psi.astype(np.float64), cfg, precision=64)
    # FP32 round-off of the larger transforms: about 2e-5 of the largest force (a block transformed
    # from the wrong lines would differ at order one)
    assert G.rel(acc, ref) < 1e-4


@pytest.mark.parametrize("n, nr, method, W, grad", [
    (32, 2, "fourier", 4, 0), (32, 3, "fourier", 3, 4), (24, 2, "fourier", 2, 6), (32, 2, "linear", 3, 4),
    (24, 3, "linear", 4, 0)])
@pytest.mark.parametrize("prec", [32, 64])
def test_resampling_vs_oracle(n, nr, method, W, grad, prec):
    """Sheet resampling (nr^3 - 1 extra point sets per interlacing pass) on meshes of 2n cells, against
    the oracle's FP64 solver-form force; the transforms on the particle grid run through the library's
    FFT passes (n = 24: radix-3 sizes)."""
    res, L = 2 * n, 100.0
    dt = np.float32 if prec == 32 else np.float64
    psi = _random_psi(n, L, 2.0, 7 * nr + W, dt)
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=grad, antialias=1,
                     deconvolve=grad == 0, n_resample=nr, resampling=method)
    acc = D.pm_acceleration(psi, cfg, precision=prec)
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=grad, antialias=1, deconvolve=grad == 0, n_resample=nr,
                       resampling=method)
    ref = O.pm_acceleration(psi.astype(np.float64), ocfg, "solver")
    assert G.rel(acc, ref) < (1e-12 if prec == 64 else 5e-5)


@pytest.mark.parametrize("name", G.CONFIGS)
def test_acc_ini_fp32(name):
    # FP32 mesh / FFT / forces against the FP64 golden acceleration
    acc = D.pm_acceleration(G.particles("ics/psi_ini").astype(np.float32), _cfg(name), precision=32)
    assert G.rel(acc, G.particles(f"{name}/acc_ini")) < 2e-4


def test_lean_equals_solver_and_errors():
    n, res, L = 24, 48, 120.0
    psi = 3 * np.random.default_rng(5).normal(size=(3, n, n, n))
    for W, g in ((2, 4), (3, 2), (4, 6)):
        cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=g)
        a = D.pm_acceleration(psi, cfg, precision=64)
        with D.PMSolverHIP(cfg, precision=64, lean=True) as S:
            b = S.pm_acceleration(psi)
        assert G.rel(b, a) < 1e-12
        ref = O.pm_acceleration(psi, O.pm_config(res, n, L, worder=W, grad_order=g), "lean")
        assert G.rel(b, ref) < 1e-12
    with pytest.raises(D.DJError):
        D.PMSolverHIP(D.PMConfig(res_pm=res, n_part=n, boxsize=L, antialias=1), lean=True)


# ---------------------------------------------------------------- determinism and failure handling
def test_fixed_deposit_run_is_bit_reproducible():
    n, res, L = 32, 64, 200.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    psi = _random_psi(n, L, 2.0, 3, np.float32)
    outs = []
    for _ in range(2):
        out = D.run_nbody(c, psi, psi / 0.05, boxsize=L, a_ini=0.05, a_end=0.5, n_steps=4, res_pm=res,
                          return_displacement=True, deposit="fixed64", antialias=1)
        outs.append(out)
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_callback_when_and_resume():
    # callback fires at exactly the selected steps (pm_lean_tests.jl:30-33); a run resumed from a
    # callback checkpoint equals the uninterrupted one bit for bit in the fixed-point mode
    n, L = 16, 100.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    psi = _random_psi(n, L, 0.5, 2, np.float32)
    pi = psi / 0.05
    seen, snaps = [], {}
    cfg = D.PMConfig(res_pm=32, n_part=n, boxsize=L)
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 0.5, 10)
    with D.PMSolverHIP(cfg, deposit="fixed64") as S:
        S.set_state(psi, pi)
        want = np.zeros(11, np.uint8)
        want[[0, 4, 10]] = 1
        S.run(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"], want=want,
              callback=lambda k: (seen.append(k), snaps.__setitem__(k, S.get_state())))
        full = S.get_state()
    assert seen == [0, 4, 10]
    with D.PMSolverHIP(cfg, deposit="fixed64") as S:
        S.set_state(*snaps[4])
        S.run(co["ddrift1"][4:], co["alpha"][4:], co["beta"][4:], co["ddrift2"][4:])
        resumed = S.get_state()
    assert np.array_equal(resumed[0], full[0]) and np.array_equal(resumed[1], full[1])


def test_nonfinite_state_detected():
    n = 16
    with D.PMSolverHIP(D.PMConfig(res_pm=32, n_part=n, boxsize=100.0)) as S:
        psi = np.zeros((3, n, n, n), np.float32)
        S.set_state(psi)
        S.check_finite()
        psi[1, 3, 4, 5] = np.nan
        S.set_state(psi)
        with pytest.raises(D.DJError):
            S.check_finite()


# -------------------------------------------------------------------- FP32 trajectory vs FP64 oracle
def test_fp32_run_vs_fp64_oracle_c1_shape():
    # 32^3 particles on a 64^3 mesh, 2LPT ICs from the golden fphi upsampled is not available, so
    # the ICs are the oracle's 2LPT of an N-GenIC seed-42 field (the C1 recipe at reduced size)
    n, res, L = 32, 64, 125.0
    c = D.Cosmology.named("Planck18EEBAOSN")
    oc = O.Cosmology.named("Planck18EEBAOSN")
    k, pk = O.linear_power_spectrum(oc)
    w = O.ngenic_wnoise_real(42, n)
    fphi = O.white_noise_to_fphi(O.ic_operator(n, L, k, pk), w)
    psi1, psi2 = O.compute_core2(fphi, L)
    psi, pi = O.nbody_ics_lpt2(oc, psi1, psi2, 0.05)
    ref, _ = O.run_nbody(oc, psi, pi, O.pm_config(res, n, L, worder=2, grad_order=4, antialias=1),
                         0.05, 1.0, 10)
    got, _, _ = D.run_nbody(c, psi.astype(np.float32), pi.astype(np.float32), boxsize=L, a_ini=0.05,
                            a_end=1.0, n_steps=10, res_pm=res, antialias=1, return_displacement=True)
    rms = np.sqrt(np.mean((got.astype(np.float64) - ref) ** 2)) / (L / n)
    assert rms < 1e-3


def test_snapshots_through_the_step_callback(tmp_path):
    """run_nbody's callback writes snapshots at chosen step times; the last equals the returned state
    and a run restarted from a middle snapshot reaches the same final state."""
    from discodj_amd import snapshots as SN
    kw = G.kwargs("fastpm_D_tsc")
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, kw["stepper"], kw["a_ini"], kw["a_end"], kw["n_steps"], "D")
    targets = [co["all_a"][3], co["all_a"][-1]]
    w = SN.SnapshotWriter(str(tmp_path), targets, G.BOXSIZE)
    run = dict(boxsize=G.BOXSIZE, res_pm=kw["res_pm"], stepper=kw["stepper"], worder=kw["worder"],
               grad_kernel_order=kw["grad_kernel_order"], return_displacement=True, precision=64)
    psi, P, _ = D.run_nbody(c, G.particles("ics/psi_ini"), G.particles("ics/pi_ini"), a_ini=kw["a_ini"],
                            a_end=kw["a_end"], n_steps=kw["n_steps"], step_callback=w, callback_when=w.when, **run)
    assert len(w.written) == 2
    last = SN.read_snapshot(w.written[-1])
    assert np.allclose(last["psi"], psi, rtol=0, atol=1e-12) and np.array_equal(last["P"], P)
    mid = SN.read_snapshot(w.written[0])
    assert mid["step"] == 3
    # restart: canonical P back to Pi, the remaining steps on the same grid
    co2 = {k: np.asarray(v)[3:] for k, v in co.items() if k != "all_a"}
    with D.PMSolverHIP(D.PMConfig(res_pm=kw["res_pm"], n_part=G.N_PART, boxsize=G.BOXSIZE, worder=kw["worder"],
                                  grad_order=kw["grad_kernel_order"]), precision=64) as S:
        S.set_state(mid["psi"], mid["P"] / c.Fplus(mid["a"]))
        S.run(co2["ddrift1"], co2["alpha"], co2["beta"], co2["ddrift2"])
        psi2, _ = S.get_state()
    assert G.rel(psi2, psi) < 1e-12


def test_plane_wave_linear_growth_on_the_device():
    """nbody_parity_tests.jl:68-90 on the device (FP64): a tiny fundamental plane wave follows D(a) under
    every stepper (BullFrog 1 and 16 steps, FastPM 16, symplectic 64), stays one-dimensional, and its
    growth error converges with the lattice (16^3 -> 32^3 at least 3x smaller); the device run also
    equals the oracle's."""
    c = D.Cosmology.named("Planck18EEBAOSN")
    oc = O.Cosmology.named("Planck18EEBAOSN")
    L, a0 = 100.0, 0.05
    D0, D1 = c.Dplus(a0), c.Dplus(1.0)
    err = {}
    runs = (("bullfrog", 1), ("bullfrog", 16), ("fastpm", 16), ("symplectic", 64))
    for n in (16, 32):
        q = np.arange(n) * L / n
        psi1 = np.zeros((3, n, n, n))
        psi1[0] = 1e-3 * np.sin(2 * np.pi * q / L)[None, None, :]
        for st, ns in runs:
            psi, _, _ = D.run_nbody(c, D0 * psi1, psi1.copy(), boxsize=L, a_ini=a0, a_end=1.0, n_steps=ns,
                                    res_pm=2 * n, stepper=st, return_displacement=True, precision=64)
            assert np.max(np.abs(psi[1:])) < 1e-12
            err[(n, st, ns)] = abs(np.sum(psi[0] * psi1[0]) / np.sum(psi1[0] ** 2) / D1 - 1)
            if n == 16:
                ref, _ = O.run_nbody(oc, D0 * psi1, psi1.copy(), O.pm_config(res_pm=2 * n, n_part=n, boxsize=L),
                                     a0, 1.0, ns, stepper=st)
                assert G.rel(psi, ref) < 1e-11
    for st, ns in runs:
        assert err[(32, st, ns)] < 1e-2
        assert err[(16, st, ns)] / err[(32, st, ns)] > 3
