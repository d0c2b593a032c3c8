"""Learned force correction: the device net (MFMA BF16 convolutions) against the numpy oracle, its
parameter gradients against the oracle's, and what training gains over plain PM; the fitted transfer
function of the Poisson kernel (identity, linearity, argument checks, held-out P(k) after the fit)."""
import numpy as np
import pytest

import nn_ref as R

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")


def _blockrel(x, y):
    """max over layers of max|x - y| / max|y| within the layer's real parameters"""
    out = []
    for l in range(4):
        a = x[l * R.PER_LAYER:(l + 1) * R.PER_LAYER]
        b = y[l * R.PER_LAYER:(l + 1) * R.PER_LAYER]
        out.append(float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    return max(out)


def _density(N, n, seed):
    rng = np.random.default_rng(seed)
    return (rng.lognormal(sigma=1.0, size=(N, N, N)) * (n / N) ** 3).astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2])
def test_nn_forward_and_gradient_match_oracle(seed):
    n, N, L, a, s_out = 32, 64, 100.0, 0.4, 0.7
    params = D.learned.init_params(seed, out_std=1.0)
    params[27 * 256:R.PER_LAYER] = np.random.default_rng(seed).normal(scale=0.1, size=16)  # biases of layer 1
    dens = _density(N, n, seed)
    gphi = np.random.default_rng(seed + 10).normal(size=(N, N, N)).astype(np.float32)
    with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=n, boxsize=L), precision=32) as S:
        S.nn_set_params(params)
        S.nn_enable(True, s_out)
        out = S.nn_forward(dens, a)
        grad = S.nn_backward(dens, a, gphi)
    ref, cache = R.forward(params, dens, (N / n) ** 3, a, s_out)
    gref = R.backward(cache, gphi, s_out)
    assert float(np.max(np.abs(out - ref)) / np.max(np.abs(ref))) < 1e-2
    assert _blockrel(grad, gref) < 2e-2


def test_training_reduces_force_error():
    """The learned-correction workload at reduced size: 64^3 particles on a 64^3 mesh (window
    deconvolution) against a 256^3 mesh, BullFrog 10 steps; the net trained on six seeds (every step),
    evaluated on a seventh.  Asserted: the held-out force error falls; the net alone raises the P(k)
    error above the deconvolved PM's and misses 2 % (a force-MSE optimum is smoother than the fine
    force), within a bounded factor; a transfer function fitted on top of the net (two seeds) takes back
    most of what the net loses (to 3.1 % on this seed; the transfer function alone meets 2 %,
    test_fitted_transfer_function_matches_fine_mesh_power)."""
    c = D.Cosmology.named("Planck18EEBAOSN")
    work = D.learned.Workload(c, n=64, N=64, L=128.0, ratio=4, n_steps=10, deconvolve=True)
    data = work.snapshots([1, 2, 4, 5, 6, 7], every=1)
    params, scale, losses = D.learned.train(work, data, iters=8000, lr=3e-3)
    # losses are relative to the uncorrected force error at each snapshot (1 = no gain)
    assert np.mean(losses[-500:]) < 0.9, np.mean(losses[-500:])
    held = work.snapshots([3], every=3)
    e_plain, e_nn = D.learned.force_errors(work, held, params, scale)
    ratio = e_nn / e_plain
    print("held-out force MSE, corrected / plain, per snapshot:", np.round(ratio, 3))
    assert np.mean(ratio) < 0.9 and np.all(ratio < 1.0), ratio
    fine = D.learned.fine_states(work, [1, 2, 3])
    pk_plain, pk_nn, _ = D.learned.power_spectrum_errors(work, 3, params, scale, psi_fine=fine[2])
    g, kmax = D.learned.fit_transfer(work, [1, 2], psi_fine=fine[:2], params=params, out_scale=scale)
    _, pk_both, _ = D.learned.power_spectrum_errors(work, 3, params, scale, kgain=(g, kmax), psi_fine=fine[2])
    print(f"max |P/P_fine - 1| to k_Nyq/2: deconvolved {pk_plain:.4f}, net {pk_nn:.4f}, net + transfer "
          f"function {pk_both:.4f}")
    # what the net itself does to P(k), stated as it is: trained on forces, it raises the P(k) error of the
    # deconvolved PM (0.034 -> 0.071 on this seed: its force-MSE optimum removes small-scale power) and
    # misses the 2 % bar; the rise is bounded, and the transfer function fitted on top of the net takes
    # the error back to the deconvolved PM's level
    assert pk_plain < pk_nn < 4 * pk_plain and pk_nn > 0.02, (pk_plain, pk_nn)
    assert pk_both < 0.5 * pk_nn and pk_both < 0.05


# ------------------------------------------------------------------ fitted transfer function
def _psi(n, L, seed):
    return np.random.default_rng(seed).normal(scale=2.0, size=(3, n, n, n)).astype(np.float32)


def test_training_repeats_bit_for_bit():
    """The weight gradient is summed in a fixed order (per-block partial sums, then the blocks in order) and
    the force cotangent is deposited in fixed point, so two training runs from the same start give the same
    parameters and losses bit for bit."""
    c = D.Cosmology.named("Planck18EEBAOSN")
    work = D.learned.Workload(c, n=32, N=64, L=64.0, ratio=2, n_steps=4, deconvolve=True)
    data = work.snapshots([1, 2], every=2)
    p1, s1, l1 = D.learned.train(work, data, iters=40, lr=3e-3)
    p2, s2, l2 = D.learned.train(work, data, iters=40, lr=3e-3)
    assert s1 == s2
    assert np.array_equal(p1, p2)
    assert np.array_equal(l1, l2)
    assert not np.array_equal(p1, D.learned.init_params(0))


@pytest.mark.parametrize("prec", [32, 64])
def test_transfer_function_identity_and_linearity(prec):
    """g = 1 leaves the force as it is; g = 2 doubles it (every later step is linear and the factor is a
    power of two); a g rising beyond kmax / 2 changes it.  FP32 to the bit (its deposit sums integers,
    the same in every call); FP64 to 1e-12 (its LDS float atomics sum in a varying order)."""
    n, N, L = 32, 64, 100.0
    dt = np.float32 if prec == 32 else np.float64
    psi = _psi(n, L, 4).astype(dt)
    cfg = D.PMConfig(res_pm=N, n_part=n, boxsize=L, worder=3, grad_order=4, antialias=1)
    kmax = np.pi * N / L

    def same(a, b):
        return np.array_equal(a, b) if prec == 32 else np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))

    with D.PMSolverHIP(cfg, precision=prec) as S:
        plain = S.pm_acceleration(psi).copy()
        S.set_kgain(np.ones(17), kmax)
        assert same(S.pm_acceleration(psi), plain)
        S.set_kgain(np.full(5, 2.0), kmax)
        assert same(S.pm_acceleration(psi), 2 * plain)
        S.set_kgain(None)
        assert same(S.pm_acceleration(psi), plain)
        # beyond kmax the last value holds: g = 1 up to kmax / 2, rising to 3 at kmax
        S.set_kgain(np.array([1.0, 1.0, 3.0]), kmax)
        hi = S.pm_acceleration(psi)
        assert not np.allclose(hi, plain) and np.all(np.isfinite(hi))


def test_transfer_function_rejects_bad_tables():
    with D.PMSolverHIP(D.PMConfig(res_pm=32, n_part=16, boxsize=50.0), precision=32) as S:
        for g, kmax in (([1.0], 1.0), ([1.0, 2.0], 0.0), ([1.0, np.nan], 1.0)):
            with pytest.raises(Exception):
                S.set_kgain(np.array(g), kmax)


def test_fitted_transfer_function_matches_fine_mesh_power():
    """The transfer function fitted on two seeds (64^3 particles, 64^3 mesh with window deconvolution,
    against the 256^3-mesh run) brings the held-out seed's P(k) within 2 % of the fine mesh's to
    k_Nyq/2 (plain PM: 10 %)."""
    c = D.compute_timetables(D.Cosmology(Omega_c=0.2621, Omega_b=0.049, h=0.6766, sigma8=0.8102, n_s=0.9665))
    work = D.learned.Workload(c, n=64, N=64, L=128.0, ratio=4, n_steps=10, deconvolve=True)
    fine = D.learned.fine_states(work, [1, 2, 3])
    g, kmax = D.learned.fit_transfer(work, [1, 2], psi_fine=fine[:2])
    pk_plain, pk_corr, _ = D.learned.power_spectrum_errors(work, 3, kgain=(g, kmax), psi_fine=fine[2])
    plain_work = D.learned.Workload(c, n=64, N=64, L=128.0, ratio=4, n_steps=10)
    pk_undeconvolved, _, _ = D.learned.power_spectrum_errors(plain_work, 3, psi_fine=fine[2])
    print(f"max |P/P_fine - 1| to k_Nyq/2: plain {pk_undeconvolved:.4f}, deconvolved {pk_plain:.4f}, "
          f"fitted transfer function {pk_corr:.4f}")
    assert pk_corr < 0.02 < pk_undeconvolved
