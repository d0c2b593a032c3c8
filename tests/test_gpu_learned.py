"""Learned force correction: the device net (MFMA BF16 convolutions) against the numpy oracle, its
parameter gradients against the oracle's, and what training gains over plain PM."""
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
    """The learned-correction workload at reduced size: 64^3 particles on a 64^3 mesh against a 256^3
    mesh, BullFrog 10 steps; trained on six seeds (every step), judged on a seventh.  The gain asserted
    is the held-out force error; the P(k) errors to k_Nyq/2 are printed (the net narrows the gap at
    intermediate k, and the last bins stay about as far as plain PM's)."""
    c = D.Cosmology.named("Planck18EEBAOSN")
    work = D.learned.Workload(c, n=64, N=64, L=128.0, ratio=4, n_steps=10)
    data = work.snapshots([1, 2, 4, 5, 6, 7], every=1)
    params, scale, losses = D.learned.train(work, data, iters=8000, lr=3e-3)
    # losses are relative to plain PM's force error at each snapshot (1 = no gain)
    assert np.mean(losses[-500:]) < 0.8, np.mean(losses[-500:])
    held = work.snapshots([3], every=3)
    e_plain, e_nn = D.learned.force_errors(work, held, params, scale)
    ratio = e_nn / e_plain
    print("held-out force MSE, corrected / plain, per snapshot:", np.round(ratio, 3))
    assert np.mean(ratio) < 0.85 and np.all(ratio < 1.0), ratio
    pk_plain, pk_nn, _ = D.learned.power_spectrum_errors(work, 3, params, scale)
    print(f"max |P/P_fine - 1| to k_Nyq/2: plain {pk_plain:.4f}, corrected {pk_nn:.4f}")
    assert pk_nn < pk_plain + 0.02
