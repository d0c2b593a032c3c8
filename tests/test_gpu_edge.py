"""Edge cases of the tiled pipeline against the oracle: extreme clustering (one tile holding a large share
of the particles: list overflow, the exact recount, tiles with several chunks of particles), meshes whose
size is not a power of two (radix-3 FFT sizes with power-of-two tiles), sparse particles (empty tiles),
and the fixed-point deposit's bit-exactness under clustering."""
import numpy as np
import pytest

import golden as G
import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")


def _clustered(n, L, frac, radius, seed):
    """frac of the particles moved into a ball of the given radius around the box centre"""
    rng = np.random.default_rng(seed)
    psi = rng.normal(scale=0.3, size=(3, n, n, n))
    q = np.arange(n) * (L / n)
    X = np.stack([np.broadcast_to(q[None, None, :], (n, n, n)), np.broadcast_to(q[None, :, None], (n, n, n)),
                  np.broadcast_to(q[:, None, None], (n, n, n))])
    sel = rng.random((n, n, n)) < frac
    target = L / 2 + rng.normal(scale=radius, size=(3, n, n, n))
    psi = np.where(sel[None], target - X, psi)
    return psi


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("W, order", [(2, 4), (3, 6), (4, 2)])
def test_extreme_clustering_vs_oracle(prec, W, order):
    n, res, L = 32, 64, 100.0
    psi = _clustered(n, L, 0.6, 0.8, 10 * W + order)
    dt = np.float32 if prec == 32 else np.float64
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=W, grad_order=order, lap_order=order, antialias=1)
    with D.PMSolverHIP(cfg, precision=prec) as S:
        acc = S.pm_acceleration(psi.astype(dt))
        again = S.pm_acceleration(psi.astype(dt))
    ocfg = O.pm_config(res, n, L, worder=W, grad_order=order, lap_order=order, antialias=1)
    ref = O.pm_acceleration(psi, ocfg, "solver")
    assert G.rel(acc, ref) < (1e-12 if prec == 64 else 5e-5)
    assert G.rel(again, ref) < (1e-12 if prec == 64 else 5e-5)


def test_clustered_fixed_deposit_bit_exact():
    n, res, L = 32, 64, 100.0
    psi = _clustered(n, L, 0.7, 0.5, 3)
    with D.PMSolverHIP(D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3), precision=64, deposit="fixed64") as S:
        got = S.deposit(psi)
    ref = O.pm_scatter_fixed(O.lagrangian_positions(psi, L), res, L, 3, 32)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n, res", [(48, 96), (24, 96), (16, 48)])
def test_radix3_meshes_vs_oracle(n, res):
    L = 120.0
    psi = np.random.default_rng(res + n).normal(scale=1.5, size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=4, antialias=1)
    acc = D.pm_acceleration(psi, cfg, precision=64)
    ref = O.pm_acceleration(psi, O.pm_config(res, n, L, worder=3, grad_order=4, antialias=1), "solver")
    assert G.rel(acc, ref) < 1e-12


def test_sparse_particles_empty_tiles():
    n, res, L = 8, 128, 64.0
    psi = np.random.default_rng(1).normal(scale=0.5, size=(3, n, n, n))
    cfg = D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=2, grad_order=2)
    for prec in (32, 64):
        acc = D.pm_acceleration(psi.astype(np.float32 if prec == 32 else np.float64), cfg, precision=prec)
        ref = O.pm_acceleration(psi, O.pm_config(res, n, L, worder=2, grad_order=2), "solver")
        assert G.rel(acc, ref) < (1e-12 if prec == 64 else 2e-5)


def test_clustered_run_stays_finite_and_reproducible():
    """Ten steps from a clustered state: the tile lists grow and shrink between steps (reserved lists,
    recounts); the fixed-point run is bit-reproducible."""
    n, res, L = 32, 64, 100.0
    psi = _clustered(n, L, 0.5, 1.0, 7)
    pi = np.zeros_like(psi)
    c = D.Cosmology.named("Planck18EEBAOSN")
    co = D.stepper_coefficients(c, "bullfrog", 0.3, 1.0, 10, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    out = []
    for _ in range(2):
        with D.PMSolverHIP(D.PMConfig(res_pm=res, n_part=n, boxsize=L, worder=3, grad_order=4, antialias=1),
                           precision=32, deposit="fixed64") as S:
            S.set_state(psi.astype(np.float32), pi.astype(np.float32))
            S.run(*coef)
            S.check_finite()
            out.append(S.get_state()[0])
    assert np.array_equal(out[0], out[1])
