"""Learned force correction of the PM step: training data, training and evaluation on the device.

Two learned parts, used alone or together: a fitted transfer function g(|k|) of the Poisson kernel
(fit_transfer: P(k) of the coarse run matched to the fine-mesh run's), and a residual convolution net
that adds a potential correction (train: force matched to the fine-mesh force at snapshots).

The corrected force is that of the PM solver on an N^3 mesh with a residual convolution net's potential
correction (dj_nn_*): it is trained on pairs of runs the method generates itself, the same particles and
ICs on the N^3 mesh and on a finer mesh (4N)^3, by minimising the mean squared difference between the
corrected N^3 force and the fine-mesh force at snapshots of the fine run.  Plain PM stays the fallback
(nn_enable(False)) and the baseline of the correction's P(k) error.  There is no counterpart in the reference.
"""
import numpy as np

from .cosmology import linear_power_spectrum, stepper_coefficients
from .pm import PMConfig, PMSolverHIP, check_solver
from ._lib import NN_PARAMS

PER_LAYER = 27 * 16 * 16 + 16


def init_params(seed=0, out_std=1e-2):
    """He-normal hidden layers (fan-in 27 x inputs), a small output layer, zero biases and padding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NN_PARAMS, np.float32)
    for l, (co, ci) in enumerate([(16, 2), (16, 16), (16, 16), (1, 16)]):
        std = np.sqrt(2.0 / (27 * ci)) if l < 3 else out_std * np.sqrt(1.0 / (27 * ci))
        W = np.zeros((27, 16, 16), np.float32)
        W[:, :co, :ci] = rng.normal(scale=std, size=(27, co, ci))
        p[l * PER_LAYER:l * PER_LAYER + 27 * 256] = W.reshape(-1)
    return p


def force_scale_factors(cosmo, co, time_var="D"):
    """Scale factor at each step's force evaluation (the midpoint of the step in the time variable D;
    the geometric mean of the step's ends otherwise)."""
    a = np.asarray(co["all_a"], np.float64)
    if time_var == "D":
        D = np.array([cosmo.Dplus(x) for x in a])
        return np.array([cosmo.a_of_Dplus(0.5 * (D[k] + D[k + 1])) for k in range(len(a) - 1)])
    return np.sqrt(a[:-1] * a[1:])


class Workload:
    """Particles n^3 in a box L, the coarse mesh N and the fine mesh ratio*N, a stepper from a_ini to a_end."""

    def __init__(self, cosmo, n, N, L, ratio=4, n_steps=10, a_ini=0.05, a_end=1.0, stepper="bullfrog",
                 worder=2, grad_order=4, deconvolve=False):
        self.cosmo, self.n, self.N, self.L, self.ratio = cosmo, n, N, float(L), ratio
        self.co = stepper_coefficients(cosmo, stepper, a_ini, a_end, n_steps, "D")
        self.coef = [self.co["ddrift1"], self.co["alpha"], self.co["beta"], self.co["ddrift2"]]
        self.a_force = force_scale_factors(cosmo, self.co)
        self.D_ini = cosmo.Dplus(a_ini)
        self.k_tab, self.pk_tab = linear_power_spectrum(cosmo)
        kw = dict(n_part=n, boxsize=self.L, worder=worder, grad_order=grad_order)
        # deconvolve: the coarse solve divides by the mass-assignment window (pm.jl:150-151), and the net
        # learns the residual; the fine-mesh reference stays plain PM
        self.cfg_lo = PMConfig(res_pm=N, deconvolve=deconvolve, **kw)
        self.cfg_hi = PMConfig(res_pm=ratio * N, **kw)

    def run(self, S, seed, want=None, cb=None):
        """ICs of seed (2LPT, N-GenIC noise) and the steps on solver S."""
        S.ics_lpt(self.k_tab, self.pk_tab, self.D_ini, order=2, seed=seed)
        S.run(*self.coef, want=want, callback=cb)

    def snapshots(self, seeds, every=1):
        """Fine-mesh trajectories: (psi, acc_fine, a) at the start of every `every`-th step."""
        out = []
        with PMSolverHIP(self.cfg_hi, precision=32) as S:
            for seed in seeds:
                steps = len(self.coef[0])
                want = np.zeros(steps + 1, np.uint8)
                want[:steps:every] = 1

                def cb(k, S=S):
                    psi, _ = S.get_state()
                    out.append((psi, S.pm_acceleration(psi), float(self.co["all_a"][k])))

                self.run(S, seed, want, cb)
        return out


def train(work, data, params=None, iters=600, lr=1e-3, seed=0, solver=None, log=None, cosine=True):
    """Adam on the coarse-mesh corrected force against the fine-mesh force of the snapshots in data;
    -> (params, out_scale, losses).  Each snapshot's squared error is weighted by the inverse of the
    plain coarse PM's (the force error grows by orders of magnitude from the first snapshot to the last),
    so a loss of 1 is no gain and the losses are relative.  out_scale = rms force difference x mesh
    spacing, so that the net's output is of order one.  cosine: the learning rate decays as
    lr (1 + cos(pi t / iters)) / 2."""
    own = solver is None
    if not own:
        check_solver(solver, work.cfg_lo, 32, deposit="float")
    S = solver or PMSolverHIP(work.cfg_lo, precision=32)
    try:
        e0 = [float(np.mean((S.pm_acceleration(p).astype(np.float64) - f) ** 2)) for p, f, _ in data]
        out_scale = float(np.mean(np.sqrt(e0))) * work.L / work.N
        S.nn_set_params(init_params(seed) if params is None else params)
        S.nn_enable(True, out_scale)
        rng = np.random.default_rng(seed)
        losses = []
        for it in range(iters):
            i = rng.integers(len(data))
            p, f, a = data[i]
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / iters)) if cosine else lr
            losses.append(S.nn_train_step(p, f, a, lr_t, it + 1, weight=1.0 / e0[i]))
            if log and (it % 50 == 0 or it == iters - 1):
                log(f"iteration {it}: loss {losses[-1]:.4e}")
        return S.nn_get_params(), out_scale, np.array(losses)
    finally:
        if own:
            S.close()


def force_errors(work, data, params, out_scale):
    """Mean squared force errors against the fine mesh per snapshot: (plain coarse PM, corrected)."""
    with PMSolverHIP(work.cfg_lo, precision=32) as S:
        S.nn_set_params(params)
        e_plain, e_nn = [], []
        for p, f, a in data:
            S.nn_enable(False)
            e_plain.append(np.mean((S.pm_acceleration(p).astype(np.float64) - f) ** 2))
            S.nn_enable(True, out_scale)
            S.nn_set_schedule([a])
            e_nn.append(np.mean((S.pm_acceleration(p).astype(np.float64) - f) ** 2))
    return np.array(e_plain), np.array(e_nn)


def fine_states(work, seeds):
    """Displacements at a_end of the fine-mesh runs of seeds."""
    out = []
    with PMSolverHIP(work.cfg_hi, precision=32) as S:
        for seed in seeds:
            work.run(S, seed)
            out.append(S.get_state()[0])
    return out


def fit_transfer(work, seeds, iters=8, power=0.5, psi_fine=None, params=None, out_scale=None, log=None):
    """Transfer function g(|k|) of the coarse Poisson kernel fitted on the device so that the coarse run's
    P(k) at a_end matches the fine-mesh run's for the training seeds (all P(k) with the coarse mesh's CIC
    estimator, linear bins of width kf up to k_Nyq/2): fixed-point updates g <- g (P_fine / P_coarse)^power
    averaged over the seeds, on the grid |k| = i kf, held at its last value beyond k_Nyq/2.  With params
    (and out_scale) the coarse runs include the conv net, and g is fitted on top of it.
    -> (g, kmax) for PMSolverHIP.set_kgain."""
    psi_fine = psi_fine if psi_fine is not None else fine_states(work, seeds)
    with PMSolverHIP(work.cfg_lo, precision=32) as S:
        if params is not None:
            S.nn_set_params(params)
            S.nn_enable(True, out_scale)
            S.nn_set_schedule(work.a_force)
        spectra = [S.power_spectrum(p, W=2) for p in psi_fine]
        k, nm = spectra[0][0], spectra[0][2]
        ok = nm > 0
        kf = 2 * np.pi / work.L
        nb = int(np.ceil(k[ok][-1] / kf)) + 1
        kmax = (nb - 1) * kf
        grid = np.arange(nb) * kf
        g = np.ones(nb)
        for it in range(iters):
            S.set_kgain(g, kmax)
            ratio = np.zeros(ok.sum())
            for seed, (_, p_hi, _) in zip(seeds, spectra):
                work.run(S, seed)
                _, p_lo, _ = S.power_spectrum(W=2)
                ratio += p_hi[ok] / p_lo[ok] / len(seeds)
            if log:
                log(f"transfer fit {it}: max |P_coarse / P_fine - 1| = {np.max(np.abs(1 / ratio - 1)):.4f}")
            g = g * np.interp(grid, k[ok], ratio ** power)
    return g, kmax


def power_spectrum_errors(work, seed, params=None, out_scale=None, kgain=None, psi_fine=None):
    """max |P/P_fine - 1| for kf <= k <= k_Nyq/2 of the coarse mesh at a_end, for the plain and the
    corrected coarse runs of seed (all P(k) measured with the coarse mesh's CIC estimator).  The
    correction is the net (params, out_scale), the transfer function kgain = (g, kmax), or both."""
    if psi_fine is None:
        psi_fine = fine_states(work, [seed])[0]
    with PMSolverHIP(work.cfg_lo, precision=32) as S:
        _, p_hi, nm = S.power_spectrum(psi_fine, W=2)
        work.run(S, seed)
        _, p_plain, _ = S.power_spectrum(W=2)
        if params is not None:
            S.nn_set_params(params)
            S.nn_enable(True, out_scale)
            S.nn_set_schedule(work.a_force)
        if kgain is not None:
            S.set_kgain(*kgain)
        work.run(S, seed)
        k, p_nn, _ = S.power_spectrum(W=2)
    ok = nm > 0
    return (float(np.max(np.abs(p_plain[ok] / p_hi[ok] - 1))), float(np.max(np.abs(p_nn[ok] / p_hi[ok] - 1))),
            (k[ok], p_hi[ok], p_plain[ok], p_nn[ok]))
