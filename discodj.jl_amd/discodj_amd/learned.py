"""Learned force correction of the PM step: training data, training and evaluation on the device.

The corrected force is that of the PM solver on an N^3 mesh with a residual convolution net's potential
correction (dj_nn_*): it is trained on pairs of runs the method generates itself, the same particles and
ICs on the N^3 mesh and on a finer mesh (4N)^3, by minimising the mean squared difference between the
corrected N^3 force and the fine-mesh force at snapshots of the fine run.  Plain PM stays the fallback
(nn_enable(False)) and the judge of the correction's P(k) error.  There is no counterpart in the reference.
"""
import numpy as np

from .cosmology import linear_power_spectrum, stepper_coefficients
from .pm import PMConfig, PMSolverHIP
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
                 worder=2, grad_order=4):
        self.cosmo, self.n, self.N, self.L, self.ratio = cosmo, n, N, float(L), ratio
        self.co = stepper_coefficients(cosmo, stepper, a_ini, a_end, n_steps, "D")
        self.coef = [self.co["ddrift1"], self.co["alpha"], self.co["beta"], self.co["ddrift2"]]
        self.a_force = force_scale_factors(cosmo, self.co)
        self.D_ini = cosmo.Dplus(a_ini)
        self.k_tab, self.pk_tab = linear_power_spectrum(cosmo)
        kw = dict(n_part=n, boxsize=self.L, worder=worder, grad_order=grad_order)
        self.cfg_lo = PMConfig(res_pm=N, **kw)
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


def power_spectrum_errors(work, seed, params, out_scale):
    """max |P/P_fine - 1| for kf <= k <= k_Nyq/2 of the coarse mesh at a_end, for the plain and the
    corrected coarse runs of seed (all P(k) measured with the coarse mesh's CIC estimator)."""
    with PMSolverHIP(work.cfg_hi, precision=32) as S:
        work.run(S, seed)
        psi_hi, _ = S.get_state()
    with PMSolverHIP(work.cfg_lo, precision=32) as S:
        _, p_hi, nm = S.power_spectrum(psi_hi, W=2)
        work.run(S, seed)
        _, p_plain, _ = S.power_spectrum(W=2)
        S.nn_set_params(params)
        S.nn_enable(True, out_scale)
        S.nn_set_schedule(work.a_force)
        work.run(S, seed)
        k, p_nn, _ = S.power_spectrum(W=2)
    ok = nm > 0
    return (float(np.max(np.abs(p_plain[ok] / p_hi[ok] - 1))), float(np.max(np.abs(p_nn[ok] / p_hi[ok] - 1))),
            (k[ok], p_hi[ok], p_plain[ok], p_nn[ok]))
