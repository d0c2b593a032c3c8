"""run_nbody and bullfrog_ics on the MI355X library (NR/src/nbody/run_nbody.jl:124-172, 180-225).

The host computes the FP64 stepper coefficients once; the particle state stays on the device for the
whole loop (dj_run), with callbacks only at the steps callback_when selects.
"""
import numpy as np

from .cosmology import stepper_coefficients
from .pm import PMConfig, PMSolverHIP, check_solver


def lagrangian_q(n, boxsize, dtype):
    """q[i] = (i-1) L/n in FP64, rounded to the working precision (pm.jl:186)."""
    return (np.arange(n) * (boxsize / n)).astype(dtype)


def add_q(psi, boxsize):
    """X = q + psi for (3,n,n,n) arrays [c,k,j,i]."""
    n = psi.shape[1]
    q = lagrangian_q(n, boxsize, psi.dtype)
    X = np.empty_like(psi)
    X[0] = psi[0] + q[None, None, :]
    X[1] = psi[1] + q[None, :, None]
    X[2] = psi[2] + q[:, None, None]
    return X


def run_nbody(c, psi_ini, pi_ini, *, boxsize, a_ini, a_end, n_steps, res_pm, time_var="D",
              stepper="bullfrog", antialias=0, grad_kernel_order=4, laplace_kernel_order=0, worder=2,
              deconvolve=False, n_resample=1, resampling_method="fourier", return_displacement=False,
              step_callback=None, callback_when=None, lean=False, precision=None, deposit="float",
              solver=None):
    """-> (X or Psi, P, a_end) with P = Pi * F+(a) (run_nbody.jl:164-171).  psi_ini / pi_ini are
    (3,n,n,n) arrays; precision defaults to their dtype (32 or 64)."""
    psi_ini = np.asarray(psi_ini)
    n = psi_ini.shape[1]
    if precision is None:
        precision = 64 if psi_ini.dtype == np.float64 else 32
    T = np.float64 if precision == 64 else np.float32
    co = stepper_coefficients(c, stepper, a_ini, a_end, n_steps, time_var)
    cfg = PMConfig(res_pm=res_pm, n_part=n, boxsize=float(boxsize), worder=worder, deconvolve=deconvolve,
                   antialias=antialias, grad_order=grad_kernel_order, lap_order=laplace_kernel_order,
                   n_resample=n_resample, resampling=resampling_method)
    own = solver is None
    if not own:
        check_solver(solver, cfg, precision, deposit=deposit, lean=lean)
    S = solver or PMSolverHIP(cfg, precision=precision, deposit=deposit, lean=lean)
    try:
        mom = np.asarray(pi_ini, np.float64)
        if stepper == "symplectic":  # run_nbody.jl:104
            mom = mom * c.Fplus(a_ini)
        S.set_state(np.asarray(psi_ini, T), mom.astype(T))

        def canon(M, a):  # run_nbody.jl:149
            M = M.astype(np.float64)
            if stepper == "symplectic":
                M = M / c.Fplus(a)
            return (M * c.Fplus(a)).astype(T)

        want = None
        cb = None
        if step_callback is not None:
            want = np.array([callback_when is None or bool(callback_when(k, co["all_a"][k]))
                             for k in range(n_steps + 1)], np.uint8)

            def cb(k):
                psi, mom_k = S.get_state()
                step_callback(k, co["all_a"][k], add_q(psi, boxsize), canon(mom_k, co["all_a"][k]))
        S.run(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"], want=want, callback=cb)
        psi, mom = S.get_state()
    finally:
        if own:
            S.close()
    P = canon(mom, a_end)
    if return_displacement:
        return psi, P, a_end
    X = add_q(psi, boxsize)
    L = T(boxsize)
    return np.mod(X + L, L), P, a_end


def nbody_ics_lpt(c, psi1, psi2, a):
    """EdS 2LPT initial conditions Psi = D psi1 + D^2 psi2, Pi = psi1 + 2 D psi2 (run_nbody.jl:200-206)."""
    D = c.Dplus(a)
    return (0.0 + D * psi1) + (D * D) * psi2, (0.0 + 1.0 * psi1) + (2 * D) * psi2


# the discreteness-suppressing defaults of bullfrog_ics (run_nbody.jl:219-220)
BULLFROG_IC_SETTINGS = dict(time_var="D", stepper="bullfrog", antialias=1, grad_kernel_order=0,
                            laplace_kernel_order=0, worder=4, deconvolve=True, n_resample=2,
                            resampling_method="fourier")


def bullfrog_ics(c, psi1, a_ini, *, boxsize, res_pm, precision=None, **settings):
    """-> (Psi, Pi): one BullFrog step from a = 0 to a_ini starting from 1LPT (Psi = D(0) psi1,
    Pi = psi1), PCS with deconvolution, ik gradient, -1/k^2, interlacing and 2^3 Fourier sheet
    resampling unless settings override them (run_nbody.jl:217-225).  psi1 is (3,n,n,n)."""
    s = dict(BULLFROG_IC_SETTINGS)
    s.update(settings)
    psi1 = np.asarray(psi1)
    if precision is None:
        precision = 64 if psi1.dtype == np.float64 else 32
    T = np.float64 if precision == 64 else np.float32
    psi1 = psi1.astype(T)
    psi, P, _ = run_nbody(c, T(c.Dplus(0.0)) * psi1, psi1.copy(), boxsize=boxsize, a_ini=0.0, a_end=a_ini,
                          n_steps=1, res_pm=res_pm, return_displacement=True, precision=precision, **s)
    return psi, P / T(c.Fplus(a_ini))
