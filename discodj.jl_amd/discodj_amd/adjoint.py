"""Gradient of a density loss with respect to the white-noise IC field, by reverse mode on the device.

The forward chain is the path of run_nbody: white noise -> phi(k) -> 2LPT ICs (dj_ics_lpt) -> n_steps DKD
steps (dj_run_checkpointed keeps the state before every step) -> the deposit of the final positions on the
PM mesh, delta = mesh / rho_bar - 1 with rho_bar = n^3 / N^3.  For L = sum (delta - target)^2 the reverse
pass is: deltabar = 2 (delta - target) -> Psibar of the deposit (dj_vjp_deposit) -> back through the steps,
recomputing each force from its checkpoint (dj_vjp_run) -> back through the ICs (dj_vjp_ics_lpt).  The
reference has no adjoint of its PM path (NR/README.md:169); the pieces follow its CIC and FFT rrules
(NR/src/field/sheet_deposit.jl:118-132, NR/src/lpt/nlpt_core.jl:682-713).
"""
import numpy as np

from .cosmology import linear_power_spectrum, stepper_coefficients
from .device import DeviceArray
from .pm import PMConfig, PMSolverHIP, check_solver


def density_loss_gradient(cosmo, white, target, *, boxsize, res_pm, a_ini=0.05, a_end=1.0, n_steps=10,
                          stepper="bullfrog", time_var="D", order=2, worder=2, grad_kernel_order=4,
                          laplace_kernel_order=0, antialias=0, deconvolve=False, precision=32, k_tab=None,
                          pk_tab=None, solver=None, on_device=False):
    """-> (loss, white_bar, delta): L = sum (delta_final - target)^2 on the res_pm mesh and its gradient with
    respect to the (n,n,n) real white noise [k,j,i] of the ICs (FP64 array).  The momentum variable of the
    steps is Pi (bullfrog, fastpm); the symplectic stepper's velocity scaling is applied and reversed.
    on_device: white and target are (or are uploaded to) DeviceArrays, the loss and its cotangent are
    formed on the device (dj_density_loss_device) and white_bar is returned as a DeviceArray (delta as
    None): nothing of size n^3 crosses PCIe once the inputs are resident (bullfrog and fastpm steppers)."""
    if on_device:
        return _density_loss_gradient_device(cosmo, white, target, boxsize=boxsize, res_pm=res_pm, a_ini=a_ini,
                                             a_end=a_end, n_steps=n_steps, stepper=stepper, time_var=time_var,
                                             order=order, worder=worder, grad_kernel_order=grad_kernel_order,
                                             laplace_kernel_order=laplace_kernel_order, antialias=antialias,
                                             deconvolve=deconvolve, precision=precision, k_tab=k_tab, pk_tab=pk_tab,
                                             solver=solver)
    white = np.ascontiguousarray(white, np.float64)
    n = white.shape[0]
    if k_tab is None:
        k_tab, pk_tab = linear_power_spectrum(cosmo)
    cfg = PMConfig(res_pm=res_pm, n_part=n, boxsize=float(boxsize), worder=worder, deconvolve=deconvolve,
                   antialias=antialias, grad_order=grad_kernel_order, lap_order=laplace_kernel_order)
    co = stepper_coefficients(cosmo, stepper, a_ini, a_end, n_steps, time_var)
    D = cosmo.Dplus(a_ini)
    own = solver is None
    if not own:
        # the loss is built on the float deposit (the fixed-point one returns integer sums)
        check_solver(solver, cfg, precision, deposit="float")
    S = solver or PMSolverHIP(cfg, precision=precision)
    try:
        S.ics_lpt(k_tab, pk_tab, D, order=order, white=white)
        fp = cosmo.Fplus(a_ini) if stepper == "symplectic" else 1.0  # run_nbody.jl:104
        if fp != 1.0:
            psi0, pi0 = S.get_state()
            S.set_state(psi0, (pi0.astype(np.float64) * fp).astype(S.dtype))
        S.run_checkpointed(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"])
        rho = n**3 / res_pm**3
        delta = S.deposit().astype(np.float64) / rho - 1.0
        r = delta - np.asarray(target, np.float64)
        loss = float(np.sum(r * r))
        if fp == 1.0:
            # cotangents stay on the device from the deposit through the steps to the ICs
            S.vjp_deposit(None, (2.0 / rho) * r, resident=True)
            S.vjp_run()
            white_bar = S.vjp_ics_lpt(k_tab, pk_tab, D, order=order, white=white)
        else:
            psi_bar = S.vjp_deposit(None, (2.0 / rho) * r)
            psi_bar, pi_bar = S.vjp_run(psi_bar, np.zeros_like(psi_bar))
            pi_bar = (pi_bar.astype(np.float64) * fp).astype(S.dtype)
            white_bar = S.vjp_ics_lpt(k_tab, pk_tab, D, psi_bar, pi_bar, order=order, white=white)
    finally:
        if own:
            S.close()
    return loss, white_bar, delta


def _density_loss_gradient_device(cosmo, white, target, *, boxsize, res_pm, a_ini, a_end, n_steps, stepper,
                                  time_var, order, worder, grad_kernel_order, laplace_kernel_order, antialias,
                                  deconvolve, precision, k_tab, pk_tab, solver):
    if stepper == "symplectic":
        raise ValueError("on_device: the symplectic stepper's velocity scaling runs on the host; use on_device=False")
    dev = solver.device if solver is not None else 0
    w = white if isinstance(white, DeviceArray) else DeviceArray.from_numpy(np.asarray(white, np.float64), dev)
    tgt = None
    if target is not None:
        tgt = target if isinstance(target, DeviceArray) else DeviceArray.from_numpy(np.asarray(target, np.float64), dev)
    n = w.shape[0]
    if k_tab is None:
        k_tab, pk_tab = linear_power_spectrum(cosmo)
    cfg = PMConfig(res_pm=res_pm, n_part=n, boxsize=float(boxsize), worder=worder, deconvolve=deconvolve,
                   antialias=antialias, grad_order=grad_kernel_order, lap_order=laplace_kernel_order)
    co = stepper_coefficients(cosmo, stepper, a_ini, a_end, n_steps, time_var)
    D = cosmo.Dplus(a_ini)
    own = solver is None
    if not own:
        check_solver(solver, cfg, precision, deposit="float")
    S = solver or PMSolverHIP(cfg, precision=precision)
    try:
        S.ics_lpt_device(k_tab, pk_tab, D, w, order=order)
        S.run_checkpointed(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"])
        loss = S.density_loss_device(tgt)
        S.vjp_run()
        white_bar = S.vjp_ics_lpt_device(k_tab, pk_tab, D, w, DeviceArray(w.shape, np.float64, dev), order=order)
    finally:
        if own:
            S.close()
    return loss, white_bar, None
