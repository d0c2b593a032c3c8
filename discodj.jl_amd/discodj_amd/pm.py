"""PM force on the MI355X library, mirroring NR/src/nbody/pm.jl's interface.

    PMConfig(res_pm=, n_part=, boxsize=, worder=2, deconvolve=False, antialias=0, grad_order=4,
             lap_order=0, n_resample=1, resampling="fourier")               (pm.jl:35-46)
    S = PMSolverHIP(cfg, precision=32)                                        (PMSolver, pm.jl:385)
    acc = S.pm_acceleration(psi)                                              (pm_acceleration!, pm.jl:507)
    acc = pm_acceleration(psi, cfg)                                           (pm.jl:528)

Particle arrays are numpy (3, n, n, n) in C order, i.e. the reference's column-major (n,n,n,3)
[i,j,k,c] with c slowest; they are handed to the library without reordering.  Invalid options
raise DJError (the reference raises ErrorException from the same checks, pm.jl:136,146,570).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceArray
from ._lib import DJConfig, DJError, KernelStat, check


@dataclass
class PMConfig:
    res_pm: int
    n_part: int
    boxsize: float
    worder: int = 2
    deconvolve: bool = False
    antialias: int = 0
    grad_order: int = 4
    lap_order: int = 0
    n_resample: int = 1
    resampling: str = "fourier"


def _c_config(cfg, precision, deposit, lean, device, fixed_frac_bits):
    if cfg.resampling not in ("fourier", "linear"):
        raise DJError(-1, "resampling must be :fourier or :linear")
    return DJConfig(cfg.res_pm, cfg.n_part, float(cfg.boxsize), cfg.worder, int(bool(cfg.deconvolve)),
                    cfg.antialias, cfg.grad_order, cfg.lap_order, cfg.n_resample,
                    0 if cfg.resampling == "fourier" else 1, precision, _lib.DEPOSIT[deposit],
                    fixed_frac_bits, int(bool(lean)), device)


class PMSolverHIP:
    """Device-resident PM solver (a library context): owns the particle state (Psi, Pi), the mesh and
    spectrum buffers and the k-space multipliers.  precision 32 keeps mesh, FFT, forces and the
    particle state in FP32 (FP64 step coefficients); precision 64 runs the same kernels in FP64."""

    def __init__(self, cfg, precision=32, deposit="float", lean=False, device=0, fixed_frac_bits=0):
        self.cfg = cfg
        self.precision = precision
        self.dtype = np.float32 if precision == 32 else np.float64
        self.n = cfg.n_part
        self.shape = (3, self.n, self.n, self.n)
        self.deposit_mode = deposit
        self.lean = bool(lean)
        self.device = int(device)
        L = _lib.lib()
        ptr = C.c_void_p()
        c = _c_config(cfg, precision, deposit, lean, device, fixed_frac_bits)
        check(L.dj_create(C.byref(c), C.byref(ptr)))
        self._ctx = ptr
        self._L = L

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.dj_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            a = a.reshape(self.shape)
        return a

    def _chk(self, rc):
        return check(rc, self._ctx)

    # state ------------------------------------------------------------------------------------
    def set_state(self, psi, pi=None):
        psi = self._arr(psi)
        pi = None if pi is None else self._arr(pi)
        self._chk(self._L.dj_set_state(self._ctx, psi.ctypes.data, None if pi is None else pi.ctypes.data))

    def get_state(self):
        psi = np.empty(self.shape, self.dtype)
        pi = np.empty(self.shape, self.dtype)
        self._chk(self._L.dj_get_state(self._ctx, psi.ctypes.data, pi.ctypes.data))
        return psi, pi

    # force ------------------------------------------------------------------------------------
    def pm_acceleration(self, psi=None):
        """pm_acceleration!(S, psi) (pm.jl:507-521): -grad phi at q + psi, nabla^2 phi = delta."""
        acc = np.empty(self.shape, self.dtype)
        src = None if psi is None else self._arr(psi)
        self._chk(self._L.dj_acceleration(self._ctx, None if src is None else src.ctypes.data, acc.ctypes.data))
        return acc

    def deposit(self, psi=None, shift=None):
        """pm_scatter! of q + psi (+ shift): unnormalised unit-mass density (pm.jl:97-106); int64
        fixed-point values in the fixed64 deposit mode."""
        N = self.cfg.res_pm
        out = np.empty((N, N, N), np.int64 if self.deposit_mode == "fixed64" else self.dtype)
        src = None if psi is None else self._arr(psi)
        sh = None if shift is None else np.ascontiguousarray(shift, np.float64)
        self._chk(self._L.dj_deposit(self._ctx, None if src is None else src.ctypes.data,
                                     None if sh is None else sh.ctypes.data, out.ctypes.data))
        return out

    # stepping ---------------------------------------------------------------------------------
    def step(self, dd1, alpha, beta, dd2):
        self._chk(self._L.dj_step(self._ctx, float(dd1), float(alpha), float(beta), float(dd2)))

    def run(self, dd1, alpha, beta, dd2, want=None, callback=None):
        """n_steps DKD steps with FP64 coefficient arrays; callback(k) after the steps want[k] marks."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (dd1, alpha, beta, dd2)]
        n = len(arrs[0])
        w = None if want is None else np.ascontiguousarray(want, np.uint8)
        cb = _lib.STEP_CB(lambda k, u: callback(int(k))) if callback else _lib.STEP_CB()
        self._chk(self._L.dj_run(self._ctx, n, *(a.ctypes.data for a in arrs),
                                 None if w is None else w.ctypes.data, cb, None))

    def ics_lpt(self, k_tab, pk_tab, D, order=2, seed=42, white=None):
        """Initial conditions on the device (N-GenIC noise of `seed`, or the given real white noise
        (n,n,n) [k,j,i]), 1LPT/2LPT at growth factor D, into the solver's state (dj_ics_lpt)."""
        kt = np.ascontiguousarray(k_tab, np.float64)
        pt = np.ascontiguousarray(pk_tab, np.float64)
        w = None if white is None else np.ascontiguousarray(white, np.float64)
        self._chk(self._L.dj_ics_lpt(self._ctx, seed, None if w is None else w.ctypes.data, kt.ctypes.data,
                                     pt.ctypes.data, len(kt), float(D), order))

    # device-resident arrays (DeviceArray: memory of the HIP runtime the library links) ----------------
    def _dptr(self, t, nbytes, what):
        """the device address of t, a DeviceArray of nbytes on the solver's device (None passes through)"""
        if t is None:
            return None
        if not isinstance(t, DeviceArray):
            raise TypeError(f"{what}: a DeviceArray is expected")
        if t.nbytes != nbytes:
            raise ValueError(f"{what}: {nbytes} bytes expected, the array has {t.nbytes}")
        if t.device != self.device:
            raise ValueError(f"{what}: on device {t.device}, the solver runs on {self.device}")
        return t.ptr

    def density_loss_device(self, target=None):
        """L = sum (delta - target)^2 of the solver's Psi on the mesh (delta = mesh / rho - 1; target: an
        (N,N,N) FP64 DeviceArray or None for zero); its cotangent stays resident for vjp_run (dj_density_loss_device)"""
        N = self.cfg.res_pm
        loss = C.c_double()
        self._chk(self._L.dj_density_loss_device(self._ctx, self._dptr(target, 8 * N**3, "target"), C.byref(loss)))
        return loss.value

    def ics_lpt_device(self, k_tab, pk_tab, D, white, order=2, seed=42):
        """ics_lpt with the (n,n,n) FP64 white noise a DeviceArray (None: the N-GenIC field of seed)"""
        kt = np.ascontiguousarray(k_tab, np.float64)
        pt = np.ascontiguousarray(pk_tab, np.float64)
        n = self.cfg.n_part
        w = self._dptr(white, 8 * n**3, "white")
        self._chk(self._L.dj_ics_lpt_device(self._ctx, seed, w, kt.ctypes.data, pt.ctypes.data, len(kt), float(D),
                                            order))

    def deposit_device(self, out):
        """deposit() of the solver's Psi into the DeviceArray out (N^3 values of the solver's precision, int64
        in the fixed64 mode)"""
        N = self.cfg.res_pm
        es = 8 if self.deposit_mode == "fixed64" else np.dtype(self.dtype).itemsize
        self._chk(self._L.dj_deposit_device(self._ctx, self._dptr(out, es * N**3, "mesh")))
        return out

    def vjp_deposit_device(self, mesh_bar):
        """vjp_deposit(None, mesh_bar, resident=True) with mesh_bar (N^3, the solver's precision) a DeviceArray"""
        N = self.cfg.res_pm
        mb = self._dptr(mesh_bar, np.dtype(self.dtype).itemsize * N**3, "mesh_bar")
        self._chk(self._L.dj_vjp_deposit_device(self._ctx, mb))

    def vjp_ics_lpt_device(self, k_tab, pk_tab, D, white, white_bar, order=2, seed=42):
        """vjp_ics_lpt from the resident cotangents into the (n,n,n) FP64 DeviceArray white_bar"""
        kt = np.ascontiguousarray(k_tab, np.float64)
        pt = np.ascontiguousarray(pk_tab, np.float64)
        n = self.cfg.n_part
        w = self._dptr(white, 8 * n**3, "white")
        wb = self._dptr(white_bar, 8 * n**3, "white_bar")
        self._chk(self._L.dj_vjp_ics_lpt_device(self._ctx, seed, w, kt.ctypes.data, pt.ctypes.data, len(kt),
                                                float(D), order, wb))
        return white_bar

    def power_spectrum(self, psi=None, W=2, nbins=None):
        """(k, P, nmodes) of the CIC density of q + psi on the solver's mesh, window-deconvolved, linear
        bins of width kf up to k_Nyq/2 by default (dj_power_spectrum)."""
        nbins = nbins or self.cfg.res_pm // 4 - 1
        k, pk, nm = np.empty(nbins), np.empty(nbins), np.empty(nbins, np.int64)
        src = None if psi is None else self._arr(psi)
        self._chk(self._L.dj_power_spectrum(self._ctx, None if src is None else src.ctypes.data, W, nbins,
                                            k.ctypes.data, pk.ctypes.data, nm.ctypes.data))
        return k, pk, nm

    def field_power_spectrum(self, field, bins=20, logarithmic=True):
        """evaluate_power_spectrum(field, boxsize; bins, logarithmic) (power_spectrum.jl:25-79) of a real
        (N,N,N) field [iz,iy,ix] on the solver's mesh -> (k, Pk, N_modes)."""
        N = self.cfg.res_pm
        f = np.ascontiguousarray(field, self.dtype).reshape(N, N, N)
        k, pk, nm = np.empty(bins), np.empty(bins), np.empty(bins, np.int64)
        self._chk(self._L.dj_field_power_spectrum(self._ctx, f.ctypes.data, bins, int(bool(logarithmic)),
                                                  k.ctypes.data, pk.ctypes.data, nm.ctypes.data))
        return k, pk, nm

    def field_cross_power_spectrum(self, field1, field2, bins=20, logarithmic=True):
        """evaluate_cross_power_spectrum (power_spectrum.jl:86-123) -> (k, Pk12)."""
        N = self.cfg.res_pm
        f1 = np.ascontiguousarray(field1, self.dtype).reshape(N, N, N)
        f2 = np.ascontiguousarray(field2, self.dtype).reshape(N, N, N)
        k, pk = np.empty(bins), np.empty(bins)
        self._chk(self._L.dj_field_cross_power_spectrum(self._ctx, f1.ctypes.data, f2.ctypes.data, bins,
                                                        int(bool(logarithmic)), k.ctypes.data, pk.ctypes.data))
        return k, pk

    def power_spectra_binned(self, edges, psi=None, density=True, displacement=True):
        """pk_full of the study estimator (measure.jl:403-447) of q + psi (None: the solver's state) over the
        given bin edges -> (k centres, P_delta or None, P_Psi or None)."""
        e = np.ascontiguousarray(edges, np.float64)
        nb = len(e) - 1
        k = np.empty(nb)
        pd = np.empty(nb) if density else None
        pp = np.empty(nb) if displacement else None
        src = None if psi is None else self._arr(psi)
        self._chk(self._L.dj_power_spectra_binned(self._ctx, None if src is None else src.ctypes.data, nb,
                                                  e.ctypes.data, k.ctypes.data,
                                                  None if pd is None else pd.ctypes.data,
                                                  None if pp is None else pp.ctypes.data))
        return k, pd, pp

    def check_finite(self):
        self._chk(self._L.dj_check_finite(self._ctx))

    def set_kgain(self, gain, kmax=None):
        """Multiply the Poisson kernel by the transfer function g(|k|) given at len(gain) points
        |k| = i kmax / (len - 1) (linear in between, the last value beyond; dj_set_kgain); None removes it."""
        if gain is None:
            self._chk(self._L.dj_set_kgain(self._ctx, None, 0, 1.0))
            return
        g = np.ascontiguousarray(gain, np.float64)
        self._chk(self._L.dj_set_kgain(self._ctx, g.ctypes.data, len(g), float(kmax)))

    # learned correction -----------------------------------------------------------------------
    def nn_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == _lib.NN_PARAMS
        self._chk(self._L.dj_nn_set_params(self._ctx, p.ctypes.data))

    def nn_get_params(self):
        p = np.empty(_lib.NN_PARAMS, np.float32)
        self._chk(self._L.dj_nn_get_params(self._ctx, p.ctypes.data))
        return p

    def nn_enable(self, on=True, out_scale=1.0):
        self._chk(self._L.dj_nn_enable(self._ctx, int(on), float(out_scale)))

    def nn_set_schedule(self, a_force):
        a = np.ascontiguousarray(a_force, np.float64)
        self._chk(self._L.dj_nn_set_schedule(self._ctx, len(a), a.ctypes.data))

    def nn_forward(self, dens, a):
        """The potential correction (N,N,N) of a density mesh (the unnormalised deposit) at scale factor a."""
        N = self.cfg.res_pm
        d = np.ascontiguousarray(dens, np.float32).reshape(N, N, N)
        out = np.empty((N, N, N), np.float32)
        self._chk(self._L.dj_nn_forward(self._ctx, d.ctypes.data, float(a), out.ctypes.data))
        return out

    def nn_backward(self, dens, a, gphi):
        """Parameter gradient of <gphi, nn_forward(dens, a)>."""
        N = self.cfg.res_pm
        d = np.ascontiguousarray(dens, np.float32).reshape(N, N, N)
        g = np.ascontiguousarray(gphi, np.float32).reshape(N, N, N)
        out = np.empty(_lib.NN_PARAMS, np.float32)
        self._chk(self._L.dj_nn_backward(self._ctx, d.ctypes.data, float(a), g.ctypes.data, out.ctypes.data))
        return out

    def nn_train_step(self, psi, acc_target, a, lr, step, weight=1.0):
        """One Adam step on weight x the mean squared force error of the corrected PM at snapshot psi;
        -> the loss before the step."""
        loss = C.c_double()
        self._chk(self._L.dj_nn_train_step(self._ctx, self._arr(psi).ctypes.data, self._arr(acc_target).ctypes.data,
                                           float(a), float(weight), float(lr), int(step), C.byref(loss)))
        return loss.value

    # reverse mode -----------------------------------------------------------------------------
    def vjp_acceleration(self, psi, acc_bar):
        """(d acc / d psi)^T acc_bar of pm_acceleration at psi (None: the solver's state)."""
        out = np.empty(self.shape, self.dtype)
        src = None if psi is None else self._arr(psi)
        ab = self._arr(acc_bar)
        self._chk(self._L.dj_vjp_acceleration(self._ctx, None if src is None else src.ctypes.data, ab.ctypes.data,
                                              out.ctypes.data))
        return out

    def vjp_deposit(self, psi, mesh_bar, resident=False):
        """(d mesh / d psi)^T mesh_bar of the unshifted deposit at psi (None: the solver's state); resident:
        keep it on the device as the cotangents that vjp_run / vjp_ics_lpt take with None."""
        out = None if resident else np.empty(self.shape, self.dtype)
        src = None if psi is None else self._arr(psi)
        N = self.cfg.res_pm
        mb = np.ascontiguousarray(mesh_bar, self.dtype).reshape(N, N, N)
        self._chk(self._L.dj_vjp_deposit(self._ctx, None if src is None else src.ctypes.data, mb.ctypes.data,
                                         None if out is None else out.ctypes.data))
        return out

    def run_checkpointed(self, dd1, alpha, beta, dd2):
        """run() keeping the state before every step on the device, for vjp_run."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (dd1, alpha, beta, dd2)]
        self._chk(self._L.dj_run_checkpointed(self._ctx, len(arrs[0]), *(a.ctypes.data for a in arrs)))

    def vjp_run(self, psi_bar=None, pi_bar=None):
        """Cotangents of the state before the last run_checkpointed, given those of the state after it
        (both None: the resident cotangents, which stay on the device; returns None then)."""
        if psi_bar is None and pi_bar is None:
            self._chk(self._L.dj_vjp_run(self._ctx, None, None))
            return None
        pb, mb = self._arr(psi_bar).copy(), self._arr(pi_bar).copy()
        self._chk(self._L.dj_vjp_run(self._ctx, pb.ctypes.data, mb.ctypes.data))
        return pb, mb

    def vjp_ics_lpt(self, k_tab, pk_tab, D, psi_bar=None, pi_bar=None, order=2, seed=42, white=None):
        """Cotangent (n,n,n) FP64 of the real white noise of ics_lpt, given those of the ICs (Psi, Pi)
        (both None: the resident cotangents)."""
        kt = np.ascontiguousarray(k_tab, np.float64)
        pt = np.ascontiguousarray(pk_tab, np.float64)
        w = None if white is None else np.ascontiguousarray(white, np.float64)
        res = psi_bar is None and pi_bar is None
        pb, mb = (None, None) if res else (self._arr(psi_bar), self._arr(pi_bar))
        out = np.empty((self.n,) * 3, np.float64)
        self._chk(self._L.dj_vjp_ics_lpt(self._ctx, seed, None if w is None else w.ctypes.data, kt.ctypes.data,
                                         pt.ctypes.data, len(kt), float(D), order, None if res else pb.ctypes.data,
                                         None if res else mb.ctypes.data, out.ctypes.data))
        return out

    # measurement ------------------------------------------------------------------------------
    def synchronize(self):
        self._chk(self._L.dj_synchronize(self._ctx))

    def set_profiling(self, on=True):
        self._chk(self._L.dj_set_profiling(self._ctx, int(on)))

    def reset_stats(self):
        self._chk(self._L.dj_reset_stats(self._ctx))

    def kernel_stats(self):
        arr = (KernelStat * 32)()
        n = C.c_int32()
        self._chk(self._L.dj_kernel_stats(self._ctx, arr, 32, C.byref(n)))
        return {arr[i].name.decode(): {"calls": arr[i].calls, "ms_total": arr[i].ms_total,
                                       "bytes_per_call": arr[i].bytes_per_call} for i in range(n.value)}


def check_solver(S, cfg, precision, deposit=None, lean=None):
    """A solver handed to a driver in place of the one its arguments describe must compute the same
    model: the same PMConfig, precision and deposit mode (and lean form).  Raises ValueError otherwise."""
    if S.cfg != cfg:
        raise ValueError(f"the solver's configuration {S.cfg} differs from the requested {cfg}")
    if S.precision != precision:
        raise ValueError(f"the solver's precision {S.precision} differs from the requested {precision}")
    if deposit is not None and S.deposit_mode != deposit:
        raise ValueError(f"the solver's deposit mode {S.deposit_mode!r} differs from the requested {deposit!r}")
    if lean is not None and bool(S.lean) != bool(lean):
        raise ValueError(f"the solver's lean form ({S.lean}) differs from the requested one ({lean})")


def pm_acceleration(psi, cfg, precision=None, **kw):
    """pm_acceleration(psi, cfg) (pm.jl:528-529) on the library; precision follows psi's dtype."""
    psi = np.asarray(psi)
    if precision is None:
        precision = 64 if psi.dtype == np.float64 else 32
    with PMSolverHIP(cfg, precision=precision, **kw) as S:
        return S.pm_acceleration(psi)


def fft3d(x, direction=-1, device=0):
    """The library's 3-D transform (tests): real (N,N,N) [iz,iy,ix] -> half spectrum (N,N,N//2+1)
    [kz,ky,kx] (x halved), or the unnormalised inverse."""
    L = _lib.lib()
    x = np.ascontiguousarray(x)
    if direction < 0:
        prec = 64 if x.dtype == np.float64 else 32
        N = x.shape[0]
        out = np.empty((N, N, N // 2 + 1), np.complex128 if prec == 64 else np.complex64)
    else:
        prec = 64 if x.dtype == np.complex128 else 32
        N = x.shape[0]
        out = np.empty((N, N, N), np.float64 if prec == 64 else np.float32)
    check(L.dj_fft3d(N, prec, direction, x.ctypes.data, out.ctypes.data, device))
    return out


def ngenic_field_on_device(seed, res, device=0):
    """The N-GenIC Fourier white noise generated on the GPU: (res//2+1, res, res) complex [kz, j, i]."""
    out = np.empty((res // 2 + 1, res, res), np.complex128)
    check(_lib.lib().dj_ngenic_field_on_device(seed, res, device, out.ctypes.data))
    return out


def ngenic_field(seed, res):
    """N-GenIC Fourier white noise (host): (res//2+1, res, res) complex [kz, j, i] (halved along z)."""
    out = np.empty((res // 2 + 1, res, res), np.complex128)
    check(_lib.lib().dj_ngenic_field(seed, res, out.ctypes.data))
    return out
