"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

The oracle is the checker of the HIP library and the timed CPU baseline; product code never
imports this module.  Array conventions (numpy, C order):
    particle fields  (3, n, n, n)      [c, k, j, i]  == the reference's column-major (n,n,n,3) [i,j,k,c]
    meshes           (N, N, N)         [iz, iy, ix]  == column-major (N,N,N) [ix,iy,iz]
    half spectra     (N//2+1, N, N)    [kz, ky, kx]  complex128/complex64 (z halved, [3,1,2] region)
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

STEPPERS = {"bullfrog": 0, "fastpm": 1, "symplectic": 2}
TIME_VARS = {"a": 0, "log_a": 1, "D": 2, "superconft": 3}
TABLES = {"a": 0, "chi": 1, "D1": 2, "D2": 3, "f1": 4, "D3a": 5, "D3b": 6, "D3c": 7, "f2": 8,
          "f3a": 9, "f3b": 10, "superconft": 11, "Dplusda": 12, "D2plusda": 13}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


lib = _load()
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class PMConfigC(C.Structure):
    _fields_ = [("res_pm", C.c_int32), ("n_part", C.c_int32), ("boxsize", C.c_double),
                ("worder", C.c_int32), ("deconvolve", C.c_int32), ("antialias", C.c_int32),
                ("grad_order", C.c_int32), ("lap_order", C.c_int32), ("n_resample", C.c_int32),
                ("resampling", C.c_int32)]


def pm_config(res_pm, n_part, boxsize, worder=2, deconvolve=False, antialias=0, grad_order=4,
              lap_order=0, n_resample=1, resampling="fourier"):
    """PMConfig of NR/src/nbody/pm.jl:35-46 (defaults as there)."""
    return PMConfigC(res_pm, n_part, boxsize, worder, int(bool(deconvolve)), antialias, grad_order,
                     lap_order, n_resample, 0 if resampling == "fourier" else 1)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("orc_cosmo_create", C.c_void_p, [C.c_double] * 8 + [C.c_int, C.c_double])
_sig("orc_cosmo_destroy", None, [C.c_void_p])
_sig("orc_cosmo_npts", C.c_int, [C.c_void_p])
_sig("orc_cosmo_table", None, [C.c_void_p, C.c_int, _dp])
_sig("orc_cosmo_D1_unnormed_at_1", C.c_double, [C.c_void_p])
for _n in ("orc_Omega_m",):
    _sig(_n, C.c_double, [C.c_void_p])
for _n in ("orc_hubble_E", "orc_Dplus", "orc_D2plus", "orc_Dplusda", "orc_D2plusda", "orc_superconft",
           "orc_a_of_superconft", "orc_a_of_Dplus", "orc_Fplus"):
    _sig(_n, C.c_double, [C.c_void_p, C.c_double])
_sig("orc_stepper_coefficients", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_void_p, _dp, _dp, _dp, _dp, _dp])
_sig("orc_eisenstein_hu", None, [C.c_void_p, _dp, C.c_int, C.c_double, _dp])
_sig("orc_linear_power_spectrum", None, [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp])
_sig("orc_eh98_scales", None, [C.c_void_p, C.c_double, _dp])
_sig("orc_fft1_d", None, [C.c_int, _dp, _dp, C.c_int])
_sig("orc_fft1_f", None, [C.c_int, _fp, _fp, C.c_int])
_sig("orc_rfft3_d", None, [C.c_int, _dp, _dp])
_sig("orc_rfft3_f", None, [C.c_int, _fp, _fp])
_sig("orc_brfft3_d", None, [C.c_int, _dp, _dp])
_sig("orc_brfft3_f", None, [C.c_int, _fp, _fp])
for _s, _p in (("d", _dp), ("f", _fp)):
    for _n in ("orc_pm_acceleration_ref", "orc_pm_acceleration", "orc_pm_acceleration_lean"):
        _sig(f"{_n}_{_s}", C.c_int, [C.POINTER(PMConfigC), _p, _p])
    _sig(f"orc_pm_scatter_{_s}", None, [_p, _p, C.c_int64, C.c_int,
                                         C.c_double if _s == "d" else C.c_float, C.c_int])
    _sig(f"orc_pm_scatter_fixed_{_s}", None, [_i64p, _p, C.c_int64, C.c_int,
                                               C.c_double if _s == "d" else C.c_float, C.c_int, C.c_int])
    _sig(f"orc_pm_gather_{_s}", None, [_p, _p, _p, C.c_int64, C.c_int,
                                        C.c_double if _s == "d" else C.c_float, C.c_int])
    _sig(f"orc_pm_cells_{_s}", None, [_p, C.c_int64, C.c_int, C.c_double if _s == "d" else C.c_float,
                                       C.c_int, _i32p, _p])
    _sig(f"orc_run_nbody_{_s}", C.c_int, [C.c_void_p, C.POINTER(PMConfigC), C.c_int, C.c_double, C.c_double,
                                          C.c_int, C.c_int, C.c_void_p, C.c_int, _p, _p, _p, _p,
                                          C.c_void_p, C.c_void_p])
_sig("orc_bullfrog_ics_d", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, _dp, _dp, _dp])
_sig("orc_compute_core2_d", C.c_int, [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp])
_sig("orc_nbody_ics_lpt2_d", None, [C.c_void_p, C.c_int64, C.c_double, _dp, _dp, _dp, _dp])
_sig("orc_ic_operator_d", None, [C.c_int, C.c_double, _dp, _dp, C.c_int, _dp])
_sig("orc_white_noise_to_fphi_d", None, [C.c_int, _dp, _dp, _dp])
_sig("orc_ranlxd1_uniforms", None, [C.c_uint32, C.c_int, _dp])
_sig("orc_ngenic_field_gsl", None, [C.c_uint32, C.c_int, _dp])
_sig("orc_ngenic_wnoise_real", None, [C.c_uint32, C.c_int, _dp])
_sig("orc_power_spectrum_d", C.c_int, [C.c_int, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp, _i64p])
_sig("orc_power_spectrum_lin_d", C.c_int, [C.c_int, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp, _i64p])
_sig("orc_cross_power_spectrum_d", C.c_int, [C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp])
_sig("orc_pk_binned_d", None, [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_int, _dp, C.c_int,
                               C.c_void_p, C.c_void_p])

STEP_CB = C.CFUNCTYPE(None, C.c_int, C.c_double, C.c_void_p)


class Cosmology:
    """Cosmology + compute_timetables (NR/src/cosmology/Cosmology.jl:22-115)."""

    def __init__(self, Omega_c, Omega_b, h, sigma8, n_s, Omega_k=0.0, w0=-1.0, wa=0.0, n_pts=2500,
                 a_ini=1e-10):
        self.ptr = lib.orc_cosmo_create(Omega_c, Omega_b, h, sigma8, n_s, Omega_k, w0, wa, n_pts, a_ini)
        self.n = n_pts

    @classmethod
    def named(cls, name):
        p = {"Planck18EEBAOSN": (0.259622, 0.0488911, 0.67742, 0.8105, 0.96822),
             "Planck15": (0.2589, 0.0486, 0.6774, 0.8159, 0.9667),
             "Quijote": (0.3175, 0.049, 0.6711, 0.834, 0.9624),
             "CamelsCV": (0.2514, 0.049, 0.6711, 0.818, 0.9624)}[name]
        return cls(*p)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.orc_cosmo_destroy(self.ptr)
            self.ptr = None

    def table(self, name):
        out = np.empty(self.n)
        lib.orc_cosmo_table(self.ptr, TABLES[name], out)
        return out

    @property
    def D1_unnormed_at_1(self):
        return lib.orc_cosmo_D1_unnormed_at_1(self.ptr)

    def __getattr__(self, name):
        f = getattr(lib, "orc_" + name, None)
        if f is None or name.startswith("_"):
            raise AttributeError(name)
        return lambda *a: f(self.ptr, *a)


def stepper_coefficients(cosmo, stepper, a_ini, a_end, n_steps, time_var="D"):
    """NR/src/nbody/run_nbody.jl:48-101 -> dict(alpha, beta, ddrift1, ddrift2, all_a)."""
    out = {k: np.empty(n_steps) for k in ("alpha", "beta", "ddrift1", "ddrift2")}
    out["all_a"] = np.empty(n_steps + 1)
    if isinstance(time_var, str):
        tv, arr = TIME_VARS[time_var], None
    else:
        tv, arr = 4, np.ascontiguousarray(time_var, dtype=np.float64)
    rc = lib.orc_stepper_coefficients(cosmo.ptr, STEPPERS[stepper], a_ini, a_end, n_steps, tv,
                                      None if arr is None else arr.ctypes.data, out["alpha"], out["beta"],
                                      out["ddrift1"], out["ddrift2"], out["all_a"])
    if rc:
        raise ValueError("invalid stepper / time_var")
    return out


def fft1(x, sign=-1):
    x = np.ascontiguousarray(x)
    if x.dtype == np.complex64:
        out = np.empty_like(x)
        lib.orc_fft1_f(len(x), x.view(np.float32), out.view(np.float32), sign)
    else:
        x = x.astype(np.complex128)
        out = np.empty_like(x)
        lib.orc_fft1_d(len(x), x.view(np.float64), out.view(np.float64), sign)
    return out


def rfft3(field):
    N = field.shape[0]
    if field.dtype == np.float32:
        out = np.empty((N // 2 + 1, N, N), np.complex64)
        lib.orc_rfft3_f(N, np.ascontiguousarray(field), out.view(np.float32))
    else:
        out = np.empty((N // 2 + 1, N, N), np.complex128)
        lib.orc_rfft3_d(N, np.ascontiguousarray(field, dtype=np.float64), out.view(np.float64))
    return out


def brfft3(spec):
    N = spec.shape[1]
    if spec.dtype == np.complex64:
        out = np.empty((N, N, N), np.float32)
        lib.orc_brfft3_f(N, np.ascontiguousarray(spec).view(np.float32), out)
    else:
        out = np.empty((N, N, N), np.float64)
        lib.orc_brfft3_d(N, np.ascontiguousarray(spec, dtype=np.complex128).view(np.float64), out)
    return out


def _sfx(a):
    return "f" if a.dtype == np.float32 else "d"


def pm_acceleration(psi, cfg, form="solver"):
    """acc (3,n,n,n) of psi (3,n,n,n): form 'reference' (pm.jl:253), 'solver' (pm.jl:507), 'lean' (pm.jl:623)."""
    psi = np.ascontiguousarray(psi)
    acc = np.empty_like(psi)
    name = {"reference": "orc_pm_acceleration_ref", "solver": "orc_pm_acceleration",
            "lean": "orc_pm_acceleration_lean"}[form]
    rc = getattr(lib, f"{name}_{_sfx(psi)}")(C.byref(cfg), psi, acc)
    if rc:
        raise ValueError("invalid PM configuration")
    return acc


def pm_scatter(X, res, L, W, mesh=None):
    X = np.ascontiguousarray(X)
    np_ = X.size // 3
    if mesh is None:
        mesh = np.zeros((res, res, res), X.dtype)
    getattr(lib, f"orc_pm_scatter_{_sfx(X)}")(mesh, X, np_, res, L, W)
    return mesh


def pm_scatter_fixed(X, res, L, W, frac_bits=32, mesh=None):
    X = np.ascontiguousarray(X)
    if mesh is None:
        mesh = np.zeros((res, res, res), np.int64)
    getattr(lib, f"orc_pm_scatter_fixed_{_sfx(X)}")(mesh, X, X.size // 3, res, L, W, frac_bits)
    return mesh


def pm_gather(mesh, X, L, W):
    X = np.ascontiguousarray(X)
    out = np.empty(X.size // 3, X.dtype)
    getattr(lib, f"orc_pm_gather_{_sfx(X)}")(out, np.ascontiguousarray(mesh, X.dtype), X, X.size // 3,
                                              mesh.shape[0], L, W)
    return out


def pm_cells(X, res, L, W):
    X = np.ascontiguousarray(X)
    np_ = X.size // 3
    base = np.empty((3, np_), np.int32)
    w = np.empty((3, np_, W), X.dtype)
    getattr(lib, f"orc_pm_cells_{_sfx(X)}")(X, np_, res, L, W, base, w)
    return base, w


def run_nbody(cosmo, psi_ini, pi_ini, cfg, a_ini, a_end, n_steps, stepper="bullfrog", time_var="D",
              lean=False, callback=None):
    """run_nbody(...; return_displacement=true) -> (Psi, P)  (NR/src/nbody/run_nbody.jl:124-172)."""
    psi_ini = np.ascontiguousarray(psi_ini)
    pi_ini = np.ascontiguousarray(pi_ini, dtype=psi_ini.dtype)
    psi = np.empty_like(psi_ini)
    p = np.empty_like(psi_ini)
    if isinstance(time_var, str):
        tv, arr = TIME_VARS[time_var], None
    else:
        tv, arr = 4, np.ascontiguousarray(time_var, dtype=np.float64)
    cb = STEP_CB(lambda k, a, u: callback(k, a)) if callback else None
    rc = getattr(lib, f"orc_run_nbody_{_sfx(psi_ini)}")(
        cosmo.ptr, C.byref(cfg), STEPPERS[stepper], a_ini, a_end, n_steps, tv,
        None if arr is None else arr.ctypes.data, int(lean), psi_ini, pi_ini, psi, p,
        C.cast(cb, C.c_void_p) if cb else None, None)
    if rc:
        raise ValueError("invalid run_nbody configuration")
    return psi, p


def bullfrog_ics(cosmo, psi1, a_ini, boxsize, res_pm):
    psi1 = np.ascontiguousarray(psi1, dtype=np.float64)
    n = psi1.shape[1]
    psi, pi = np.empty_like(psi1), np.empty_like(psi1)
    if lib.orc_bullfrog_ics_d(cosmo.ptr, n, boxsize, res_pm, a_ini, psi1, psi, pi):
        raise ValueError("bullfrog_ics failed")
    return psi, pi


def compute_core2(fphi, boxsize, ext=0):
    """fphi (n//2+1, n, n) complex -> psi1, psi2 (3,n,n,n) (nlpt_core.jl:311, n_order = 2)."""
    n = fphi.shape[1]
    psi1 = np.empty((3, n, n, n))
    psi2 = np.empty((3, n, n, n))
    if lib.orc_compute_core2_d(n, boxsize, ext, np.ascontiguousarray(fphi, np.complex128).view(np.float64),
                               psi1, psi2):
        raise ValueError("compute_core failed")
    return psi1, psi2


def nbody_ics_lpt2(cosmo, psi1, psi2, a):
    psi, pi = np.empty_like(psi1), np.empty_like(psi1)
    lib.orc_nbody_ics_lpt2_d(cosmo.ptr, psi1.size, a, psi1, psi2, psi, pi)
    return psi, pi


def linear_power_spectrum(cosmo, kmin=1e-4, kmax=50.0, n_modes=512):
    k = np.empty(n_modes)
    pk = np.empty(n_modes)
    lib.orc_linear_power_spectrum(cosmo.ptr, kmin, kmax, n_modes, k, pk)
    return k, pk


def eisenstein_hu(cosmo, k, Tcmb=2.72548):
    k = np.ascontiguousarray(k, np.float64)
    T = np.empty_like(k)
    lib.orc_eisenstein_hu(cosmo.ptr, k, len(k), Tcmb, T)
    return T


def eh98_scales(cosmo, Tcmb=2.72548):
    """(k_eq [h/Mpc], z_eq, z_d, sound horizon s [Mpc/h]) of the EH98 fit."""
    out = np.empty(4)
    lib.orc_eh98_scales(cosmo.ptr, Tcmb, out)
    return tuple(out)


def ic_operator(res, boxsize, k_tab, pk_tab):
    scale = np.empty((res // 2 + 1, res, res))
    lib.orc_ic_operator_d(res, boxsize, np.ascontiguousarray(k_tab, np.float64),
                          np.ascontiguousarray(pk_tab, np.float64), len(k_tab), scale)
    return scale


def white_noise_to_fphi(scale, white):
    res = white.shape[0]
    f = np.empty((res // 2 + 1, res, res), np.complex128)
    lib.orc_white_noise_to_fphi_d(res, scale, np.ascontiguousarray(white, np.float64), f.view(np.float64))
    return f


def ranlxd1_uniforms(seed, n):
    out = np.empty(n)
    lib.orc_ranlxd1_uniforms(seed, n, out)
    return out


def ngenic_field_gsl(seed, res):
    out = np.empty((res // 2 + 1, res, res), np.complex128)
    lib.orc_ngenic_field_gsl(seed, res, out.view(np.float64))
    return out


def ngenic_wnoise_real(seed, res):
    out = np.empty((res, res, res))
    lib.orc_ngenic_wnoise_real(seed, res, out)
    return out


def power_spectrum(field, boxsize, bins=20, logarithmic=True):
    res = field.shape[0]
    k, pk, nm = np.empty(bins), np.empty(bins), np.empty(bins, np.int64)
    lib.orc_power_spectrum_d(res, np.ascontiguousarray(field, np.float64), boxsize, bins, int(logarithmic),
                             k, pk, nm)
    return k, pk, nm


def power_spectrum_lin(field, boxsize, W=2, nbins=None):
    res = field.shape[0]
    nbins = nbins or res // 4 - 1
    k, pk, nm = np.empty(nbins), np.empty(nbins), np.empty(nbins, np.int64)
    lib.orc_power_spectrum_lin_d(res, np.ascontiguousarray(field, np.float64), boxsize, W, nbins, k, pk, nm)
    return k, pk, nm


def cross_power_spectrum(f1, f2, boxsize, bins=20, logarithmic=True):
    """evaluate_cross_power_spectrum (power_spectrum.jl:86-123) -> (k, Pk12)."""
    res = f1.shape[0]
    k, pk = np.empty(bins), np.empty(bins)
    lib.orc_cross_power_spectrum_d(res, np.ascontiguousarray(f1, np.float64), np.ascontiguousarray(f2, np.float64),
                                   boxsize, bins, int(logarithmic), k, pk)
    return k, pk


def pk_binned(boxsize, edges, delta=None, psi=None, deconv_W=2):
    """pk_full of the study estimator (measure.jl:403-447) over the given edges -> (P_delta, P_psi); delta is a
    density contrast (N,N,N) on the deposit mesh, psi the displacement (3,n,n,n); either may be None."""
    edges = np.ascontiguousarray(edges, np.float64)
    nb = len(edges) - 1
    d = None if delta is None else np.ascontiguousarray(delta, np.float64)
    p = None if psi is None else np.ascontiguousarray(psi, np.float64)
    pd, pp = np.zeros(nb), np.zeros(nb)
    lib.orc_pk_binned_d(0 if d is None else d.shape[0], None if d is None else d.ctypes.data,
                        0 if p is None else p.shape[1], None if p is None else p.ctypes.data, boxsize, nb, edges,
                        deconv_W, pd.ctypes.data, pp.ctypes.data)
    return (None if d is None else pd), (None if p is None else pp)


def lagrangian_positions(psi, boxsize):
    """X = q + psi with q = (i-1) L/n computed in FP64 and cast (pm.jl:186, 512)."""
    n = psi.shape[1]
    q = (np.arange(n) * (boxsize / n)).astype(psi.dtype)
    X = np.empty_like(psi)
    X[0] = psi[0] + q[None, None, :]
    X[1] = psi[1] + q[None, :, None]
    X[2] = psi[2] + q[:, None, None]
    return X
