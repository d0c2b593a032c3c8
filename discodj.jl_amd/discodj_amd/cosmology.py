"""Host-side FP64 cosmology for the PM stepper: timetables, jnp.interp lookups, stepper coefficients.

These stay on the host in FP64, as in the reference (the device only sees per-step coefficients):
    Cosmology / compute_timetables ........ NR/src/cosmology/Cosmology.jl:22-115, growth.jl:23-96
    lookups (_jinterp, Fplus, ...) ........ NR/src/nbody/timetables.jl:10-45
    stepper_coefficients .................. NR/src/nbody/run_nbody.jl:22-101
    Eisenstein-Hu, linear_power_spectrum .. NR/src/cosmology/transfer.jl:24-77, 121-153
"""
import math
from dataclasses import dataclass, field

import numpy as np

PREDEFINED = {
    "Planck18EEBAOSN": dict(Omega_c=0.259622, Omega_b=0.0488911, h=0.67742, sigma8=0.8105, n_s=0.96822),
    "Planck15": dict(Omega_c=0.2589, Omega_b=0.0486, h=0.6774, sigma8=0.8159, n_s=0.9667),
    "Quijote": dict(Omega_c=0.3175, Omega_b=0.049, h=0.6711, sigma8=0.834, n_s=0.9624),
    "CamelsCV": dict(Omega_c=0.2514, Omega_b=0.049, h=0.6711, sigma8=0.818, n_s=0.9624),
}


@dataclass
class Cosmology:
    Omega_c: float = 0.259622
    Omega_b: float = 0.0488911
    h: float = 0.67742
    sigma8: float = 0.8105
    n_s: float = 0.96822
    Omega_k: float = 0.0
    w0: float = -1.0
    wa: float = 0.0
    tables: dict = field(default_factory=dict)
    D1_unnormed_at_1: float = float("nan")

    @classmethod
    def named(cls, name):
        return compute_timetables(cls(**PREDEFINED[name]))

    @property
    def Omega_m(self):
        return self.Omega_c + self.Omega_b

    def omega_de(self, a):
        return (1 - self.Omega_m - self.Omega_k) * a ** (-3 * (1 + self.w0 + self.wa)) * math.exp(-3 * self.wa * (1 - a))

    def hubble_E(self, a):
        return math.sqrt(self.Omega_m * a**-3.0 + self.Omega_k * a**-2.0 + self.omega_de(a))

    # timetables.jl lookups ----------------------------------------------------------------------
    def _need(self):
        if not self.tables:
            raise ValueError("N-body timetables missing: build the cosmology with compute_timetables")

    def Dplus(self, a):
        self._need()
        return jinterp(a, self.tables["a"], self.tables["D1"])

    def D2plus(self, a):
        self._need()
        return jinterp(a, self.tables["a"], self.tables["D2"])

    def Dplusda(self, a):
        self._need()
        return jinterp(a, self.tables["a"], self.tables["Dplusda"])

    def D2plusda(self, a):
        self._need()
        return jinterp(a, self.tables["a"], self.tables["D2plusda"])

    def superconft(self, a):
        self._need()
        return jinterp(a, self.tables["a"], self.tables["superconft"])

    def a_of_superconft(self, s):
        self._need()
        return jinterp(s, self.tables["superconft"], self.tables["a"])

    def a_of_Dplus(self, D):
        self._need()
        return jinterp(D, self.tables["D1"], self.tables["a"])

    def Fplus(self, a):
        return a * a * a * self.hubble_E(a) * self.Dplusda(a)


def jinterp(x, xp, fp):
    """jnp.interp semantics: clamp outside, searchsortedlast inside (timetables.jl:10-17)."""
    if x <= xp[0]:
        return float(fp[0])
    if x >= xp[-1]:
        return float(fp[-1])
    i = int(np.searchsorted(xp, x, side="right")) - 1
    t = (x - xp[i]) / (xp[i + 1] - xp[i])
    return float(fp[i] + t * (fp[i + 1] - fp[i]))


def _growth_derivs(c, lna, y):
    a = math.exp(lna)
    E = c.hubble_E(a)
    Esq = E * E
    Om_a = (c.Omega_m * a**-3.0) / Esq
    Ox_a = c.omega_de(a) / Esq
    w_a = c.w0 + c.wa * (1 - a)
    drag = 1 - 0.5 * (Om_a + (1 + 3 * w_a) * Ox_a)
    y1, f1, y2, f2, y3a, f3a, y3b, f3b, y3c = y
    D1 = math.exp(y1)
    D1sq = D1 * D1
    D1cu = D1sq * D1
    D2, D3a, D3b, D3c = -math.exp(y2), math.exp(y3a), -math.exp(y3b), math.exp(y3c)
    return (f1, 1.5 * Om_a - f1 * f1 - f1 * drag,
            f2, 1.5 * Om_a * (1 - D1sq / D2) - f2 * f2 - f2 * drag,
            f3a, 1.5 * Om_a * (1 + 2 * D1cu / D3a) - f3a * f3a - f3a * drag,
            f3b, 1.5 * Om_a * (1 + (2 * D1 * D2 - 2 * D1cu) / D3b) - f3b * f3b - f3b * drag,
            D1 * D2 / D3c * (f1 - f2))


def compute_timetables(c, n_pts=2500, a_ini=1e-10):
    """Geometric a grid, RK4 growth ODE in ln a, reverse-trapezoid superconformal time."""
    lo = math.log10(a_ini)
    t = np.arange(n_pts) / (n_pts - 1)
    a = 10.0 ** ((1 - t) * lo + t * 0.0)
    a[-1] = 1.0
    amin = a[0]
    y = (math.log(amin), 1.0, math.log(3.0 / 7.0 * (amin * amin)), 2.0, math.log(1.0 / 3.0 * amin**3), 3.0,
         math.log(10.0 / 21.0 * amin**3), 3.0, math.log(1.0 / 7.0 * amin**3))
    rows = np.empty((n_pts, 9))
    rows[0] = y
    lna = np.log(a)
    for i in range(1, n_pts):
        h = lna[i] - lna[i - 1]
        x = lna[i - 1]
        k1 = _growth_derivs(c, x, y)
        k2 = _growth_derivs(c, x + h / 2, tuple(y[j] + h / 2 * k1[j] for j in range(9)))
        k3 = _growth_derivs(c, x + h / 2, tuple(y[j] + h / 2 * k2[j] for j in range(9)))
        k4 = _growth_derivs(c, x + h, tuple(y[j] + h * k3[j] for j in range(9)))
        y = tuple(y[j] + h / 6 * (k1[j] + 2 * k2[j] + 2 * k3[j] + k4[j]) for j in range(9))
        rows[i] = y
    D1 = np.exp(rows[:, 0])
    D1_at1 = float(D1[-1])
    tab = {"a": a, "D1": D1 / D1_at1, "f1": rows[:, 1], "D2": -np.exp(rows[:, 2]) / D1_at1**2,
           "f2": rows[:, 3], "D3a": np.exp(rows[:, 4]) / D1_at1**3, "f3a": rows[:, 5],
           "D3b": -np.exp(rows[:, 6]) / D1_at1**3, "f3b": rows[:, 7], "D3c": np.exp(rows[:, 8]) / D1_at1**3}
    f = np.array([1.0 / (c.hubble_E(ai) * ai**3) for ai in a])
    s = np.zeros(n_pts)
    for i in range(n_pts - 2, -1, -1):
        s[i] = s[i + 1] - (a[i + 1] - a[i]) * (f[i] + f[i + 1]) / 2
    tab["superconft"] = s - s[-1]
    tab["Dplusda"] = tab["f1"] * tab["D1"] / a
    tab["D2plusda"] = tab["f2"] * tab["D2"] / a
    c.tables = tab
    c.D1_unnormed_at_1 = D1_at1
    return c


def _nplinspace(s, e, num):
    st = (e - s) / (num - 1)
    y = [s + i * st for i in range(num)]
    y[-1] = e
    return y


_GL5 = None


def _gl5():
    global _GL5
    if _GL5 is None:
        r = math.sqrt(10.0 / 7.0)
        x1, x2 = math.sqrt(5 - 2 * r) / 3, math.sqrt(5 + 2 * r) / 3
        w1, w2 = (322 + 13 * math.sqrt(70.0)) / 900, (322 - 13 * math.sqrt(70.0)) / 900
        _GL5 = ((-x2, -x1, 0.0, x1, x2), (w2, w1, 128.0 / 225.0, w1, w2))
    return _GL5


def stepper_coefficients(c, stepper, a_ini, a_end, n_steps, time_var="D"):
    """Per-step (alpha, beta, ddrift1, ddrift2) and all_a of the DKD steppers (run_nbody.jl:48-101)."""
    if not isinstance(time_var, str):
        tv = np.asarray(time_var, np.float64)
        if len(tv) != n_steps + 1:
            raise ValueError("time_var vector must have n_steps + 1 entries")
        tgrid = np.array(_nplinspace(0.0, float(n_steps), n_steps + 1))
        itoa = lambda x: jinterp(x, tgrid, tv)  # noqa: E731
        all_int = _nplinspace(0.0, float(n_steps), n_steps + 1)
    else:
        to_int = {"a": lambda a: a, "log_a": math.log, "D": c.Dplus, "superconft": c.superconft}
        to_a = {"a": lambda x: x, "log_a": math.exp, "D": c.a_of_Dplus, "superconft": c.a_of_superconft}
        if time_var not in to_int:
            raise ValueError("time_var must be :a, :log_a, :D, :superconft or a vector of scale factors")
        itoa = to_a[time_var]
        all_int = _nplinspace(to_int[time_var](float(a_ini)), to_int[time_var](float(a_end)), n_steps + 1)
    all_a = [itoa(x) for x in all_int]
    a_mid = [itoa((all_int[i] + all_int[i + 1]) / 2) for i in range(n_steps)]
    alpha, beta, dd1, dd2 = (np.empty(n_steps) for _ in range(4))
    if stepper == "bullfrog":
        D1 = c.D1_unnormed_at_1
        for i in range(n_steps):
            ab, ae = all_a[i], all_a[i + 1]
            D_b, D_e = c.Dplus(ab) * D1, c.Dplus(ae) * D1
            Dda_b, Dda_e = c.Dplusda(ab) * D1, c.Dplusda(ae) * D1
            D2_b = c.D2plus(ab) * (D1 * D1)
            D2da_b, D2da_e = c.D2plusda(ab) * (D1 * D1), c.D2plusda(ae) * (D1 * D1)
            dD = D_e - D_b
            xi = D_b / dD
            bracket = (D2_b + D2da_b / Dda_b * dD / 2) / ((xi + 0.5) * dD) - (xi + 0.5) * dD
            alpha[i] = (D2da_e / Dda_e - bracket) / (D2da_b / Dda_b - bracket)
            D_mid = c.Dplus(a_mid[i])
            beta[i] = (1 - alpha[i]) / D_mid
            dd1[i] = D_mid - D_b / D1
            dd2[i] = D_e / D1 - D_mid
    elif stepper == "fastpm":
        for i in range(n_steps):
            alpha[i] = c.Fplus(all_a[i]) / c.Fplus(all_a[i + 1])
            D_mid = c.Dplus(a_mid[i])
            beta[i] = (1 - alpha[i]) / D_mid
            dd1[i] = D_mid - c.Dplus(all_a[i])
            dd2[i] = c.Dplus(all_a[i + 1]) - D_mid
    elif stepper == "symplectic":
        gx, gw = _gl5()
        for i in range(n_steps):
            b, e = all_a[i], all_a[i + 1]
            s = 0.0
            for q in range(5):
                aa = 0.5 * (e - b) * gx[q] + 0.5 * (b + e)
                s += gw[q] * (1 / (aa * aa * c.hubble_E(aa)))
            alpha[i] = 1.0
            beta[i] = 1.5 * c.Omega_m * (0.5 * (e - b) * s)
            scm = c.superconft(a_mid[i])
            dd1[i] = scm - c.superconft(b)
            dd2[i] = c.superconft(e) - scm
    else:
        raise ValueError("stepper must be :bullfrog, :fastpm or :symplectic")
    return {"alpha": alpha, "beta": beta, "ddrift1": dd1, "ddrift2": dd2, "all_a": np.array(all_a)}


def eisenstein_hu(c, k, Tcmb=2.72548):
    """EH98 transfer function with baryons (transfer.jl:24-77), k in h/Mpc."""
    k = np.asarray(k, np.float64)
    th = Tcmb / 2.7
    th2, th4 = th * th, th**4
    omh2 = c.Omega_m * c.h**2
    f_b = c.Omega_b / c.Omega_m
    obh2 = omh2 * f_b
    z_eq = 2.50e4 * omh2 / th4
    k_eq = 0.0746 * omh2 / th2 / c.h
    z_d1 = 0.313 * omh2**-0.419 * (1 + 0.607 * omh2**0.674)
    z_d2 = 0.238 * omh2**0.223
    z_d = 1291.0 * omh2**0.251 / (1 + 0.659 * omh2**0.828) * (1 + z_d1 * obh2**z_d2)
    R_d = 31.5 * obh2 / th4 * (1000 / (1 + z_d))
    R_eq = 31.5 * obh2 / th4 * (1000 / z_eq)
    s = 2 / 3 / k_eq * math.sqrt(6 / R_eq) * math.log((math.sqrt(1 + R_d) + math.sqrt(R_d + R_eq)) / (1 + math.sqrt(R_eq)))
    k_silk = 1.6 * obh2**0.52 * omh2**0.73 * (1 + (10.4 * omh2) ** -0.95) / c.h
    a1 = (46.9 * omh2) ** 0.670 * (1 + (32.1 * omh2) ** -0.532)
    a2 = (12.0 * omh2) ** 0.424 * (1 + (45.0 * omh2) ** -0.582)
    alpha_c = a1**-f_b * a2 ** -(f_b**3)
    b1 = 0.944 / (1 + (458 * omh2) ** -0.708)
    b2 = (0.395 * omh2) ** -0.0266
    beta_c = 1 / (1 + b1 * ((1 - f_b) ** b2 - 1))
    yG = (1 + z_eq) / (1 + z_d)
    Gy = yG * (-6 * math.sqrt(1 + yG) + (2 + 3 * yG) * math.log((math.sqrt(1 + yG) + 1) / (math.sqrt(1 + yG) - 1)))
    alpha_b = 2.07 * k_eq * s * (1 + R_d) ** -0.75 * Gy
    beta_b = 0.5 + f_b + (3 - 2 * f_b) * math.sqrt((17.2 * omh2) ** 2 + 1)
    beta_node = 8.41 * omh2**0.435
    q = k / (13.41 * k_eq)
    ks = k * s
    Cf = lambda ac: 14.2 / ac + 386 / (1 + 69.9 * q**1.08)  # noqa: E731
    lt = lambda b: np.log(math.e + 1.8 * b * q)  # noqa: E731
    T0t = lambda ac, bc: lt(bc) / (lt(bc) + Cf(ac) * q**2)  # noqa: E731
    f = 1 / (1 + (ks / 5.4) ** 4)
    T_c = f * T0t(1.0, beta_c) + (1 - f) * T0t(alpha_c, beta_c)
    s_tilde = s / (1 + (beta_node / ks) ** 3) ** (1 / 3)
    Tb1 = T0t(1.0, 1.0) / (1 + (ks / 5.2) ** 2)
    Tb2 = alpha_b / (1 + (beta_b / ks) ** 3) * np.exp(-((k / k_silk) ** 1.4))
    T_b = np.sinc(k * s_tilde / np.pi) * (Tb1 + Tb2)
    return f_b * T_b + (1 - f_b) * T_c


def linear_power_spectrum(c, kmin=1e-4, kmax=50.0, n_modes=512):
    """sigma8-normalised P(k) = k^ns T^2 on a log grid (transfer.jl:121-153) -> (k, Pk)."""
    t = np.arange(n_modes) / (n_modes - 1)
    k = 10.0 ** ((1 - t) * math.log10(kmin) + t * math.log10(kmax))
    pk = k**c.n_s * eisenstein_hu(c, k) ** 2
    x = k * 8.0
    W = 3 * (np.sin(x) - x * np.cos(x)) / x**3
    integ = k**3 * pk * W**2
    s8 = np.sum(0.5 * (integ[:-1] + integ[1:]) * np.diff(np.log(k))) / (2 * np.pi**2)
    return k, pk * (c.sigma8**2 / s8)
