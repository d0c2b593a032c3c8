/*
 * Oracle (test infrastructure only): FP64 host cosmology, growth timetables, jnp.interp lookups,
 * stepper coefficients and the Eisenstein-Hu linear power spectrum.
 * Restates NR/src/cosmology/Cosmology.jl, growth.jl, transfer.jl, NR/src/nbody/timetables.jl and
 * NR/src/nbody/run_nbody.jl:22-101.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

struct orc_cosmo {
    double Omega_c, Omega_b, h, sigma8, n_s, Omega_k, w0, wa;
    int n;
    double *tab[ORC_NTAB];
    double D1_at1;
};

double orc_Omega_m(const orc_cosmo *c) { return c->Omega_c + c->Omega_b; }

/* Cosmology.jl:48 */
static double omega_de(const orc_cosmo *c, double a) {
    return (1 - orc_Omega_m(c) - c->Omega_k) * pow(a, -3 * (1 + c->w0 + c->wa)) * exp(-3 * c->wa * (1 - a));
}

/* Cosmology.jl:55-58 */
double orc_hubble_E(const orc_cosmo *c, double a) {
    double Om = orc_Omega_m(c);
    return sqrt(Om * pow(a, -3.0) + c->Omega_k * pow(a, -2.0) + omega_de(c, a));
}

/* Cosmology.jl:67-71 */
static double dchi_da(const orc_cosmo *c, double a) { return 2997.92458 / (a * a * orc_hubble_E(c, a)); }

/* growth.jl:23-44: state y = [ln D1, f1, ln|D2|, f2, ln D3a, f3a, ln|D3b|, f3b, ln D3c] in t = ln a */
static void growth_derivs(const orc_cosmo *c, double lna, const double *y, double *dy) {
    double a = exp(lna);
    double E = orc_hubble_E(c, a);
    double Esq = E * E;
    double Om_a = (orc_Omega_m(c) * pow(a, -3.0)) / Esq;
    double Ox_a = omega_de(c, a) / Esq;
    double w_a = c->w0 + c->wa * (1 - a);
    double drag = 1 - 0.5 * (Om_a + (1 + 3 * w_a) * Ox_a);
    double y1 = y[0], f1 = y[1], y2 = y[2], f2 = y[3], y3a = y[4], f3a = y[5], y3b = y[6], f3b = y[7],
           y3c = y[8];
    double D1 = exp(y1), D1sq = D1 * D1, D1cu = D1sq * D1;
    double D2 = -exp(y2), D3a = exp(y3a), D3b = -exp(y3b), D3c = exp(y3c);
    dy[0] = f1;
    dy[1] = 1.5 * Om_a - f1 * f1 - f1 * drag;
    dy[2] = f2;
    dy[3] = 1.5 * Om_a * (1 - D1sq / D2) - f2 * f2 - f2 * drag;
    dy[4] = f3a;
    dy[5] = 1.5 * Om_a * (1 + 2 * D1cu / D3a) - f3a * f3a - f3a * drag;
    dy[6] = f3b;
    dy[7] = 1.5 * Om_a * (1 + (2 * D1 * D2 - 2 * D1cu) / D3b) - f3b * f3b - f3b * drag;
    dy[8] = D1 * D2 / D3c * (f1 - f2);
}

/* growth.jl:48-56 */
static void rk4_growth(const orc_cosmo *c, double lna, double *y, double h) {
    double k1[9], k2[9], k3[9], k4[9], t[9];
    growth_derivs(c, lna, y, k1);
    for (int i = 0; i < 9; i++) t[i] = y[i] + h / 2 * k1[i];
    growth_derivs(c, lna + h / 2, t, k2);
    for (int i = 0; i < 9; i++) t[i] = y[i] + h / 2 * k2[i];
    growth_derivs(c, lna + h / 2, t, k3);
    for (int i = 0; i < 9; i++) t[i] = y[i] + h * k3[i];
    growth_derivs(c, lna + h, t, k4);
    for (int i = 0; i < 9; i++) y[i] = y[i] + h / 6 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
}

/* Cosmology.jl:86-115 (compute_timetables), growth.jl:57-96 (growth tables, superconformal time) */
orc_cosmo *orc_cosmo_create(double Omega_c, double Omega_b, double h, double sigma8, double n_s,
                            double Omega_k, double w0, double wa, int n_pts, double a_ini) {
    orc_cosmo *c = (orc_cosmo *)calloc(1, sizeof(orc_cosmo));
    c->Omega_c = Omega_c; c->Omega_b = Omega_b; c->h = h; c->sigma8 = sigma8; c->n_s = n_s;
    c->Omega_k = Omega_k; c->w0 = w0; c->wa = wa; c->n = n_pts;
    for (int t = 0; t < ORC_NTAB; t++) c->tab[t] = (double *)calloc((size_t)n_pts, sizeof(double));
    double *a = c->tab[ORC_TAB_A];
    double lo = log10(a_ini);
    /* exp10.(LinRange(lo, 0, n)): LinRange element j = (1-t)*lo + t*0, t = j/(n-1) */
    for (int j = 0; j < n_pts; j++) {
        double t = (double)j / (double)(n_pts - 1);
        a[j] = exp10((1 - t) * lo + t * 0.0);
    }
    /* chi: RK4 over each [a_{i-1}, a_i] */
    double *chi = c->tab[ORC_TAB_CHI];
    double *fwd = (double *)malloc(sizeof(double) * (size_t)n_pts);
    fwd[0] = 0;
    for (int i = 1; i < n_pts; i++) {
        double ai = a[i - 1], da = a[i] - a[i - 1];
        double k1 = dchi_da(c, ai), k2 = dchi_da(c, ai + da / 2), k3 = dchi_da(c, ai + da / 2),
               k4 = dchi_da(c, ai + da);
        fwd[i] = fwd[i - 1] + da * (k1 + 2 * k2 + 2 * k3 + k4) / 6;
    }
    for (int i = 0; i < n_pts; i++) chi[i] = fwd[n_pts - 1] - fwd[i];
    free(fwd);
    /* growth: EdS initial conditions at a_min */
    double amin = a[0];
    double y[9] = {log(amin), 1.0, log(3.0 / 7.0 * (amin * amin)), 2.0,
                   log(1.0 / 3.0 * (amin * amin * amin)), 3.0, log(10.0 / 21.0 * (amin * amin * amin)), 3.0,
                   log(1.0 / 7.0 * (amin * amin * amin))};
    double *D1 = c->tab[ORC_TAB_D1], *D2 = c->tab[ORC_TAB_D2], *D3a = c->tab[ORC_TAB_D3A],
           *D3b = c->tab[ORC_TAB_D3B], *D3c = c->tab[ORC_TAB_D3C], *F1 = c->tab[ORC_TAB_F1],
           *F2 = c->tab[ORC_TAB_F2], *F3a = c->tab[ORC_TAB_F3A], *F3b = c->tab[ORC_TAB_F3B];
    for (int i = 0; i < n_pts; i++) {
        if (i > 0) rk4_growth(c, log(a[i - 1]), y, log(a[i]) - log(a[i - 1]));
        D1[i] = exp(y[0]); F1[i] = y[1]; F2[i] = y[3]; F3a[i] = y[5]; F3b[i] = y[7];
        D2[i] = -exp(y[2]); D3a[i] = exp(y[4]); D3b[i] = -exp(y[6]); D3c[i] = exp(y[8]);
    }
    /* normalise by D1(a = 1); a[end] = 1 exactly, so the linear interpolation returns D1[end] */
    double D1_at1 = orc_jinterp(1.0, a, D1, n_pts);
    c->D1_at1 = D1_at1;
    for (int i = 0; i < n_pts; i++) {
        D1[i] /= D1_at1;
        D2[i] /= D1_at1 * D1_at1;
        D3a[i] /= D1_at1 * D1_at1 * D1_at1;
        D3b[i] /= D1_at1 * D1_at1 * D1_at1;
        D3c[i] /= D1_at1 * D1_at1 * D1_at1;
    }
    /* superconformal time: reverse cumulative trapezoid of 1/(E a^3), zero at a = 1 */
    double *s = c->tab[ORC_TAB_SUPERCONFT];
    double *f = (double *)malloc(sizeof(double) * (size_t)n_pts);
    for (int i = 0; i < n_pts; i++) f[i] = 1.0 / (orc_hubble_E(c, a[i]) * (a[i] * a[i] * a[i]));
    s[n_pts - 1] = 0;
    for (int i = n_pts - 2; i >= 0; i--) s[i] = s[i + 1] - (a[i + 1] - a[i]) * (f[i] + f[i + 1]) / 2;
    double send = s[n_pts - 1];
    for (int i = 0; i < n_pts; i++) s[i] -= send;
    free(f);
    /* timetables.jl:17-18: dD+/da = f1 D1 / a, dD2/da = f2 D2 / a */
    for (int i = 0; i < n_pts; i++) {
        c->tab[ORC_TAB_DPLUSDA][i] = F1[i] * D1[i] / a[i];
        c->tab[ORC_TAB_D2PLUSDA][i] = F2[i] * D2[i] / a[i];
    }
    return c;
}

void orc_cosmo_destroy(orc_cosmo *c) {
    if (!c) return;
    for (int t = 0; t < ORC_NTAB; t++) free(c->tab[t]);
    free(c);
}

int orc_cosmo_npts(const orc_cosmo *c) { return c->n; }
void orc_cosmo_table(const orc_cosmo *c, int which, double *out) {
    memcpy(out, c->tab[which], sizeof(double) * (size_t)c->n);
}
double orc_cosmo_D1_unnormed_at_1(const orc_cosmo *c) { return c->D1_at1; }

/* timetables.jl:10-17: jnp.interp semantics (clamp outside, searchsortedlast inside) */
double orc_jinterp(double x, const double *xp, const double *fp, int n) {
    if (x <= xp[0]) return fp[0];
    if (x >= xp[n - 1]) return fp[n - 1];
    int lo = 0, hi = n - 1; /* invariant xp[lo] <= x < xp[hi] */
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (xp[mid] <= x) lo = mid; else hi = mid;
    }
    double t = (x - xp[lo]) / (xp[lo + 1] - xp[lo]);
    return fp[lo] + t * (fp[lo + 1] - fp[lo]);
}

/* timetables.jl:19-45 */
double orc_Dplus(const orc_cosmo *c, double a) { return orc_jinterp(a, c->tab[ORC_TAB_A], c->tab[ORC_TAB_D1], c->n); }
double orc_D2plus(const orc_cosmo *c, double a) { return orc_jinterp(a, c->tab[ORC_TAB_A], c->tab[ORC_TAB_D2], c->n); }
double orc_Dplusda(const orc_cosmo *c, double a) { return orc_jinterp(a, c->tab[ORC_TAB_A], c->tab[ORC_TAB_DPLUSDA], c->n); }
double orc_D2plusda(const orc_cosmo *c, double a) { return orc_jinterp(a, c->tab[ORC_TAB_A], c->tab[ORC_TAB_D2PLUSDA], c->n); }
double orc_superconft(const orc_cosmo *c, double a) { return orc_jinterp(a, c->tab[ORC_TAB_A], c->tab[ORC_TAB_SUPERCONFT], c->n); }
double orc_a_of_superconft(const orc_cosmo *c, double s) { return orc_jinterp(s, c->tab[ORC_TAB_SUPERCONFT], c->tab[ORC_TAB_A], c->n); }
double orc_a_of_Dplus(const orc_cosmo *c, double D) { return orc_jinterp(D, c->tab[ORC_TAB_D1], c->tab[ORC_TAB_A], c->n); }
double orc_Fplus(const orc_cosmo *c, double a) { return a * a * a * orc_hubble_E(c, a) * orc_Dplusda(c, a); }

/* run_nbody.jl:27-40 */
static double a_to_internal(const orc_cosmo *c, int tv, double a) {
    switch (tv) {
    case ORC_TV_A: return a;
    case ORC_TV_LOG_A: return log(a);
    case ORC_TV_D: return orc_Dplus(c, a);
    case ORC_TV_SUPERCONFT: return orc_superconft(c, a);
    }
    return NAN;
}
static double internal_to_a(const orc_cosmo *c, int tv, double x) {
    switch (tv) {
    case ORC_TV_A: return x;
    case ORC_TV_LOG_A: return exp(x);
    case ORC_TV_D: return orc_a_of_Dplus(c, x);
    case ORC_TV_SUPERCONFT: return orc_a_of_superconft(c, x);
    }
    return NAN;
}

/* run_nbody.jl:22: numpy.linspace */
static void nplinspace(double s, double e, int num, double *y) {
    double st = (e - s) / (num - 1);
    for (int i = 0; i < num; i++) y[i] = s + i * st;
    y[num - 1] = e;
}

/* 5-point Gauss-Legendre nodes and weights on [-1, 1] (closed forms) */
static void gl5(double *x, double *w) {
    double r = sqrt(10.0 / 7.0);
    double x1 = sqrt(5 - 2 * r) / 3, x2 = sqrt(5 + 2 * r) / 3;
    double w1 = (322 + 13 * sqrt(70.0)) / 900, w2 = (322 - 13 * sqrt(70.0)) / 900;
    x[0] = -x2; x[1] = -x1; x[2] = 0; x[3] = x1; x[4] = x2;
    w[0] = w2; w[1] = w1; w[2] = 128.0 / 225.0; w[3] = w1; w[4] = w2;
}

/* run_nbody.jl:48-101 */
int orc_stepper_coefficients(const orc_cosmo *c, int stepper, double a_ini, double a_end, int n_steps,
                             int time_var, const double *tv_array, double *alpha, double *beta,
                             double *dd1, double *dd2, double *all_a) {
    int n = n_steps;
    double *all_int = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *tgrid = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *a_mid = (double *)malloc(sizeof(double) * (size_t)n);
    if (time_var == ORC_TV_ARRAY) {
        nplinspace(0.0, (double)n, n + 1, tgrid);
        nplinspace(0.0, (double)n, n + 1, all_int);
        for (int i = 0; i <= n; i++) all_a[i] = orc_jinterp(all_int[i], tgrid, tv_array, n + 1);
        for (int i = 0; i < n; i++)
            a_mid[i] = orc_jinterp((all_int[i] + all_int[i + 1]) / 2, tgrid, tv_array, n + 1);
    } else {
        if (time_var < 0 || time_var > ORC_TV_SUPERCONFT) goto bad;
        double s = a_to_internal(c, time_var, a_ini), e = a_to_internal(c, time_var, a_end);
        nplinspace(s, e, n + 1, all_int);
        for (int i = 0; i <= n; i++) all_a[i] = internal_to_a(c, time_var, all_int[i]);
        for (int i = 0; i < n; i++) a_mid[i] = internal_to_a(c, time_var, (all_int[i] + all_int[i + 1]) / 2);
    }
    if (stepper == ORC_STEPPER_BULLFROG) {
        double D1 = c->D1_at1;
        for (int i = 0; i < n; i++) {
            double ab = all_a[i], ae = all_a[i + 1];
            double D_b = orc_Dplus(c, ab) * D1, D_e = orc_Dplus(c, ae) * D1;
            double Dda_b = orc_Dplusda(c, ab) * D1, Dda_e = orc_Dplusda(c, ae) * D1;
            double D2_b = orc_D2plus(c, ab) * (D1 * D1);
            double D2da_b = orc_D2plusda(c, ab) * (D1 * D1), D2da_e = orc_D2plusda(c, ae) * (D1 * D1);
            double dD = D_e - D_b;
            double xi = D_b / dD;
            double bracket = (D2_b + D2da_b / Dda_b * dD / 2) / ((xi + 0.5) * dD) - (xi + 0.5) * dD;
            alpha[i] = (D2da_e / Dda_e - bracket) / (D2da_b / Dda_b - bracket);
            double D_mid = orc_Dplus(c, a_mid[i]);
            beta[i] = (1 - alpha[i]) / D_mid;
            dd1[i] = D_mid - D_b / D1;
            dd2[i] = D_e / D1 - D_mid;
        }
    } else if (stepper == ORC_STEPPER_FASTPM) {
        for (int i = 0; i < n; i++) {
            alpha[i] = orc_Fplus(c, all_a[i]) / orc_Fplus(c, all_a[i + 1]);
            double D_mid = orc_Dplus(c, a_mid[i]);
            beta[i] = (1 - alpha[i]) / D_mid;
            dd1[i] = D_mid - orc_Dplus(c, all_a[i]);
            dd2[i] = orc_Dplus(c, all_a[i + 1]) - D_mid;
        }
    } else if (stepper == ORC_STEPPER_SYMPLECTIC) {
        double gx[5], gw[5];
        gl5(gx, gw);
        double Om = orc_Omega_m(c);
        for (int i = 0; i < n; i++) {
            double b = all_a[i], e = all_a[i + 1], sum = 0;
            for (int q = 0; q < 5; q++) {
                double aa = 0.5 * (e - b) * gx[q] + 0.5 * (b + e);
                sum += gw[q] * (1 / (aa * aa * orc_hubble_E(c, aa)));
            }
            alpha[i] = 1.0;
            beta[i] = 1.5 * Om * (0.5 * (e - b) * sum);
            double sc_b = orc_superconft(c, b), sc_e = orc_superconft(c, e), scm = orc_superconft(c, a_mid[i]);
            dd1[i] = scm - sc_b;
            dd2[i] = sc_e - scm;
        }
    } else {
        goto bad;
    }
    free(all_int); free(tgrid); free(a_mid);
    return 0;
bad:
    free(all_int); free(tgrid); free(a_mid);
    return -1;
}

/* transfer.jl:24-77: Eisenstein & Hu (1998) with baryons */
/* The scales of EH98 that T(k) below is built from, for the tests that pin it (sound horizon against
   the paper's approximation, k_eq against its definition): out = (k_eq [h/Mpc], z_eq, z_d, s [Mpc/h]) */
void orc_eh98_scales(const orc_cosmo *c, double Tcmb, double *out) {
    double th = Tcmb / 2.7, th2 = th * th, th4 = th2 * th2;
    double Om = orc_Omega_m(c);
    double omh2 = Om * c->h * c->h;
    double f_b = c->Omega_b / Om;
    double obh2 = omh2 * f_b;
    double z_eq = 2.50e4 * omh2 / th4;
    double k_eq = 0.0746 * omh2 / th2 / c->h;
    double z_d1 = 0.313 * pow(omh2, -0.419) * (1 + 0.607 * pow(omh2, 0.674));
    double z_d2 = 0.238 * pow(omh2, 0.223);
    double z_d = 1291.0 * pow(omh2, 0.251) / (1 + 0.659 * pow(omh2, 0.828)) * (1 + z_d1 * pow(obh2, z_d2));
    double R_d = 31.5 * obh2 / th4 * (1000 / (1 + z_d));
    double R_eq = 31.5 * obh2 / th4 * (1000 / z_eq);
    out[0] = k_eq;
    out[1] = z_eq;
    out[2] = z_d;
    out[3] = 2.0 / 3.0 / k_eq * sqrt(6 / R_eq) * log((sqrt(1 + R_d) + sqrt(R_d + R_eq)) / (1 + sqrt(R_eq)));
}

void orc_eisenstein_hu(const orc_cosmo *c, const double *k, int n, double Tcmb, double *T) {
    double th = Tcmb / 2.7, th2 = th * th, th4 = th2 * th2;
    double Om = orc_Omega_m(c);
    double omh2 = Om * c->h * c->h;
    double f_b = c->Omega_b / Om;
    double obh2 = omh2 * f_b;
    double z_eq = 2.50e4 * omh2 / th4;
    double k_eq = 0.0746 * omh2 / th2 / c->h;
    double z_d1 = 0.313 * pow(omh2, -0.419) * (1 + 0.607 * pow(omh2, 0.674));
    double z_d2 = 0.238 * pow(omh2, 0.223);
    double z_d = 1291.0 * pow(omh2, 0.251) / (1 + 0.659 * pow(omh2, 0.828)) * (1 + z_d1 * pow(obh2, z_d2));
    double R_d = 31.5 * obh2 / th4 * (1000 / (1 + z_d));
    double R_eq = 31.5 * obh2 / th4 * (1000 / z_eq);
    double s = 2.0 / 3.0 / k_eq * sqrt(6 / R_eq) * log((sqrt(1 + R_d) + sqrt(R_d + R_eq)) / (1 + sqrt(R_eq)));
    double k_silk = 1.6 * pow(obh2, 0.52) * pow(omh2, 0.73) * (1 + pow(10.4 * omh2, -0.95)) / c->h;
    double a1 = pow(46.9 * omh2, 0.670) * (1 + pow(32.1 * omh2, -0.532));
    double a2 = pow(12.0 * omh2, 0.424) * (1 + pow(45.0 * omh2, -0.582));
    double alpha_c = pow(a1, -f_b) * pow(a2, -(f_b * f_b * f_b));
    double b1 = 0.944 / (1 + pow(458 * omh2, -0.708));
    double b2 = pow(0.395 * omh2, -0.0266);
    double beta_c = 1 / (1 + b1 * (pow(1 - f_b, b2) - 1));
    double yG = (1 + z_eq) / (1 + z_d);
    double Gy = yG * (-6 * sqrt(1 + yG) + (2 + 3 * yG) * log((sqrt(1 + yG) + 1) / (sqrt(1 + yG) - 1)));
    double alpha_b = 2.07 * k_eq * s * pow(1 + R_d, -0.75) * Gy;
    double beta_b = 0.5 + f_b + (3 - 2 * f_b) * sqrt((17.2 * omh2) * (17.2 * omh2) + 1);
    double beta_node = 8.41 * pow(omh2, 0.435);
    for (int i = 0; i < n; i++) {
        double kk = k[i];
        double q = kk / (13.41 * k_eq);
        double ks = kk * s;
#define CF(ac) (14.2 / (ac) + 386 / (1 + 69.9 * pow(q, 1.08)))
#define LT(b) (log(exp(1.0) + 1.8 * (b) * q))
#define T0T(ac, bc) (LT(bc) / (LT(bc) + CF(ac) * q * q))
        double f = 1 / (1 + pow(ks / 5.4, 4));
        double T_c = f * T0T(1.0, beta_c) + (1 - f) * T0T(alpha_c, beta_c);
        double s_tilde = s / pow(1 + pow(beta_node / ks, 3), 1.0 / 3.0);
        double Tb1 = T0T(1.0, 1.0) / (1 + (ks / 5.2) * (ks / 5.2));
        double Tb2 = alpha_b / (1 + pow(beta_b / ks, 3)) * exp(-pow(kk / k_silk, 1.4));
        double xs = kk * s_tilde; /* sinc(x/pi) = sin(x)/x */
        double sinc = xs == 0 ? 1.0 : sin(xs) / xs;
        double T_b = sinc * (Tb1 + Tb2);
        T[i] = f_b * T_b + (1 - f_b) * T_c;
#undef CF
#undef LT
#undef T0T
    }
}

/* transfer.jl:121-153: P = k^ns T^2, sigma8-normalised by a trapezoid in ln k with R = 8 Mpc/h */
void orc_linear_power_spectrum(const orc_cosmo *c, double kmin, double kmax, int n_modes, double *k,
                               double *pk) {
    double lo = log10(kmin), hi = log10(kmax);
    for (int j = 0; j < n_modes; j++) {
        double t = (double)j / (double)(n_modes - 1);
        k[j] = exp10((1 - t) * lo + t * hi);
    }
    double *T = (double *)malloc(sizeof(double) * (size_t)n_modes);
    double *integ = (double *)malloc(sizeof(double) * (size_t)n_modes);
    orc_eisenstein_hu(c, k, n_modes, 2.72548, T);
    for (int j = 0; j < n_modes; j++) pk[j] = pow(k[j], c->n_s) * (T[j] * T[j]);
    for (int j = 0; j < n_modes; j++) {
        double x = k[j] * 8.0;
        double W = 3 * (sin(x) - x * cos(x)) / (x * x * x);
        integ[j] = k[j] * k[j] * k[j] * pk[j] * (W * W);
    }
    double s8 = 0;
    for (int j = 0; j + 1 < n_modes; j++) s8 += 0.5 * (integ[j] + integ[j + 1]) * (log(k[j + 1]) - log(k[j]));
    s8 /= 2 * M_PI * M_PI;
    double fac = c->sigma8 * c->sigma8 / s8;
    for (int j = 0; j < n_modes; j++) pk[j] *= fac;
    free(T);
    free(integ);
}
