/*
 * Oracle (test infrastructure only): power spectra of a mesh.
 *   evaluate_power_spectrum ........ NR/src/analysis/power_spectrum.jl:25-79
 *   linear-bin, window-deconvolved P(k) that measures the PM path's accuracy, following the
 *   study estimator studies/eft_shell_crossing/src/measure.jl:403-447 (mode weights 1 on the kz = 0
 *   and kz = N/2 planes and 2 elsewhere, |W(k)|^2 = prod sinc(k dx / 2)^(2W) divided out).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int orc_power_spectrum_d(int res, const double *field, double boxsize, int bins, int logarithmic,
                         double *k_out, double *pk_out, int64_t *nmodes_out) {
    int nh = res / 2 + 1;
    size_t H = (size_t)res * res * nh;
    double *F = malloc(sizeof(double) * 2 * H);
    orc_rfft3_d(res, field, F);
    double V = boxsize * boxsize * boxsize, dk = 2 * M_PI / boxsize, kny = M_PI * res / boxsize;
    double r6 = (double)res * res * res;
    r6 = r6 * r6;
    double kmin = dk, kmax = kny * sqrt(3.0);
    double *edges = malloc(sizeof(double) * (bins + 1));
    for (int b = 0; b <= bins; b++) {
        double t = (double)b / bins;
        edges[b] = logarithmic ? exp10((1 - t) * log10(kmin) + t * log10(kmax)) : (1 - t) * kmin + t * kmax;
    }
    double *sk = calloc(bins, sizeof(double)), *sp = calloc(bins, sizeof(double)), *sw = calloc(bins, sizeof(double));
    int64_t *cnt = calloc(bins, sizeof(int64_t));
    for (int z = 0; z < nh; z++)
        for (int y = 0; y < res; y++)
            for (int x = 0; x < res; x++) {
                size_t i = x + (size_t)res * (y + (size_t)res * z);
                double kx = (x <= res / 2 ? x : x - res) * dk, ky = (y <= res / 2 ? y : y - res) * dk, kz = z * dk;
                double k = sqrt(kx * kx + ky * ky + kz * kz);
                double p = (F[2 * i] * F[2 * i] + F[2 * i + 1] * F[2 * i + 1]) * (V / r6);
                double w = (z == 0 || z == nh - 1) ? 0.5 : 1.0;
                for (int b = 0; b < bins; b++)
                    if (k >= edges[b] && k < edges[b + 1]) {
                        sk[b] += k; sp[b] += p * w; sw[b] += w; cnt[b]++;
                        break;
                    }
            }
    for (int b = 0; b < bins; b++) {
        k_out[b] = cnt[b] ? sk[b] / cnt[b] : sqrt(edges[b] * edges[b + 1]);
        pk_out[b] = cnt[b] ? sp[b] / sw[b] : 0.0;
        nmodes_out[b] = cnt[b];
    }
    free(F); free(edges); free(sk); free(sp); free(sw); free(cnt);
    return bins;
}

int orc_power_spectrum_lin_d(int res, const double *field, double boxsize, int W, int nbins,
                             double *k_out, double *pk_out, int64_t *nmodes_out) {
    int nh = res / 2 + 1;
    size_t H = (size_t)res * res * nh;
    double *F = malloc(sizeof(double) * 2 * H);
    orc_rfft3_d(res, field, F);
    double kf = 2 * M_PI / boxsize, dx = boxsize / res;
    double *sp = calloc(nbins, sizeof(double)), *sw = calloc(nbins, sizeof(double));
    int64_t *cnt = calloc(nbins, sizeof(int64_t));
    for (int z = 0; z < nh; z++)
        for (int y = 0; y < res; y++)
            for (int x = 0; x < res; x++) {
                size_t i = x + (size_t)res * (y + (size_t)res * z);
                double kv[3] = {(x < res / 2 ? x : x - res) * kf, (y < res / 2 ? y : y - res) * kf, z * kf};
                double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                int b = (int)floor(k / kf - 0.5); /* bin b covers [kf (b+0.5), kf (b+1.5)) */
                if (k == 0 || b < 0 || b >= nbins) continue;
                double win = 1.0;
                if (W > 0)
                    for (int a = 0; a < 3; a++) {
                        double u = kv[a] * dx / 2, s = u == 0 ? 1.0 : sin(u) / u;
                        win *= pow(s, 2 * W);
                    }
                double w = (z == 0 || z == nh - 1) ? 1.0 : 2.0;
                sp[b] += w * (F[2 * i] * F[2 * i] + F[2 * i + 1] * F[2 * i + 1]) / win;
                sw[b] += w;
                cnt[b]++;
            }
    double n6 = (double)res * res * res;
    n6 = n6 * n6;
    double V = boxsize * boxsize * boxsize;
    for (int b = 0; b < nbins; b++) {
        k_out[b] = kf * (b + 1);
        pk_out[b] = sw[b] > 0 ? V * sp[b] / sw[b] / n6 : 0.0;
        nmodes_out[b] = cnt[b];
    }
    free(F); free(sp); free(sw); free(cnt);
    return nbins;
}

/* evaluate_cross_power_spectrum (NR/src/analysis/power_spectrum.jl:86-123): Re(conj(f1_k) f2_k) V / N^6
   averaged over the modes of each bin without weights (the reference's sum(cross[mask]) / n_k); bins,
   k grid and bin centres as evaluate_power_spectrum's. */
int orc_cross_power_spectrum_d(int res, const double *f1, const double *f2, double boxsize, int bins,
                               int logarithmic, double *k_out, double *pk_out) {
    int nh = res / 2 + 1;
    size_t H = (size_t)res * res * nh;
    double *F1 = malloc(sizeof(double) * 2 * H), *F2 = malloc(sizeof(double) * 2 * H);
    orc_rfft3_d(res, f1, F1);
    orc_rfft3_d(res, f2, F2);
    double V = boxsize * boxsize * boxsize, dk = 2 * M_PI / boxsize, kny = M_PI * res / boxsize;
    double r6 = (double)res * res * res;
    r6 = r6 * r6;
    double kmin = dk, kmax = kny * sqrt(3.0);
    double *edges = malloc(sizeof(double) * (bins + 1));
    for (int b = 0; b <= bins; b++) {
        double t = (double)b / bins;
        edges[b] = logarithmic ? exp10((1 - t) * log10(kmin) + t * log10(kmax)) : (1 - t) * kmin + t * kmax;
    }
    double *sk = calloc(bins, sizeof(double)), *sp = calloc(bins, sizeof(double));
    int64_t *cnt = calloc(bins, sizeof(int64_t));
    for (int z = 0; z < nh; z++)
        for (int y = 0; y < res; y++)
            for (int x = 0; x < res; x++) {
                size_t i = x + (size_t)res * (y + (size_t)res * z);
                double kx = (x <= res / 2 ? x : x - res) * dk, ky = (y <= res / 2 ? y : y - res) * dk, kz = z * dk;
                double k = sqrt(kx * kx + ky * ky + kz * kz);
                double c = (F1[2 * i] * F2[2 * i] + F1[2 * i + 1] * F2[2 * i + 1]) * (V / r6);
                for (int b = 0; b < bins; b++)
                    if (k >= edges[b] && k < edges[b + 1]) {
                        sk[b] += k; sp[b] += c; cnt[b]++;
                        break;
                    }
            }
    for (int b = 0; b < bins; b++) {
        k_out[b] = cnt[b] ? sk[b] / cnt[b] : sqrt(edges[b] * edges[b + 1]);
        pk_out[b] = cnt[b] ? sp[b] / cnt[b] : 0.0;
    }
    free(F1); free(F2); free(edges); free(sk); free(sp); free(cnt);
    return bins;
}

/* The study estimator pk_full (studies/eft_shell_crossing/src/measure.jl:403-447) over given bin edges
   (nbins + 1, ascending; a mode with |k| > 0 is in bin b when edges[b] <= |k| < edges[b + 1], as
   searchsortedlast in _pk_bins): mode weights 1 on the kz = 0 and kz = N/2 planes, 2 elsewhere.
     delta (res^3, the density contrast on the deposit mesh): P_delta = L^3 sum w |d_k|^2 / W2 / sum w / res^6,
       W2 = prod_a sinc(k_a dx / 2)^4 (the CIC window squared), or no division when deconv_W = 0
     psi (3, n, n, n) the displacement on the particle grid: P_psi = L^3 sum_c sum w |psi_c,k|^2 / (sum w / 3) / n^6
   Either part is skipped when its pointer is NULL. */
void orc_pk_binned_d(int res, const double *delta, int n, const double *psi, double boxsize, int nbins,
                     const double *edges, int deconv_W, double *pk_delta, double *pk_psi) {
    double L3 = boxsize * boxsize * boxsize;
    for (int part = 0; part < 2; part++) {
        const double *src = part == 0 ? delta : psi;
        double *out = part == 0 ? pk_delta : pk_psi;
        if (!src || !out) continue;
        int N = part == 0 ? res : n, ncomp = part == 0 ? 1 : 3, nh = N / 2 + 1;
        size_t H = (size_t)N * N * nh, M = (size_t)N * N * N;
        double dk = 2 * M_PI / boxsize, dx = boxsize / N;
        double *F = malloc(sizeof(double) * 2 * H), *acc = calloc(nbins, sizeof(double)), *cw = calloc(nbins, sizeof(double));
        for (int c = 0; c < ncomp; c++) {
            orc_rfft3_d(N, src + c * M, F);
            for (int z = 0; z < nh; z++)
                for (int y = 0; y < N; y++)
                    for (int x = 0; x < N; x++) {
                        size_t i = x + (size_t)N * (y + (size_t)N * z);
                        double kv[3] = {(x <= N / 2 ? x : x - N) * dk, (y <= N / 2 ? y : y - N) * dk, z * dk};
                        double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                        if (!(k > 0) || k < edges[0] || !(k < edges[nbins])) continue;
                        int b = 0;
                        while (!(k < edges[b + 1])) b++;
                        double v = F[2 * i] * F[2 * i] + F[2 * i + 1] * F[2 * i + 1];
                        if (part == 0 && deconv_W > 0) {
                            double win = 1.0;
                            for (int a = 0; a < 3; a++) {
                                double u = kv[a] * dx / 2, s = u == 0 ? 1.0 : sin(u) / u;
                                win *= pow(s, 2 * deconv_W);
                            }
                            v /= win;
                        }
                        double w = (z == 0 || z == nh - 1) ? 1.0 : 2.0;
                        acc[b] += w * v;
                        cw[b] += w;
                    }
        }
        double N6 = (double)N * N * N;
        N6 = N6 * N6;
        for (int b = 0; b < nbins; b++) out[b] = cw[b] > 0 ? L3 * acc[b] / (cw[b] / ncomp) / N6 : 0.0;
        free(F); free(acc); free(cw);
    }
}
