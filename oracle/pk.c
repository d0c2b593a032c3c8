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
