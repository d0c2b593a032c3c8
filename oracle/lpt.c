/*
 * Oracle (test infrastructure only): FP64 initial conditions of the PM path.
 *   ic_operator / white_noise_to_fphi ..... NR/src/ics/grf.jl:118-144, 190-240
 *   compute_core at n_order = 2 ........... NR/src/lpt/nlpt_core.jl:46-78, 126-177, 284-297 and the
 *                                           fast engine NR/src/lpt/nlpt_fast.jl:27-64, 208-249
 *   nbody_ics_lpt (EdS) ................... NR/src/nbody/run_nbody.jl:180-208
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double _Complex cplx;

/* nlpt_core.jl:49-53: full axis fftshift(arange(-n/2, n/2)) * dk (index n/2 holds -n/2), half axis */
static void kfull(int n, double L, double *k) {
    double dk = 2 * M_PI / L;
    for (int m = 0; m < n; m++) k[m] = (double)(m < n / 2 ? m : m - n) * dk;
}
static void khalf(int n, double L, double *k) {
    double dk = 2 * M_PI / L;
    for (int m = 0; m <= n / 2; m++) k[m] = (double)m * dk;
}

/* nlpt_core.jl:57-67: i k along an axis, Nyquist zeroed; d[0..n) full x, y; dz[0..n/2] half */
static void grad_kernels(int n, double L, cplx *dx, cplx *dz) {
    double *kf = malloc(sizeof(double) * n), *kh = malloc(sizeof(double) * (n / 2 + 1));
    kfull(n, L, kf);
    khalf(n, L, kh);
    for (int m = 0; m < n; m++) dx[m] = I * kf[m];
    dx[n / 2] = 0;
    for (int m = 0; m <= n / 2; m++) dz[m] = I * kh[m];
    dz[n / 2] = 0;
    free(kf);
    free(kh);
}

/* nlpt_core.jl:129-138: pad (orig,orig,orig/2+1) -> (ext,ext,ext/2+1), Nyquist planes dropped */
static void pad3(const cplx *f, int orig, int ext, cplx *out) {
    int h = orig / 2, he = ext / 2 + 1;
    memset(out, 0, sizeof(cplx) * (size_t)ext * ext * he);
    double sc = (double)ext / orig;
    sc = sc * sc * sc;
    for (int z = 0; z < h; z++)
        for (int y = 0; y < orig; y++) {
            if (y == h) continue;
            int yd = y < h ? y : y - orig + ext;
            for (int x = 0; x < orig; x++) {
                if (x == h) continue;
                int xd = x < h ? x : x - orig + ext;
                out[xd + (size_t)ext * (yd + (size_t)ext * z)] = f[x + (size_t)orig * (y + (size_t)orig * z)] * sc;
            }
        }
}

/* nlpt_core.jl:142-151: crop back, Nyquist planes zero */
static void crop3(const cplx *f, int orig, int ext, cplx *out) {
    int h = orig / 2, ho = orig / 2 + 1;
    memset(out, 0, sizeof(cplx) * (size_t)orig * orig * ho);
    double sc = (double)orig / ext;
    sc = sc * sc * sc;
    for (int z = 0; z < h; z++)
        for (int y = 0; y < orig; y++) {
            if (y == h) continue;
            int ys = y < h ? y : y - orig + ext;
            for (int x = 0; x < orig; x++) {
                if (x == h) continue;
                int xs = x < h ? x : x - orig + ext;
                out[x + (size_t)orig * (y + (size_t)orig * z)] = f[xs + (size_t)ext * (ys + (size_t)ext * z)] * sc;
            }
        }
}

static void irfft3(int n, const cplx *f, double *out) {
    orc_brfft3_d(n, (const double *)f, out);
    double s = 1.0 / ((double)n * n * n);
    size_t M = (size_t)n * n * n;
    for (size_t i = 0; i < M; i++) out[i] *= s;
}

int orc_compute_core2_d(int n, double boxsize, int ext, const double *fphi_r, double *psi1, double *psi2) {
    if (ext <= 0) ext = 3 * n / 2;
    if (n % 2 || ext % 2 || ext < n) return -1;
    const cplx *fphi = (const cplx *)fphi_r;
    int nh = n / 2 + 1, eh = ext / 2 + 1;
    size_t H = (size_t)n * n * nh, HE = (size_t)ext * ext * eh, M = (size_t)n * n * n, ME = (size_t)ext * ext * ext;
    cplx *d = malloc(sizeof(cplx) * n), *dzh = malloc(sizeof(cplx) * nh);
    cplx *de = malloc(sizeof(cplx) * ext), *dze = malloc(sizeof(cplx) * eh);
    grad_kernels(n, boxsize, d, dzh);
    grad_kernels(ext, boxsize, de, dze);
    double *kf = malloc(sizeof(double) * n), *kh = malloc(sizeof(double) * nh);
    kfull(n, boxsize, kf);
    khalf(n, boxsize, kh);
    /* nlpt_core.jl:71-78: -1/k^2 (DC -> -1) */
    double *il = malloc(sizeof(double) * H);
    for (int z = 0; z < nh; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                size_t i = x + (size_t)n * (y + (size_t)n * z);
                double k2 = kf[x] * kf[x] + kf[y] * kf[y] + kh[z] * kh[z];
                il[i] = -1.0 / (i == 0 ? 1.0 : k2);
            }
    /* psi1 in Fourier space: -(d_c * fphi) (nlpt_core.jl:285-287) */
    cplx *f1 = malloc(sizeof(cplx) * H * 3);
    for (int c = 0; c < 3; c++)
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) {
                    size_t i = x + (size_t)n * (y + (size_t)n * z);
                    cplx g = c == 0 ? d[x] : c == 1 ? d[y] : dzh[z];
                    f1[i + c * H] = -(g * fphi[i]);
                }
    for (int c = 0; c < 3; c++) irfft3(n, f1 + c * H, psi1 + c * M);
    /* ext_derivs (nlpt_fast.jl:27-38): A[3i+k] = irfft_ext(dk_ext * pad(f_i)) */
    double *A = malloc(sizeof(double) * ME * 9);
    cplx *P = malloc(sizeof(cplx) * HE), *Q = malloc(sizeof(cplx) * HE);
    for (int i = 0; i < 3; i++) {
        pad3(f1 + i * H, n, ext, P);
        for (int k = 0; k < 3; k++) {
            for (int z = 0; z < eh; z++)
                for (int y = 0; y < ext; y++)
                    for (int x = 0; x < ext; x++) {
                        size_t e = x + (size_t)ext * (y + (size_t)ext * z);
                        cplx g = k == 0 ? de[x] : k == 1 ? de[y] : dze[z];
                        Q[e] = g * P[e];
                    }
            irfft3(ext, Q, A + (size_t)(3 * i + k) * ME);
        }
    }
    /* mu2 symmetric source (nlpt_fast.jl:57-61) times the order-2 factor ((3-2)/2 - 2)/((2+3/2)(2-1)) */
    double fac = ((3.0 - 2) / 2 - 2 * 1 * 1) / ((2 + 3.0 / 2) * (2 - 1));
    double *mu = malloc(sizeof(double) * ME);
    for (size_t p = 0; p < ME; p++) {
#define AA(q) A[(size_t)((q)-1) * ME + p]
        double m = AA(1) * AA(5) + AA(1) * AA(9) + AA(5) * AA(9) - AA(2) * AA(4) - AA(3) * AA(7) - AA(6) * AA(8);
#undef AA
        mu[p] = fac * m;
    }
    orc_rfft3_d(ext, mu, (double *)P);
    cplx *fL = malloc(sizeof(cplx) * H);
    crop3(P, n, ext, fL);
    /* assemble psi2 = il * (d_c * fL) (nlpt_core.jl:260-271), then irfft */
    for (int c = 0; c < 3; c++) {
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) {
                    size_t i = x + (size_t)n * (y + (size_t)n * z);
                    cplx g = c == 0 ? d[x] : c == 1 ? d[y] : dzh[z];
                    f1[i] = il[i] * (g * fL[i]);
                }
        irfft3(n, f1, psi2 + c * M);
    }
    free(d); free(dzh); free(de); free(dze); free(kf); free(kh); free(il); free(f1); free(A); free(P);
    free(Q); free(mu); free(fL);
    return 0;
}

/* run_nbody.jl:200-206: Psi = D psi1 + D^2 psi2, Pi = psi1 + 2 D psi2 */
void orc_nbody_ics_lpt2_d(const orc_cosmo *c, int64_t np3, double a, const double *psi1,
                          const double *psi2, double *psi, double *pi) {
    double D = orc_Dplus(c, a);
    for (int64_t i = 0; i < np3; i++) {
        psi[i] = (0.0 + D * psi1[i]) + (D * D) * psi2[i];
        pi[i] = (0.0 + 1.0 * psi1[i]) + (2 * D) * psi2[i];
    }
}

/* grf.jl:218-240: np.interp of the P(k) table (clamped) */
static double interp_pk(double k, const double *kt, const double *pt, int n) {
    if (k <= kt[0]) return pt[0];
    if (k >= kt[n - 1]) return pt[n - 1];
    int lo = 0, hi = n - 1; /* kt[lo] < k <= kt[hi] */
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (kt[mid] < k) lo = mid; else hi = mid;
    }
    double t = (k - kt[hi - 1]) / (kt[hi] - kt[hi - 1]);
    return pt[hi - 1] * (1 - t) + pt[hi] * t;
}

/* grf.jl:118-128 with the k grid of grf.jl:190-216 (index res/2 holds +res/2) */
void orc_ic_operator_d(int res, double boxsize, const double *k_tab, const double *pk_tab, int ntab,
                       double *scale) {
    double dk = 2 * M_PI / boxsize;
    double rb = (double)res / boxsize, nrm = rb * rb * rb;
    int nh = res / 2 + 1;
    for (int z = 0; z < nh; z++)
        for (int y = 0; y < res; y++)
            for (int x = 0; x < res; x++) {
                double kx = (x <= res / 2 ? x : x - res) * dk, ky = (y <= res / 2 ? y : y - res) * dk, kz = z * dk;
                double k2 = kx * kx + ky * ky + kz * kz;
                double P = interp_pk(sqrt(k2), k_tab, pk_tab, ntab);
                scale[x + (size_t)res * (y + (size_t)res * z)] = k2 == 0 ? 0.0 : -sqrt(P * nrm) / k2;
            }
}

/* grf.jl:139-140 */
void orc_white_noise_to_fphi_d(int res, const double *scale, const double *white, double *fphi) {
    orc_rfft3_d(res, white, fphi);
    size_t H = (size_t)res * res * (res / 2 + 1);
    for (size_t i = 0; i < H; i++) { fphi[2 * i] *= scale[i]; fphi[2 * i + 1] *= scale[i]; }
}
