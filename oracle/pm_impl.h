/*
 * Oracle (test infrastructure only): particle-mesh force of the reference, instantiated for
 * REAL = double / float by pm_d.c / pm_f.c.  Restates NR/src/nbody/pm.jl:
 *   weights and stencils ............ pm.jl:49-55
 *   scatter / gather ................ pm.jl:57-95
 *   k vectors and Fourier kernels ... pm.jl:119-183
 *   sheet resampling (_spawn) ....... pm.jl:186-234
 *   interlacing shifts .............. pm.jl:237-246
 *   reference force ................. pm.jl:253-298 (pm_acceleration_reference, _acc_pm_single)
 *   PMSolver force .................. pm.jl:385-521 (combined K_phi, FD stencil, gather3)
 *   PMSolverLean force .............. pm.jl:569-641
 * Arithmetic follows the reference's operation order in REAL precision; the oracle is built with
 * -ffp-contract=off so no multiply-add is fused.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef REAL _Complex CPLX;

/* Julia mod(x, y) for floats: rem, then shift into the sign of y */
static inline REAL SFX(jmod)(REAL x, REAL L) {
    REAL r = (REAL)fmod(x, L);
    if (r == 0) return copysign(r, L);
    if ((r > 0) != (L > 0)) return r + L;
    return r;
}
static inline int SFX(imod)(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

/* pm.jl:49-51 */
static inline REAL SFX(wk)(int W, REAL d) {
    if (W == 2) return d > 1 ? (REAL)0 : 1 - d;
    if (W == 3) {
        if (d > (REAL)1.5) return 0;
        if (d > (REAL)0.5) { REAL t = (REAL)1.5 - d; return (REAL)0.5 * (t * t); }
        return (REAL)0.75 - d * d;
    }
    if (d > 2) return 0;
    if (d > 1) { REAL t = 2 - d; return t * t * t / 6; }
    return (4 - 6 * (d * d) + 3 * (d * d * d)) / 6;
}
/* pm.jl:52-55: base cell floor (even W) or ties-to-even round (odd W); stencil offsets */
static inline int SFX(wbase)(int W, REAL g) { return (W % 2 == 0) ? (int)floor(g) : (int)rint(g); }
static inline int wlo(int W) { return W == 2 ? 0 : -1; }

typedef struct {
    int b[3];
    REAL w[3][4];
    int cell[3][4];
} SFX(stencil);

static inline void SFX(make_stencil)(SFX(stencil) * st, REAL x, REAL y, REAL z, int res, REAL L, int W) {
    REAL g[3] = {SFX(jmod)(x, L) / L * (REAL)res, SFX(jmod)(y, L) / L * (REAL)res,
                 SFX(jmod)(z, L) / L * (REAL)res};
    int lo = wlo(W);
    for (int a = 0; a < 3; a++) {
        st->b[a] = SFX(wbase)(W, g[a]);
        for (int o = 0; o < W; o++) {
            int bo = st->b[a] + lo + o;
            st->w[a][o] = SFX(wk)(W, (REAL)fabs(g[a] - (REAL)bo));
            st->cell[a][o] = SFX(imod)(bo, res);
        }
    }
}

void SFX(orc_pm_cells)(const REAL *X, int64_t np, int res, REAL L, int W, int32_t *base, REAL *w) {
    for (int64_t p = 0; p < np; p++) {
        SFX(stencil) st;
        SFX(make_stencil)(&st, X[p], X[p + np], X[p + 2 * np], res, L, W);
        for (int a = 0; a < 3; a++) {
            base[p + a * np] = st.b[a];
            for (int o = 0; o < W; o++) w[(p + a * np) * W + o] = st.w[a][o];
        }
    }
}

/* deposit every particle's stencil cells whose z index lies in [z0, z1), in particle order */
static void SFX(scatter_slab)(REAL *mesh, const REAL *X, int64_t np, int res, REAL L, int W, int z0, int z1) {
    size_t R = (size_t)res;
    int lo = wlo(W);
    for (int64_t p = 0; p < np; p++) {
        REAL gz = SFX(jmod)(X[p + 2 * np], L) / L * (REAL)res;
        int bz = SFX(wbase)(W, gz), hit = 0;
        for (int o = 0; o < W; o++) {
            int kz = SFX(imod)(bz + lo + o, res);
            hit |= (kz >= z0 && kz < z1);
        }
        if (!hit) continue;
        SFX(stencil) st;
        SFX(make_stencil)(&st, X[p], X[p + np], X[p + 2 * np], res, L, W);
        for (int oz = 0; oz < W; oz++) {
            if (st.cell[2][oz] < z0 || st.cell[2][oz] >= z1) continue;
            for (int oy = 0; oy < W; oy++)
                for (int ox = 0; ox < W; ox++)
                    mesh[st.cell[0][ox] + R * (st.cell[1][oy] + R * st.cell[2][oz])] +=
                        st.w[0][ox] * st.w[1][oy] * st.w[2][oz];
        }
    }
}

/* pm.jl:57-74, 102-106.  Each thread owns a z-slab of the mesh and adds the contributions of all
   particles to it in particle order, so every cell sums in the sequential order: the result is
   bitwise identical for any thread count. */
void SFX(orc_pm_scatter)(REAL *mesh, const REAL *X, int64_t np, int res, REAL L, int W) {
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    if (nt > res) nt = res;
    if (nt <= 1 || np < 4096) {
        SFX(scatter_slab)(mesh, X, np, res, L, W, 0, res);
        return;
    }
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        SFX(scatter_slab)(mesh, X, np, res, L, W, res * t / nt, res * (t + 1) / nt);
    }
}

void SFX(orc_pm_scatter_fixed)(int64_t *mesh, const REAL *X, int64_t np, int res, REAL L, int W,
                               int frac_bits) {
    size_t R = (size_t)res;
    double scale = ldexp(1.0, frac_bits);
    for (int64_t p = 0; p < np; p++) {
        SFX(stencil) st;
        SFX(make_stencil)(&st, X[p], X[p + np], X[p + 2 * np], res, L, W);
        for (int oz = 0; oz < W; oz++)
            for (int oy = 0; oy < W; oy++)
                for (int ox = 0; ox < W; ox++) {
                    REAL w = st.w[0][ox] * st.w[1][oy] * st.w[2][oz];
                    mesh[st.cell[0][ox] + R * (st.cell[1][oy] + R * st.cell[2][oz])] +=
                        (int64_t)llrint((double)w * scale);
                }
    }
}

/* pm.jl:76-95 */
void SFX(orc_pm_gather)(REAL *out, const REAL *mesh, const REAL *X, int64_t np, int res, REAL L, int W) {
    size_t R = (size_t)res;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; p++) {
        SFX(stencil) st;
        SFX(make_stencil)(&st, X[p], X[p + np], X[p + 2 * np], res, L, W);
        REAL s = 0;
        for (int oz = 0; oz < W; oz++)
            for (int oy = 0; oy < W; oy++)
                for (int ox = 0; ox < W; ox++)
                    s += mesh[st.cell[0][ox] + R * (st.cell[1][oy] + R * st.cell[2][oz])] *
                         (st.w[0][ox] * st.w[1][oy] * st.w[2][oz]);
        out[p] = s;
    }
}

/* ------------------------------------------------------------------ Fourier kernels (pm.jl:119-183) */
typedef struct {
    int res, nh;
    CPLX *g[3];    /* gradient kernels: g[0] (res) along x, g[1] (res) along y, g[2] (nh) along z */
    REAL *invlap;  /* (res,res,nh) */
    REAL *mak;     /* (res,res,nh) or NULL */
} SFX(kernels);

static void SFX(kvecs)(int res, double L, REAL *kf, REAL *kh) {
    for (int m = 0; m < res; m++) kf[m] = (REAL)((double)(m < res / 2 ? m : m - res) * 2 * M_PI / L);
    for (int m = 0; m <= res / 2; m++) kh[m] = (REAL)((double)m * 2 * M_PI / L);
}

static REAL SFX(maxabs)(const REAL *k, int n) {
    REAL m = 0;
    for (int i = 0; i < n; i++) if ((REAL)fabs(k[i]) > m) m = (REAL)fabs(k[i]);
    return m;
}

/* pm.jl:125-138 */
static int SFX(grad1d)(const REAL *k, int n, int order, int half, CPLX *g) {
    if (order == 0) {
        for (int i = 0; i < n; i++) g[i] = I * k[i];
        if (half) g[n - 1] = 0; else g[n / 2] = 0;
        return 0;
    }
    if (order != 2 && order != 4 && order != 6) return -1;
    REAL m = SFX(maxabs)(k, n) / (REAL)M_PI;
    for (int i = 0; i < n; i++) {
        REAL s1 = (REAL)sin(1 * k[i] / m), s2 = (REAL)sin(2 * k[i] / m), s3 = (REAL)sin(3 * k[i] / m);
        REAL kern = order == 2 ? s1 : order == 4 ? (8 * s1 - s2) / 6 : (45 * s1 - 9 * s2 + s3) / 30;
        g[i] = I * (m * kern);
    }
    return 0;
}

/* pm.jl:140-147 */
static int SFX(invlap_terms)(const REAL *k, int n, int order, REAL *t) {
    REAL m = SFX(maxabs)(k, n) / (REAL)M_PI;
    for (int i = 0; i < n; i++) {
        REAL c1 = (REAL)cos(1 * k[i] / m), c2 = (REAL)cos(2 * k[i] / m), c3 = (REAL)cos(3 * k[i] / m);
        if (order == 2) t[i] = -2 * (1 - c1) * (m * m);
        else if (order == 4) t[i] = -(c2 - 16 * c1 + 15) / 6 * (m * m);
        else if (order == 6) t[i] = -(-2 * c3 + 27 * c2 - 270 * c1 + 245) / 90 * (m * m);
        else return -1;
    }
    return 0;
}

/* pm.jl:150-151 */
static void SFX(invmak1d)(const REAL *k, int n, int W, REAL *out) {
    REAL m = SFX(maxabs)(k, n) / (REAL)M_PI;
    for (int i = 0; i < n; i++) {
        double x = (double)k[i] / (2 * M_PI) / (double)m;
        REAL v = x == 0 ? (REAL)1 : (REAL)(sin(M_PI * x) / (M_PI * x));
        out[i] = (REAL)pow(v, -2 * W);
    }
}

static int SFX(kernels_init)(SFX(kernels) * K, const orc_pm_config *cfg) {
    int res = cfg->res_pm, nh = res / 2 + 1;
    K->res = res; K->nh = nh;
    REAL *kf = (REAL *)malloc(sizeof(REAL) * res), *kh = (REAL *)malloc(sizeof(REAL) * nh);
    SFX(kvecs)(res, cfg->boxsize, kf, kh);
    K->g[0] = (CPLX *)malloc(sizeof(CPLX) * res);
    K->g[1] = (CPLX *)malloc(sizeof(CPLX) * res);
    K->g[2] = (CPLX *)malloc(sizeof(CPLX) * nh);
    int err = SFX(grad1d)(kf, res, cfg->grad_order, 0, K->g[0]);
    err |= SFX(grad1d)(kf, res, cfg->grad_order, 0, K->g[1]);
    err |= SFX(grad1d)(kh, nh, cfg->grad_order, 1, K->g[2]);
    size_t M = (size_t)res * res * nh;
    K->invlap = (REAL *)malloc(sizeof(REAL) * M);
    if (cfg->lap_order == 0) {
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < res; y++)
                for (int x = 0; x < res; x++) {
                    REAL k2 = kf[x] * kf[x] + kf[y] * kf[y] + kh[z] * kh[z];
                    size_t i = x + (size_t)res * (y + (size_t)res * z);
                    K->invlap[i] = (i == 0) ? (REAL)0 : -1 / k2;
                }
    } else {
        REAL *tf = (REAL *)malloc(sizeof(REAL) * res), *th = (REAL *)malloc(sizeof(REAL) * nh);
        err |= SFX(invlap_terms)(kf, res, cfg->lap_order, tf);
        err |= SFX(invlap_terms)(kh, nh, cfg->lap_order, th);
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < res; y++)
                for (int x = 0; x < res; x++) {
                    size_t i = x + (size_t)res * (y + (size_t)res * z);
                    REAL inv = (i == 0) ? (REAL)-1 : tf[x] + tf[y] + th[z];
                    K->invlap[i] = 1 / inv;
                }
        free(tf); free(th);
    }
    K->mak = NULL;
    if (cfg->deconvolve) {
        REAL *mf = (REAL *)malloc(sizeof(REAL) * res), *mh = (REAL *)malloc(sizeof(REAL) * nh);
        SFX(invmak1d)(kf, res, cfg->worder, mf);
        SFX(invmak1d)(kh, nh, cfg->worder, mh);
        K->mak = (REAL *)malloc(sizeof(REAL) * M);
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < res; y++)
                for (int x = 0; x < res; x++) K->mak[x + (size_t)res * (y + (size_t)res * z)] = mf[x] * mf[y] * mh[z];
        free(mf); free(mh);
    }
    free(kf); free(kh);
    return err;
}

static void SFX(kernels_free)(SFX(kernels) * K) {
    free(K->g[0]); free(K->g[1]); free(K->g[2]); free(K->invlap); free(K->mak);
}

/* ------------------------------------------------------------------ shifts */
/* pm.jl:237-246: interlacing shift list (order of itertools.product) */
static int SFX(aa_shifts)(int aa, double dh, double (*s)[3]) {
    if (aa == 0) { s[0][0] = s[0][1] = s[0][2] = 0; return 1; }
    if (aa == -1) { s[0][0] = s[0][1] = s[0][2] = dh; return 1; }
    if (aa == 1) { s[0][0] = s[0][1] = s[0][2] = 0; s[1][0] = s[1][1] = s[1][2] = dh; return 2; }
    if (aa != 2 && aa != 3) return -1;
    int n = 0;
    double v[2] = {0.0, dh};
    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
        s[n][0] = v[a]; s[n][1] = v[b]; s[n][2] = v[c]; n++;
    }
    if (aa == 3) {
        double q = dh / 2, u[2] = {-q, q};
        for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
            s[n][0] = u[a]; s[n][1] = u[b]; s[n][2] = u[c]; n++;
        }
    }
    return n;
}

/* pm.jl:233-234 */
static int SFX(resample_shifts)(int nr, double (*d)[3]) {
    int n = 0;
    for (int a = 0; a < nr; a++) for (int b = 0; b < nr; b++) for (int c = 0; c < nr; c++) {
        if (a == 0 && b == 0 && c == 0) continue;
        d[n][0] = (double)a / nr; d[n][1] = (double)b / nr; d[n][2] = (double)c / nr; n++;
    }
    return n;
}

/* pm.jl:186 */
static void SFX(lagr1d)(int n, double L, REAL *q) {
    for (int i = 0; i < n; i++) q[i] = (REAL)((double)i * (L / n));
}

/* pm.jl:220-231: periodic trilinear interpolation at grid index (i+d1, j+d2, k+d3) */
static void SFX(trilinear_shift)(const REAL *f, int n, const double *d, REAL *out) {
    REAL dx = (REAL)d[0], dy = (REAL)d[1], dz = (REAL)d[2];
    size_t N = (size_t)n;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int i1 = (i + 1) % n, j1 = (j + 1) % n, k1 = (k + 1) % n;
#define F(a, b, c) f[(a) + N * ((b) + N * (c))]
                out[i + N * (j + N * k)] =
                    F(i, j, k) * (1 - dx) * (1 - dy) * (1 - dz) + F(i1, j, k) * dx * (1 - dy) * (1 - dz) +
                    F(i, j1, k) * (1 - dx) * dy * (1 - dz) + F(i, j, k1) * (1 - dx) * (1 - dy) * dz +
                    F(i1, j1, k) * dx * dy * (1 - dz) + F(i1, j, k1) * dx * (1 - dy) * dz +
                    F(i, j1, k1) * (1 - dx) * dy * dz + F(i1, j1, k1) * dx * dy * dz;
#undef F
            }
}

/* pm.jl:189-217: sheet point at Lagrangian offset d (grid cells), wrapped to [0, L) */
static int SFX(spawn)(const REAL *X, int n, const double *d, const orc_pm_config *cfg, REAL *out) {
    size_t N3 = (size_t)n * n * n, N = (size_t)n;
    REAL L = (REAL)cfg->boxsize;
    REAL *q1 = (REAL *)malloc(sizeof(REAL) * n);
    SFX(lagr1d)(n, cfg->boxsize, q1);
    REAL *psi = (REAL *)malloc(sizeof(REAL) * N3), *npsi = (REAL *)malloc(sizeof(REAL) * N3);
    int nh = n / 2 + 1;
    CPLX *ph = NULL, *fk = NULL;
    if (cfg->resampling == 0) {
        ph = (CPLX *)malloc(sizeof(CPLX) * N * N * nh);
        fk = (CPLX *)malloc(sizeof(CPLX) * N * N * nh);
        REAL tp = (REAL)(2 * M_PI);
        for (int k = 0; k < nh; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    REAL mi = (REAL)((double)(i < n / 2 ? i : i - n) / n);
                    REAL mj = (REAL)((double)(j < n / 2 ? j : j - n) / n);
                    REAL mk = (REAL)((double)k / n);
                    REAL kd = mi * (REAL)d[0] + mj * (REAL)d[1] + mk * (REAL)d[2];
                    REAL arg = tp * kd;
                    ph[i + N * (j + N * k)] = (REAL)cos(arg) + I * (REAL)sin(arg);
                }
    } else if (cfg->resampling != 1) {
        free(q1); free(psi); free(npsi);
        return -1;
    }
    for (int c = 0; c < 3; c++) {
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t p = i + N * (j + N * k);
                    REAL qc = q1[c == 0 ? i : c == 1 ? j : k];
                    psi[p] = SFX(jmod)(X[p + c * N3] - qc + L / 2, L) - L / 2;
                }
        if (cfg->resampling == 0) {
            SFX(orc_rfft3)(n, psi, (REAL *)fk);
            for (size_t i = 0; i < N * N * nh; i++) fk[i] *= ph[i];
            SFX(orc_brfft3)(n, (const REAL *)fk, npsi);
            REAL sc = (REAL)1 / (REAL)(N3);
            for (size_t i = 0; i < N3; i++) npsi[i] *= sc;
        } else {
            SFX(trilinear_shift)(psi, n, d, npsi);
        }
        REAL sh = (REAL)(d[c] * cfg->boxsize / n);
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t p = i + N * (j + N * k);
                    REAL qc = q1[c == 0 ? i : c == 1 ? j : k];
                    out[p + c * N3] = SFX(jmod)(qc + npsi[p] + sh + L, L);
                }
    }
    free(q1); free(psi); free(npsi); free(ph); free(fk);
    return 0;
}

static int SFX(check_cfg)(const orc_pm_config *cfg) {
    if (cfg->worder < 2 || cfg->worder > 4) return -1;
    if (cfg->res_pm < 2 || cfg->res_pm % 2) return -1;
    if (cfg->n_resample < 1) return -1;
    return 0;
}

/* pm.jl:273-298: one interlacing pass of the reference form; acc (np,3) accumulated */
static int SFX(acc_pm_single)(const REAL *Xs, const orc_pm_config *cfg, const SFX(kernels) * K, REAL *acc) {
    int n = cfg->n_part, res = cfg->res_pm, W = cfg->worder, nh = res / 2 + 1;
    int64_t np = (int64_t)n * n * n;
    size_t M = (size_t)res * res * res, H = (size_t)res * res * nh;
    REAL L = (REAL)cfg->boxsize;
    REAL *mesh = (REAL *)calloc(M, sizeof(REAL));
    REAL rhomean = (REAL)((double)np / ((double)res * res * res));
    SFX(orc_pm_scatter)(mesh, Xs, np, res, L, W);
    for (size_t i = 0; i < M; i++) mesh[i] = mesh[i] / rhomean - 1;
    if (cfg->n_resample > 1) {
        double d[64][3];
        int nd = SFX(resample_shifts)(cfg->n_resample, d);
        REAL *m2 = (REAL *)malloc(sizeof(REAL) * M), *Xr = (REAL *)malloc(sizeof(REAL) * 3 * np);
        for (int s = 0; s < nd; s++) {
            if (SFX(spawn)(Xs, n, d[s], cfg, Xr)) { free(m2); free(Xr); free(mesh); return -1; }
            memset(m2, 0, sizeof(REAL) * M);
            SFX(orc_pm_scatter)(m2, Xr, np, res, L, W);
            for (size_t i = 0; i < M; i++) mesh[i] += m2[i] / rhomean - 1;
        }
        REAL nr3 = (REAL)(cfg->n_resample * cfg->n_resample * cfg->n_resample);
        for (size_t i = 0; i < M; i++) mesh[i] /= nr3;
        free(m2); free(Xr);
    }
    CPLX *fd = (CPLX *)malloc(sizeof(CPLX) * H), *fa = (CPLX *)malloc(sizeof(CPLX) * H);
    SFX(orc_rfft3)(res, mesh, (REAL *)fd);
    if (K->mak) for (size_t i = 0; i < H; i++) fd[i] *= K->mak[i];
    fd[0] = 0;
    REAL *field = mesh, *out = (REAL *)malloc(sizeof(REAL) * np);
    REAL scale = (REAL)1 / (REAL)M;
    for (int c = 0; c < 3; c++) {
        for (int z = 0; z < nh; z++)
            for (int y = 0; y < res; y++)
                for (int x = 0; x < res; x++) {
                    size_t i = x + (size_t)res * (y + (size_t)res * z);
                    CPLX g = c == 0 ? K->g[0][x] : c == 1 ? K->g[1][y] : K->g[2][z];
                    fa[i] = -((g * K->invlap[i]) * fd[i]);
                }
        SFX(orc_brfft3)(res, (const REAL *)fa, field);
        for (size_t i = 0; i < M; i++) field[i] *= scale;
        SFX(orc_pm_gather)(out, field, Xs, np, res, L, W);
        for (int64_t p = 0; p < np; p++) acc[p + c * np] += out[p];
    }
    free(mesh); free(fd); free(fa); free(out);
    return 0;
}

/* pm.jl:253-271 */
int SFX(orc_pm_acceleration_ref)(const orc_pm_config *cfg, const REAL *psi, REAL *acc) {
    if (SFX(check_cfg)(cfg)) return -1;
    int n = cfg->n_part;
    int64_t np = (int64_t)n * n * n;
    size_t N = (size_t)n;
    SFX(kernels) K;
    if (SFX(kernels_init)(&K, cfg)) { SFX(kernels_free)(&K); return -1; }
    double sh[16][3];
    int ns = SFX(aa_shifts)(cfg->antialias, 0.5 * cfg->boxsize / cfg->res_pm, sh);
    if (ns < 0) { SFX(kernels_free)(&K); return -1; }
    REAL *q1 = (REAL *)malloc(sizeof(REAL) * n);
    SFX(lagr1d)(n, cfg->boxsize, q1);
    REAL *X = (REAL *)malloc(sizeof(REAL) * 3 * np), *Xs = (REAL *)malloc(sizeof(REAL) * 3 * np);
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t p = i + N * (j + N * k);
                    X[p + c * np] = psi[p + c * np] + q1[c == 0 ? i : c == 1 ? j : k];
                }
    memset(acc, 0, sizeof(REAL) * 3 * np);
    int err = 0;
    for (int s = 0; s < ns && !err; s++) {
        for (int c = 0; c < 3; c++)
            for (int64_t p = 0; p < np; p++) Xs[p + c * np] = X[p + c * np] + (REAL)sh[s][c];
        err = SFX(acc_pm_single)(Xs, cfg, &K, acc);
    }
    if (ns > 1) for (int64_t i = 0; i < 3 * np; i++) acc[i] /= (REAL)ns;
    free(q1); free(X); free(Xs);
    SFX(kernels_free)(&K);
    return err;
}

/* ------------------------------------------------------------------ PMSolver form */
static void SFX(fd_coeffs)(int order, double h, REAL *c) {
    if (order == 2) { c[0] = (REAL)(1 / (2 * h)); c[1] = 0; c[2] = 0; }
    else if (order == 4) { c[0] = (REAL)(8 / (12 * h)); c[1] = (REAL)(-1 / (12 * h)); c[2] = 0; }
    else { c[0] = (REAL)(45 / (60 * h)); c[1] = (REAL)(-9 / (60 * h)); c[2] = (REAL)(1 / (60 * h)); }
}

/* pm.jl:330-342: F_ax = -D_ax phi, periodic central differences */
static void SFX(fdgrad1)(REAL *F, const REAL *phi, int res, int ax, const REAL *c) {
    size_t R = (size_t)res;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < res; k++)
        for (int j = 0; j < res; j++)
            for (int i = 0; i < res; i++) {
                REAL d[3];
                for (int s = 1; s <= 3; s++) {
                    int ip = i, im = i, jp = j, jm = j, kp = k, km = k;
                    if (ax == 0) { ip = SFX(imod)(i + s, res); im = SFX(imod)(i - s, res); }
                    if (ax == 1) { jp = SFX(imod)(j + s, res); jm = SFX(imod)(j - s, res); }
                    if (ax == 2) { kp = SFX(imod)(k + s, res); km = SFX(imod)(k - s, res); }
                    d[s - 1] = phi[ip + R * (jp + R * kp)] - phi[im + R * (jm + R * km)];
                }
                F[i + R * (j + R * k)] = -(c[0] * d[0] + c[1] * d[1] + c[2] * d[2]);
            }
}

/* PMSolver K_phi (pm.jl:390-391) */
static REAL *SFX(kphi)(const SFX(kernels) * K, const orc_pm_config *cfg, int nsets) {
    int res = cfg->res_pm, n = cfg->n_part;
    size_t H = (size_t)res * res * (res / 2 + 1);
    REAL *Kp = (REAL *)malloc(sizeof(REAL) * H);
    REAL rho = (REAL)((double)n * n * n / ((double)res * res * res));
    REAL r3 = (REAL)res * (REAL)res * (REAL)res;
    for (size_t i = 0; i < H; i++) {
        REAL v = K->invlap[i] * (K->mak ? K->mak[i] : (REAL)1);
        Kp[i] = v / rho / (REAL)nsets / r3;
    }
    Kp[0] = 0;
    return Kp;
}

/* pm.jl:466-501: one interlacing pass of PMSolver; acc accumulated */
static int SFX(acc_single)(const REAL *Xs, const orc_pm_config *cfg, const SFX(kernels) * K,
                           const REAL *Kp, REAL *acc) {
    int n = cfg->n_part, res = cfg->res_pm, W = cfg->worder, nh = res / 2 + 1;
    int64_t np = (int64_t)n * n * n;
    size_t M = (size_t)res * res * res, H = (size_t)res * res * nh, R = (size_t)res;
    REAL L = (REAL)cfg->boxsize;
    REAL *mesh = (REAL *)calloc(M, sizeof(REAL));
    SFX(orc_pm_scatter)(mesh, Xs, np, res, L, W);
    if (cfg->n_resample > 1) {
        double d[64][3];
        int nd = SFX(resample_shifts)(cfg->n_resample, d);
        REAL *Xr = (REAL *)malloc(sizeof(REAL) * 3 * np);
        for (int s = 0; s < nd; s++) {
            if (SFX(spawn)(Xs, n, d[s], cfg, Xr)) { free(Xr); free(mesh); return -1; }
            SFX(orc_pm_scatter)(mesh, Xr, np, res, L, W);
        }
        free(Xr);
    }
    /* One force component at a time (F holds one mesh): the per-particle sums of each component are
       the same additions in the same order as in a joint three-component loop, so the result does
       not depend on this, and the 512^3 / 1024^3 FP64 force fits in ~50 GB of host memory. */
    CPLX *fd = (CPLX *)malloc(sizeof(CPLX) * H);
    CPLX *cb = cfg->grad_order == 0 ? (CPLX *)malloc(sizeof(CPLX) * H) : fd;
    SFX(orc_rfft3)(res, mesh, (REAL *)fd);
    REAL *F = (REAL *)malloc(sizeof(REAL) * M), *phi = mesh;
    REAL c[3];
    if (cfg->grad_order != 0) {
        for (size_t i = 0; i < H; i++) cb[i] = fd[i] * Kp[i];
        SFX(orc_brfft3)(res, (const REAL *)cb, phi);
        SFX(fd_coeffs)(cfg->grad_order, cfg->boxsize / res, c);
    }
    for (int ax = 0; ax < 3; ax++) {
        if (cfg->grad_order == 0) {
            for (int z = 0; z < nh; z++)
                for (int y = 0; y < res; y++)
                    for (int x = 0; x < res; x++) {
                        size_t i = x + R * (y + R * z);
                        CPLX g = ax == 0 ? K->g[0][x] : ax == 1 ? K->g[1][y] : K->g[2][z];
                        cb[i] = -((g * Kp[i]) * fd[i]);
                    }
            SFX(orc_brfft3)(res, (const REAL *)cb, F);
        } else {
            SFX(fdgrad1)(F, phi, res, ax, c);
        }
        /* gather3 (pm.jl:344-364), component ax */
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < np; p++) {
            SFX(stencil) st;
            SFX(make_stencil)(&st, Xs[p], Xs[p + np], Xs[p + 2 * np], res, L, W);
            REAL s1 = 0;
            for (int oz = 0; oz < W; oz++)
                for (int oy = 0; oy < W; oy++)
                    for (int ox = 0; ox < W; ox++) {
                        size_t cell = st.cell[0][ox] + R * (st.cell[1][oy] + R * st.cell[2][oz]);
                        REAL w = st.w[0][ox] * st.w[1][oy] * st.w[2][oz];
                        s1 += F[cell] * w;
                    }
            acc[p + ax * np] += s1;
        }
    }
    free(mesh); free(fd); free(F);
    if (cb != fd) free(cb);
    return 0;
}

/* pm.jl:507-521 */
int SFX(orc_pm_acceleration)(const orc_pm_config *cfg, const REAL *psi, REAL *acc) {
    if (SFX(check_cfg)(cfg)) return -1;
    int n = cfg->n_part;
    int64_t np = (int64_t)n * n * n;
    size_t N = (size_t)n;
    SFX(kernels) K;
    if (SFX(kernels_init)(&K, cfg)) { SFX(kernels_free)(&K); return -1; }
    double sh[16][3];
    int ns = SFX(aa_shifts)(cfg->antialias, 0.5 * cfg->boxsize / cfg->res_pm, sh);
    if (ns < 0) { SFX(kernels_free)(&K); return -1; }
    REAL *Kp = SFX(kphi)(&K, cfg, cfg->n_resample * cfg->n_resample * cfg->n_resample);
    free(K.invlap); K.invlap = NULL;
    free(K.mak); K.mak = NULL;
    REAL *q1 = (REAL *)malloc(sizeof(REAL) * n);
    SFX(lagr1d)(n, cfg->boxsize, q1);
    REAL *X = (REAL *)malloc(sizeof(REAL) * 3 * np), *Xs = (REAL *)malloc(sizeof(REAL) * 3 * np);
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t p = i + N * (j + N * k);
                    X[p + c * np] = psi[p + c * np] + q1[c == 0 ? i : c == 1 ? j : k];
                }
    memset(acc, 0, sizeof(REAL) * 3 * np);
    int err = 0;
    for (int s = 0; s < ns && !err; s++) {
        for (int c = 0; c < 3; c++)
            for (int64_t p = 0; p < np; p++) Xs[p + c * np] = X[p + c * np] + (REAL)sh[s][c];
        err = SFX(acc_single)(Xs, cfg, &K, Kp, acc);
    }
    if (ns > 1) for (int64_t i = 0; i < 3 * np; i++) acc[i] /= (REAL)ns;
    free(q1); free(X); free(Xs); free(Kp);
    SFX(kernels_free)(&K);
    return err;
}

/* pm.jl:569-641 */
int SFX(orc_pm_acceleration_lean)(const orc_pm_config *cfg, const REAL *psi, REAL *acc) {
    if (SFX(check_cfg)(cfg)) return -1;
    if (!(cfg->antialias == 0 && cfg->n_resample == 1 &&
          (cfg->grad_order == 2 || cfg->grad_order == 4 || cfg->grad_order == 6)))
        return -1;
    int n = cfg->n_part, res = cfg->res_pm, W = cfg->worder, nh = res / 2 + 1;
    int64_t np = (int64_t)n * n * n;
    size_t M = (size_t)res * res * res, H = (size_t)res * res * nh, N = (size_t)n, R = (size_t)res;
    REAL L = (REAL)cfg->boxsize;
    SFX(kernels) K;
    if (SFX(kernels_init)(&K, cfg)) { SFX(kernels_free)(&K); return -1; }
    REAL *Kp = SFX(kphi)(&K, cfg, 1);
    REAL *q1 = (REAL *)malloc(sizeof(REAL) * n);
    SFX(lagr1d)(n, cfg->boxsize, q1);
    REAL *X = (REAL *)malloc(sizeof(REAL) * 3 * np);
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t p = i + N * (j + N * k);
                    X[p + c * np] = psi[p + c * np] + q1[c == 0 ? i : c == 1 ? j : k];
                }
    REAL *mesh = (REAL *)calloc(M, sizeof(REAL)), *F = (REAL *)malloc(sizeof(REAL) * M);
    CPLX *fd = (CPLX *)malloc(sizeof(CPLX) * H);
    SFX(orc_pm_scatter)(mesh, X, np, res, L, W);
    SFX(orc_rfft3)(res, mesh, (REAL *)fd);
    for (size_t i = 0; i < H; i++) fd[i] *= Kp[i];
    SFX(orc_brfft3)(res, (const REAL *)fd, mesh);
    REAL c[3];
    SFX(fd_coeffs)(cfg->grad_order, cfg->boxsize / res, c);
    for (int ax = 0; ax < 3; ax++) {
        SFX(fdgrad1)(F, mesh, res, ax, c);
        SFX(orc_pm_gather)(acc + ax * np, F, X, np, res, L, W);
    }
    (void)R;
    free(q1); free(X); free(mesh); free(F); free(fd); free(Kp);
    SFX(kernels_free)(&K);
    return 0;
}
