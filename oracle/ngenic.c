/*
 * Oracle (test infrastructure only): N-GenIC white noise from GSL's ranlxd1 generator.
 *   ranlxd1 (RANLUX double precision, luxury p = 202; GSL gsl_rng_set / gsl_rng_uniform) ....
 *       NR/src/ics/ngenic_gsl.jl:21-98
 *   seed-table spiral and Hermitian Fourier fill (rng_ngenic) .... ngenic_gsl.jl:100-165
 *   real-space white noise = irfftn(field * res^(3/2)) ............ ngenic_gsl.jl:174-177
 * Pinned by the known answers of NR/test/runtests.jl:406-421.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    double xd[12];
    double carry;
    int ir, jr, ir_old, pr;
} ranlxd;

static const double ONE_BIT = 1.0 / 281474976710656.0; /* 2^-48 */

static void ranlxd_set(ranlxd *st, uint32_t s) {
    if (s == 0) s = 1;
    uint64_t i = (uint64_t)s & 0xFFFFFFFFu;
    int xbit[31];
    for (int k = 0; k < 31; k++) { xbit[k] = (int)(i & 1); i >>= 1; }
    int ibit = 0, jbit = 18;
    for (int k = 0; k < 12; k++) {
        double x = 0;
        for (int m = 0; m < 48; m++) {
            double y = (double)((xbit[ibit] + 1) % 2);
            x += x + y;
            xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
            ibit = (ibit + 1) % 31;
            jbit = (jbit + 1) % 31;
        }
        st->xd[k] = ONE_BIT * x;
    }
    st->carry = 0;
    st->ir = 11;
    st->jr = 7;
    st->ir_old = 0;
    st->pr = 202;
}

#define NEXT(i) (((i) + 1) % 12)

static void ranlxd_increment(ranlxd *st) {
    double *xd = st->xd, carry = st->carry, y1, y2, y3;
    int ir = st->ir, jr = st->jr, k = 0;
    while (ir > 0) {
        y1 = xd[jr] - xd[ir];
        y2 = y1 - carry;
        if (y2 < 0) { carry = ONE_BIT; y2 += 1; } else { carry = 0; }
        xd[ir] = y2;
        ir = NEXT(ir);
        jr = NEXT(jr);
        k++;
    }
    int kmax = st->pr - 12;
    /* twelve lagged subtract-with-borrow steps per block (lags 5 and 12) */
    while (k <= kmax) {
        y1 = xd[7] - xd[0]; y1 -= carry;
        y2 = xd[8] - xd[1]; if (y1 < 0) { y2 -= ONE_BIT; y1 += 1; } xd[0] = y1;
        y3 = xd[9] - xd[2]; if (y2 < 0) { y3 -= ONE_BIT; y2 += 1; } xd[1] = y2;
        y1 = xd[10] - xd[3]; if (y3 < 0) { y1 -= ONE_BIT; y3 += 1; } xd[2] = y3;
        y2 = xd[11] - xd[4]; if (y1 < 0) { y2 -= ONE_BIT; y1 += 1; } xd[3] = y1;
        y3 = xd[0] - xd[5]; if (y2 < 0) { y3 -= ONE_BIT; y2 += 1; } xd[4] = y2;
        y1 = xd[1] - xd[6]; if (y3 < 0) { y1 -= ONE_BIT; y3 += 1; } xd[5] = y3;
        y2 = xd[2] - xd[7]; if (y1 < 0) { y2 -= ONE_BIT; y1 += 1; } xd[6] = y1;
        y3 = xd[3] - xd[8]; if (y2 < 0) { y3 -= ONE_BIT; y2 += 1; } xd[7] = y2;
        y1 = xd[4] - xd[9]; if (y3 < 0) { y1 -= ONE_BIT; y3 += 1; } xd[8] = y3;
        y2 = xd[5] - xd[10]; if (y1 < 0) { y2 -= ONE_BIT; y1 += 1; } xd[9] = y1;
        y3 = xd[6] - xd[11]; if (y2 < 0) { y3 -= ONE_BIT; y2 += 1; } xd[10] = y2;
        if (y3 < 0) { carry = ONE_BIT; y3 += 1; } else { carry = 0; }
        xd[11] = y3;
        k += 12;
    }
    kmax = st->pr;
    while (k < kmax) {
        y1 = xd[jr] - xd[ir];
        y2 = y1 - carry;
        if (y2 < 0) { carry = ONE_BIT; y2 += 1; } else { carry = 0; }
        xd[ir] = y2;
        ir = NEXT(ir);
        jr = NEXT(jr);
        k++;
    }
    st->ir = ir;
    st->ir_old = ir;
    st->jr = jr;
    st->carry = carry;
}

static double ranlxd_uniform(ranlxd *st) {
    st->ir = NEXT(st->ir);
    if (st->ir == st->ir_old) ranlxd_increment(st);
    return st->xd[st->ir];
}

void orc_ranlxd1_uniforms(uint32_t seed, int n, double *out) {
    ranlxd st;
    ranlxd_set(&st, seed);
    for (int i = 0; i < n; i++) out[i] = ranlxd_uniform(&st);
}

static uint32_t draw(ranlxd *st) { return (uint32_t)floor(2147483647.0 * ranlxd_uniform(st)); }

/* ngenic_gsl.jl:103-119 */
static void seedtable(uint32_t seed, int nres, uint32_t *tab) {
    ranlxd st;
    ranlxd_set(&st, seed);
    memset(tab, 0, sizeof(uint32_t) * (size_t)nres * nres);
    int half = nres / 2;
    for (int i = 0; i < half; i++) {
        for (int j = 0; j < i; j++) tab[i * nres + j] = draw(&st);
        for (int j = 0; j <= i; j++) tab[j * nres + i] = draw(&st);
        for (int j = 0; j < i; j++) tab[(nres - 1 - i) * nres + j] = draw(&st);
        for (int j = 0; j <= i; j++) tab[(nres - 1 - j) * nres + i] = draw(&st);
        for (int j = 0; j < i; j++) tab[i * nres + (nres - 1 - j)] = draw(&st);
        for (int j = 0; j <= i; j++) tab[j * nres + (nres - 1 - i)] = draw(&st);
        for (int j = 0; j < i; j++) tab[(nres - 1 - i) * nres + (nres - 1 - j)] = draw(&st);
        for (int j = 0; j <= i; j++) tab[(nres - 1 - j) * nres + (nres - 1 - i)] = draw(&st);
    }
}

/* ngenic_gsl.jl:126-165; out (res,res,res/2+1) interleaved complex, index i + res*j + res^2*k */
void orc_ngenic_field_gsl(uint32_t seed, int res, double *outr) {
    int nresp = res / 2 + 1, half = res / 2;
    size_t R = (size_t)res;
    double _Complex *fld = (double _Complex *)outr;
    memset(fld, 0, sizeof(double _Complex) * R * R * nresp);
    uint32_t *tab = malloc(sizeof(uint32_t) * R * R);
    seedtable(seed, res, tab);
    double twopi = 2 * M_PI;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int i = 0; i < res; i++)
        for (int j = 0; j < res; j++) {
            int ii = i > 0 ? res - i : 0, jj = j > 0 ? res - j : 0;
            ranlxd st;
            ranlxd_set(&st, tab[i * res + j]);
            for (int k = 0; k <= half; k++) {
                double phase = ranlxd_uniform(&st) * twopi, ampl;
                do { ampl = ranlxd_uniform(&st); } while (ampl == 0.0 || ampl == 1.0);
                if (i == half || j == half || k == half) continue;
                if (i == 0 && j == 0 && k == 0) continue;
                ampl = sqrt(-log(ampl));
                double _Complex z = ampl * cos(phase) + I * (ampl * sin(phase));
                if (k > 0) {
                    fld[i + R * (j + R * k)] = z;
                } else if (i == 0) {
                    if (j < half) {
                        fld[i + R * j] = z;
                        fld[i + R * jj] = conj(z);
                    }
                } else if (i < half) {
                    fld[i + R * j] = z;
                    fld[ii + R * jj] = conj(z);
                }
            }
        }
    free(tab);
}

void orc_ngenic_wnoise_real(uint32_t seed, int res, double *out) {
    size_t H = (size_t)res * res * (res / 2 + 1), M = (size_t)res * res * res;
    double *f = malloc(sizeof(double) * 2 * H);
    orc_ngenic_field_gsl(seed, res, f);
    double sc = pow((double)res, 1.5);
    for (size_t i = 0; i < 2 * H; i++) f[i] *= sc;
    orc_brfft3_d(res, f, out);
    double inv = 1.0 / (double)M;
    for (size_t i = 0; i < M; i++) out[i] *= inv;
    free(f);
}
