/*
 * Oracle (test infrastructure only): mixed-radix Stockham FFT, instantiated for REAL = double/float
 * by fft_d.c / fft_f.c.  Stands in for FFTW's rfft/brfft with region [3,1,2] as used by the
 * reference (NR/src/lpt/nlpt_core.jl:35-37; NR/src/nbody/pm.jl:400-401,416-417): the real axis
 * (halved) is z, the slowest axis of the column-major (N,N,N) array.  Twiddles are built in
 * double and rounded once to REAL.
 *
 * c2r semantics follow FFTW/pocketfft for the halved axis: the complex inverses along x and y are
 * done first, then each z line is inverted as a Hermitian line whose DC and Nyquist entries
 * contribute only their real parts.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef REAL _Complex CPLX;

typedef struct {
    int n, nf, sign;
    int fac[32];
    CPLX *tw; /* tw[k] = exp(sign * 2 pi i k / n) */
} SFX(plan);

static void SFX(plan_init)(SFX(plan) * p, int n, int sign) {
    p->n = n;
    p->sign = sign;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int r = 3; m > 1; r += 2)
        while (m % r == 0) { p->fac[p->nf++] = r; m /= r; }
    p->tw = (CPLX *)malloc(sizeof(CPLX) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        double ang = sign * 2.0 * M_PI * (double)k / (double)n;
        p->tw[k] = (REAL)cos(ang) + I * (REAL)sin(ang);
    }
}

static void SFX(plan_free)(SFX(plan) * p) { free(p->tw); }

/* one line: x (length n) -> transformed in x; y is scratch of length n */
static void SFX(fft_line)(const SFX(plan) * p, CPLX *x, CPLX *y) {
    int n = p->n, s = 1;
    CPLX *src = x, *dst = y;
    CPLX a[64], b[64];
    for (int f = 0; f < p->nf; f++) {
        int R = p->fac[f], m = n / R, L = s * R, step = n / L;
        for (int j = 0; j < m; j++) {
            int k = j % s, jj = j - k;
            for (int r = 0; r < R; r++) {
                CPLX v = src[j + r * m];
                a[r] = (k && r) ? v * p->tw[k * r * step] : v;
            }
            if (R == 2) {
                b[0] = a[0] + a[1];
                b[1] = a[0] - a[1];
            } else if (R == 4) {
                CPLX t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = a[1] - a[3];
                /* multiply t3 by sign*i */
                CPLX t3r = p->sign < 0 ? (cimag(t3) - I * creal(t3)) : (-cimag(t3) + I * creal(t3));
                b[0] = t0 + t2;
                b[2] = t0 - t2;
                b[1] = t1 + t3r;
                b[3] = t1 - t3r;
            } else {
                int st = n / R; /* exp(sign 2 pi i r q / R) = tw[(r q mod R) * n/R] */
                for (int q = 0; q < R; q++) {
                    CPLX acc = a[0];
                    for (int r = 1; r < R; r++) acc += a[r] * p->tw[((r * q) % R) * st];
                    b[q] = acc;
                }
            }
            for (int q = 0; q < R; q++) dst[jj * R + k + q * s] = b[q];
        }
        s = L;
        CPLX *t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(CPLX) * (size_t)n);
}

void SFX(orc_fft1)(int n, const REAL *in, REAL *out, int sign) {
    SFX(plan) p;
    SFX(plan_init)(&p, n, sign < 0 ? -1 : 1);
    CPLX *x = (CPLX *)malloc(sizeof(CPLX) * (size_t)n), *y = (CPLX *)malloc(sizeof(CPLX) * (size_t)n);
    for (int i = 0; i < n; i++) x[i] = in[2 * i] + I * in[2 * i + 1];
    SFX(fft_line)(&p, x, y);
    for (int i = 0; i < n; i++) { out[2 * i] = creal(x[i]); out[2 * i + 1] = cimag(x[i]); }
    free(x); free(y);
    SFX(plan_free)(&p);
}

/* complex transform along axis `ax` (0 = x, 1 = y) of an (N,N,NZ) complex array, in place */
static void SFX(c2c_axis)(CPLX *a, int N, int NZ, int ax, int sign) {
    SFX(plan) p;
    SFX(plan_init)(&p, N, sign);
#pragma omp parallel
    {
        CPLX *x = (CPLX *)malloc(sizeof(CPLX) * (size_t)N), *y = (CPLX *)malloc(sizeof(CPLX) * (size_t)N);
#pragma omp for collapse(2) schedule(static)
        for (int kz = 0; kz < NZ; kz++)
            for (int o = 0; o < N; o++) {
                size_t base = (size_t)kz * N * N, stride;
                if (ax == 0) { base += (size_t)o * N; stride = 1; }
                else { base += (size_t)o; stride = (size_t)N; }
                for (int t = 0; t < N; t++) x[t] = a[base + t * stride];
                SFX(fft_line)(&p, x, y);
                for (int t = 0; t < N; t++) a[base + t * stride] = x[t];
            }
        free(x); free(y);
    }
    SFX(plan_free)(&p);
}

/* rfft(x, [3,1,2]): out (N,N,N/2+1), interleaved complex */
void SFX(orc_rfft3)(int N, const REAL *in, REAL *outr) {
    int H = N / 2 + 1;
    size_t NN = (size_t)N * N;
    CPLX *out = (CPLX *)outr;
    SFX(plan) p;
    SFX(plan_init)(&p, N, -1);
#pragma omp parallel
    {
        CPLX *x = (CPLX *)malloc(sizeof(CPLX) * (size_t)N), *y = (CPLX *)malloc(sizeof(CPLX) * (size_t)N);
        /* z lines, two real lines (ix, ix+1) per complex transform */
#pragma omp for collapse(2) schedule(static)
        for (int iy = 0; iy < N; iy++)
            for (int ix = 0; ix < N; ix += 2) {
                size_t o1 = (size_t)ix + (size_t)N * iy, o2 = o1 + 1;
                for (int t = 0; t < N; t++) x[t] = in[o1 + t * NN] + I * in[o2 + t * NN];
                SFX(fft_line)(&p, x, y);
                for (int k = 0; k < H; k++) {
                    CPLX zk = x[k], zm = conj(x[(N - k) % N]);
                    CPLX A = (zk + zm) * (REAL)0.5;
                    CPLX D = (zk - zm) * (REAL)0.5; /* = i B */
                    CPLX B = cimag(D) - I * creal(D);
                    out[o1 + k * NN] = A;
                    out[o2 + k * NN] = B;
                }
            }
        free(x); free(y);
    }
    SFX(plan_free)(&p);
    SFX(c2c_axis)(out, N, H, 0, -1);
    SFX(c2c_axis)(out, N, H, 1, -1);
}

/* brfft(f, N, [3,1,2]) (unnormalised): in (N,N,N/2+1) interleaved complex, out (N,N,N) real */
void SFX(orc_brfft3)(int N, const REAL *inr, REAL *out) {
    int H = N / 2 + 1;
    size_t NN = (size_t)N * N;
    CPLX *w = (CPLX *)malloc(sizeof(CPLX) * NN * H);
    memcpy(w, inr, sizeof(CPLX) * NN * H);
    SFX(c2c_axis)(w, N, H, 0, +1);
    SFX(c2c_axis)(w, N, H, 1, +1);
    SFX(plan) p;
    SFX(plan_init)(&p, N, +1);
#pragma omp parallel
    {
        CPLX *x = (CPLX *)malloc(sizeof(CPLX) * (size_t)N), *y = (CPLX *)malloc(sizeof(CPLX) * (size_t)N);
#pragma omp for collapse(2) schedule(static)
        for (int iy = 0; iy < N; iy++)
            for (int ix = 0; ix < N; ix += 2) {
                size_t o1 = (size_t)ix + (size_t)N * iy, o2 = o1 + 1;
                for (int k = 0; k < H; k++) {
                    CPLX A = w[o1 + k * NN], B = w[o2 + k * NN];
                    if (k == 0 || 2 * k == N) { A = creal(A); B = creal(B); }
                    x[k] = A + I * B;
                    if (k > 0 && 2 * k < N) x[N - k] = conj(A) + I * conj(B);
                }
                SFX(fft_line)(&p, x, y);
                for (int t = 0; t < N; t++) {
                    out[o1 + t * NN] = creal(x[t]);
                    out[o2 + t * NN] = cimag(x[t]);
                }
            }
        free(x); free(y);
    }
    SFX(plan_free)(&p);
    free(w);
}
