// Initial conditions of the PM path on the device, in FP64 (the IC pipeline of the reference):
//   white noise w ............ N-GenIC field (seed table on the host, ngenic.cpp; the column
//                              generators on the device) -> real space by the library's c2r
//                              (ngenic_wnoise_real, NR/src/ics/ngenic_gsl.jl:174-177)
//   phi(k) ................... rfft(w) * (-sqrt(P(k) (n/L)^3) / k^2), 0 at k = 0 (grf.jl:118-144)
//   psi_1 .................... irfft(-(i k_c) phi)  (nlpt_core.jl:285-297)
//   psi_2 (n_order = 2) ...... de-aliased on ext = 3n/2: A_ik = irfft_ext(i k_k pad(psi1_i)), the
//                              mu2 source A11 A22 + A11 A33 + A22 A33 - A12 A21 - A13 A31 - A23 A32
//                              times -3/7, crop(rfft_ext(.)), psi_2 = irfft(-1/k^2 (i k_c) .)
//                              (nlpt_fast.jl:27-64, 208-249; nlpt_core.jl:126-177)
//   Psi = D psi1 + D^2 psi2, Pi = psi1 + 2 D psi2 (run_nbody.jl:200-206)
// Spectra are in the library's x-halved B layout  B[kz + N (kx + H ky)].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "dj_internal.h"
#include "fft_passes.h"
#include "ic_modes.h"

namespace dj {


// phi(k) = F(w)(k) * scale(k), k grid of grf.jl:190-216, P interpolated as np.interp (clamped)
__global__ void k_ic_scale(cx<double> *B, int N, double L, const double *kt, const double *pt, int ntab) {
    const int H = N / 2 + 1;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)N * H * N;
    if (e >= tot) return;
    BIdx b = bidx(e, N);
    double dk = 2 * M_PI / L;
    // grf.jl indexes the full axes as i <= N/2 ? i : i - N (only k^2 enters)
    double kx = b.kx * dk, ky = (b.ky <= N / 2 ? b.ky : b.ky - N) * dk, kz = (b.kz <= N / 2 ? b.kz : b.kz - N) * dk;
    double k2 = kx * kx + ky * ky + kz * kz;
    double s = 0;
    if (k2 != 0) {
        double k = sqrt(k2), P;
        if (k <= kt[0]) P = pt[0];
        else if (k >= kt[ntab - 1]) P = pt[ntab - 1];
        else {
            int lo = 0, hi = ntab - 1;  // kt[lo] < k <= kt[hi]
            while (hi - lo > 1) {
                int mid = (lo + hi) / 2;
                if (kt[mid] < k) lo = mid; else hi = mid;
            }
            double t = (k - kt[hi - 1]) / (kt[hi] - kt[hi - 1]);
            P = pt[hi - 1] * (1 - t) + pt[hi] * t;
        }
        double rb = (double)N / L;
        s = -sqrt(P * (rb * rb * rb)) / k2;
    }
    B[e] = scale(B[e], s);
}



// zero-pad (n -> e) or crop (e -> n) between B layouts, Nyquist planes dropped, amplitude rescaled
__global__ void k_ic_pad(const cx<double> *in, cx<double> *out, int n, int ext, double s) {
    const int He = ext / 2 + 1, Hn = n / 2 + 1, h = n / 2;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)ext * He * ext;
    if (e >= tot) return;
    BIdx b = bidx(e, ext);
    auto src = [&](int m) { return m < h ? m : (m > ext - h ? m - ext + n : -1); };
    int sx = b.kx < h ? b.kx : -1, sy = src(b.ky), sz = src(b.kz);
    cx<double> v = {0, 0};
    if (sx >= 0 && sy >= 0 && sz >= 0) v = scale(in[(size_t)sz + (size_t)n * (sx + (size_t)Hn * sy)], s);
    out[e] = v;
}
__global__ void k_ic_crop(const cx<double> *in, cx<double> *out, int n, int ext, double s) {
    const int He = ext / 2 + 1, Hn = n / 2 + 1, h = n / 2;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)n * Hn * n;
    if (e >= tot) return;
    BIdx b = bidx(e, n);
    auto src = [&](int m) { return m < h ? m : (m > h ? m - n + ext : -1); };
    int sx = b.kx < h ? b.kx : -1, sy = src(b.ky), sz = src(b.kz);
    cx<double> v = {0, 0};
    if (sx >= 0 && sy >= 0 && sz >= 0) v = scale(in[(size_t)sz + (size_t)ext * (sx + (size_t)He * sy)], s);
    out[e] = v;
}

// mu op= a * b  (op 0 assign, 1 add, 2 subtract; 3 subtract, then multiply by s); the products and sums
// in the reference's order
__global__ void k_ic_mu(double *mu, const double *a, const double *b, size_t n, int op, double s = 1.0) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p = a[i] * b[i];
    mu[i] = op == 0 ? p : op == 1 ? mu[i] + p : op == 2 ? mu[i] - p : (mu[i] - p) * s;
}

// mu from the nine A_ik in one pass (A11, A22, A33, A12, A21, A13, A31, A23, A32): the products and sums of
// the six k_ic_mu calls in their order, (A11 A22 + A11 A33 + A22 A33 - A12 A21 - A13 A31 - A23 A32) s
struct IcFields9 {
    const double *f[9];
};
__global__ void k_ic_mu9(double *__restrict__ mu, IcFields9 A, size_t n, double s) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a11 = A.f[0][i], a22 = A.f[1][i], a33 = A.f[2][i];
    double m = a11 * a22;
    m = m + a11 * a33;
    m = m + a22 * a33;
    m = m - A.f[3][i] * A.f[4][i];
    m = m - A.f[5][i] * A.f[6][i];
    mu[i] = (m - A.f[7][i] * A.f[8][i]) * s;
}

template <typename T>
__global__ void k_ic_combine(const double *p1, const double *p2, T *psi, T *pi, size_t n3, double D, int order) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double a = 0.0 + D * p1[i], b = 0.0 + 1.0 * p1[i];
    if (order >= 2) {
        a = a + (D * D) * p2[i];
        b = b + (2 * D) * p2[i];
    }
    psi[i] = (T)a;
    pi[i] = (T)b;
}

// Adjoint helpers.  On the x-halved spectrum the r2c / c2r pair is adjoint up to the half-axis weights
// w(kx) = 1 at kx = 0 and kx = N/2, 2 otherwise (nlpt_core.jl:682-713): for y = c2r(Y) / M,
// Ybar = w r2c(ybar) / M; for Y = r2c(x), xbar = c2r(Ybar / w).
__global__ void k_ic_halfw(cx<double> *B, int N, int divide, double s) {
    const int H = N / 2 + 1;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)N * H * N;
    if (e >= tot) return;
    const int kx = (int)((e / N) % H);
    const double w = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
    B[e] = scale(B[e], divide ? s / w : s * w);
}
// (Psibar, Pibar) -> (psi1bar, psi2bar) of Psi = D psi1 + D^2 psi2, Pi = psi1 + 2 D psi2
template <typename T>
__global__ void k_ic_combine_adj(const T *psibar, const T *pibar, double *p1b, double *p2b, size_t n3, double D, int order) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const double a = (double)psibar[i], b = (double)pibar[i];
    p1b[i] = D * a + b;
    if (order >= 2) p2b[i] = (D * D) * a + (2 * D) * b;
}

static unsigned g1(size_t n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------- N-GenIC white noise on the device
// The column generators of ngenic_field_gsl (NR/src/ics/ngenic_gsl.jl:126-165): GSL ranlxd1 (RANLUX,
// 48-bit subtract-with-borrow in exact double arithmetic, lags 12 and 5, 202 steps per 12 numbers
// delivered; ngenic_gsl.jl:35-98), seeded from the host's seed table, one thread per (i, j) column.
// The twelve state words of each thread live in LDS (stride 256, so a wave's accesses are consecutive)
// because the recurrence indexes them by a running position.  The uniforms are the same bits as the
// host's; sqrt(-log u), cos and sin are the device's FP64 functions (within an ulp or two of the host
// libm's).
namespace {
constexpr double kNgOneBit = 1.0 / 281474976710656.0;  // 2^-48

struct NgGen {
    double *x;  // 12 words at stride 256
    double carry;
    int ir, jr, ir_old;
    __device__ __forceinline__ double &w(int k) { return x[k * 256]; }
    __device__ void seed(uint32_t s) {
        if (s == 0) s = 1;
        uint32_t bits = s & 0x7fffffffu;  // the 31 seed bits xbit[0..30]
        int ib = 0, jb = 18;
        for (int k = 0; k < 12; k++) {
            double v = 0;
            for (int m = 0; m < 48; m++) {
                const uint32_t b = (bits >> ib) & 1u;
                v += v + (double)(1u - b);
                const uint32_t nb = b ^ ((bits >> jb) & 1u);
                bits = (bits & ~(1u << ib)) | (nb << ib);
                ib = ib == 30 ? 0 : ib + 1;
                jb = jb == 30 ? 0 : jb + 1;
            }
            w(k) = kNgOneBit * v;
        }
        carry = 0;
        ir = 11;
        jr = 7;
        ir_old = 0;
    }
    __device__ __forceinline__ void swb() {
        double y = w(jr) - w(ir) - carry;
        if (y < 0) { carry = kNgOneBit; y += 1; } else { carry = 0; }
        w(ir) = y;
        ir = ir == 11 ? 0 : ir + 1;
        jr = jr == 11 ? 0 : jr + 1;
    }
    __device__ double next() {
        ir = ir == 11 ? 0 : ir + 1;
        if (ir == ir_old) {
            for (int k = 0; k < 202; k++) swb();
            ir_old = ir;
        }
        return w(ir);
    }
};
}  // namespace

__global__ __launch_bounds__(256) void k_ngenic_columns(const uint32_t *__restrict__ tab, int res, double *out) {
    __shared__ double xs[12 * 256];
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)res * res) return;  // no barrier below: each thread's words are its own
    const int i = (int)(t % res), j = (int)(t / res), half = res / 2;
    const int ii = i > 0 ? res - i : 0, jj = j > 0 ? res - j : 0;
    const size_t R = (size_t)res;
    NgGen g;
    g.x = xs + threadIdx.x;
    g.seed(tab[(size_t)i * R + j]);
    const double twopi = 2 * M_PI;
    for (int k = 0; k <= half; k++) {
        const double phase = g.next() * twopi;
        double ampl;
        do { ampl = g.next(); } while (ampl == 0.0 || ampl == 1.0);
        if (i == half || j == half || k == half) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        ampl = sqrt(-log(ampl));
        const double re = ampl * cos(phase), im = ampl * sin(phase);
        auto put = [&](int a, int b, int c, double xr, double xi) {
            const size_t e = (size_t)a + R * ((size_t)b + R * c);
            out[2 * e] = xr;
            out[2 * e + 1] = xi;
        };
        if (k > 0) {
            put(i, j, k, re, im);
        } else if (i == 0) {
            if (j < half) {
                put(0, j, 0, re, im);
                put(0, jj, 0, re, -im);
            }
        } else if (i < half) {
            put(i, j, 0, re, im);
            put(ii, jj, 0, re, -im);
        }
    }
}

__global__ void k_fill_zero_f64(double *p, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

int ngenic_field_device(uint32_t seed, int res, double *d_field, hipStream_t st) {
    const size_t R = (size_t)res, nc = R * R * (res / 2 + 1);
    std::vector<uint32_t> tab(R * R);
    ngenic_seed_table(seed, res, tab.data());
    uint32_t *d_tab = nullptr;
    hipError_t e = hipMalloc(&d_tab, sizeof(uint32_t) * R * R);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tab, tab.data(), sizeof(uint32_t) * R * R, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        k_fill_zero_f64<<<(unsigned)std::min<size_t>(8192, (2 * nc + 255) / 256), 256, 0, st>>>(d_field, 2 * nc);
        k_ngenic_columns<<<g1(R * R), 256, 0, st>>>(d_tab, res, d_field);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_tab) (void)hipFree(d_tab);
    return (int)e;
}

// N-GenIC field (z halved, (n, n, n/2+1)) -> B layout (x halved) by Hermitian symmetry, times n^(3/2)
__global__ void k_field_to_blayout(const cx<double> *__restrict__ f, int n, double sc, cx<double> *__restrict__ B) {
    const int H = n / 2 + 1;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, tot = (size_t)n * H * n;
    if (e >= tot) return;
    const BIdx k = bidx(e, n);
    cx<double> v;
    if (k.kz <= n / 2) {
        v = f[(size_t)k.kx + n * ((size_t)k.ky + (size_t)n * k.kz)];
    } else {
        const int mx = (n - k.kx) % n, my = (n - k.ky) % n, mz = n - k.kz;
        v = conj(f[(size_t)mx + n * ((size_t)my + (size_t)n * mz)]);
    }
    B[e] = scale(v, sc);
}

void ic_work_free(IcWork &ws) {
    for (void *p : ws.bufs) (void)hipFree(p);
    ws = IcWork{};
}

// the buffers of both directions for (n, order) and room for ntab table entries, kept while they fit
static hipError_t ic_work_ensure(IcWork &ws, int n, int order, int ntab) {
    if (ws.n == n && ws.order >= order && ws.ntab >= ntab) return hipSuccess;
    ic_work_free(ws);
    const int ext = 3 * n / 2, Hn = n / 2 + 1, He = ext / 2 + 1;
    const size_t M = (size_t)n * n * n, HC = (size_t)n * Hn * n, ME = (size_t)ext * ext * ext,
                 HE = (size_t)ext * He * ext;
    hipError_t err = hipSuccess;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, bytes);
        if (p) ws.bufs.push_back(p);
        return p;
    };
    auto twiddles = [&](int N) -> void * {
        std::vector<cx<double>> tw(N);
        for (int k = 0; k < N; k++) {
            double a = -2.0 * M_PI * (double)k / (double)N;
            tw[k] = {std::cos(a), std::sin(a)};
        }
        void *d = alloc(sizeof(cx<double>) * N);
        if (d && err == hipSuccess) err = hipMemcpy(d, tw.data(), sizeof(cx<double>) * N, hipMemcpyHostToDevice);
        return d;
    };
    const int cap = std::max(ntab, 4096);
    ws.twn = twiddles(n);
    ws.kt = alloc((size_t)cap * 8);
    ws.pt = alloc((size_t)cap * 8);
    ws.A = alloc(HC * 16);
    ws.Bm = alloc(HC * 16);
    ws.F = alloc(HC * 16);
    ws.Fb = alloc(HC * 16);
    ws.w = alloc(M * 8);
    ws.p1 = alloc(3 * M * 8);
    ws.p2 = alloc(3 * M * 8);
    if (order >= 2) {
        ws.twe = twiddles(ext);
        ws.fLb = alloc(HC * 16);
        ws.Ae = alloc(HE * 16);
        ws.Be = alloc(HE * 16);
        ws.mu = alloc(ME * 8);
        for (auto &f : ws.fld) f = alloc(ME * 8);
    }
    if (err != hipSuccess) {
        ic_work_free(ws);
        return err;
    }
    ws.n = n;
    ws.order = order;
    ws.ntab = cap;
    return hipSuccess;
}

// Room for the nine A_ik fields and the white-noise key, allocated once per workspace; false (and no
// retry) when the device has no room for them, the forward then computing into its three scratch fields
static bool ic_fields_ensure(IcWork &ws, int n) {
    if (ws.fk_room) return ws.fk_room > 0;
    const int ext = 3 * n / 2;
    const size_t M = (size_t)n * n * n, ME = (size_t)ext * ext * ext;
    hipError_t err = hipSuccess;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, bytes);
        if (p) ws.bufs.push_back(p);
        return p;
    };
    for (auto &f : ws.fk) f = alloc(ME * 8);
    ws.wkey = alloc(M * 8);
    ws.flag = alloc(sizeof(unsigned));
    if (err != hipSuccess) {
        (void)hipGetLastError();  // an out-of-memory here is not an error of the call
        ws.fk_room = -1;
        return false;
    }
    ws.fk_room = 1;
    return true;
}

__global__ void k_ic_differs(const unsigned long long *a, const unsigned long long *b, size_t n, unsigned *flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) atomicOr(flag, 1u);
}

// white noise into ws.w: copied from the host or the device, or the N-GenIC field of seed
static hipError_t ic_white(IcWork &ws, FftOps &fn, int n, const double *white, bool on_device, uint32_t seed,
                           hipStream_t st) {
    const size_t M = (size_t)n * n * n, HC = (size_t)n * (n / 2 + 1) * n;
    if (white)
        return hipMemcpyAsync(ws.w, white, M * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st);
    // the z-halved field (n n (n/2+1) complex values, the size of F, which is free until phi(k)), then to
    // the B layout and back to real space
    hipError_t err = (hipError_t)ngenic_field_device(seed, n, (double *)ws.F, st);
    if (err != hipSuccess) return err;
    const KMult<double> none{};
    cx<double> *Bm = (cx<double> *)ws.Bm;
    k_field_to_blayout<<<g1(HC), 256, 0, st>>>((const cx<double> *)ws.F, n, std::pow((double)n, 1.5), Bm);
    fn.zm(ZM_NONE, false, true, Bm, Bm, ws.twn, &none, st);
    fn.yi(Bm, ws.A, ws.twn, st);
    fn.xi_scaled(ws.A, ws.w, ws.twn, 1.0 / (double)M, st);
    return hipGetLastError();
}

// host orchestration; returns a HIP error code (0 = success)
int ics_lpt_device(IcWork &ws, int n, double L, int order, const double *white, bool white_on_device, uint32_t seed,
                   const double *kt_h, const double *pt_h, int ntab, double D, int prec, void *psi_out,
                   void *pi_out, hipStream_t st) {
    FftOps fn{}, fe{};
    const int ext = 3 * n / 2;
    if (!fft_dispatch<double>(n, fn)) return (int)hipErrorInvalidValue;
    if (order >= 2 && !fft_dispatch<double>(ext, fe)) return (int)hipErrorInvalidValue;
    const size_t M = (size_t)n * n * n, HC = (size_t)n * (n / 2 + 1) * n, ME = (size_t)ext * ext * ext;
    hipError_t err = ic_work_ensure(ws, n, order, ntab);
    if (err != hipSuccess) return (int)err;
    void *twn = ws.twn;
    cx<double> *A = (cx<double> *)ws.A, *Bm = (cx<double> *)ws.Bm, *F = (cx<double> *)ws.F;
    double *w = (double *)ws.w, *p1 = (double *)ws.p1, *p2 = (double *)ws.p2;
    double *kt = (double *)ws.kt, *pt = (double *)ws.pt;
    err = hipMemcpyAsync(kt, kt_h, ntab * 8, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipMemcpyAsync(pt, pt_h, ntab * 8, hipMemcpyHostToDevice, st);
    const KMult<double> none{};  // the plain transforms apply no multiplier
    auto fwd = [&](FftOps &f, const double *real, cx<double> *a, cx<double> *b, void *tw) {
        f.xf(real, false, a, tw, 1.0, st);
        f.yf(a, b, tw, st);
        f.zm(ZM_NONE, true, false, b, b, tw, &none, st);
    };
    const bool keep = order >= 2 && ic_fields_ensure(ws, n);
    ws.fk_valid = false;
    if (err == hipSuccess) err = ic_white(ws, fn, n, white, white_on_device, seed, st);
    if (err == hipSuccess && keep && white) err = hipMemcpyAsync(ws.wkey, w, M * 8, hipMemcpyDeviceToDevice, st);
    if (err != hipSuccess) goto done;
    // phi(k)
    fwd(fn, w, A, F, twn);
    k_ic_scale<<<g1(HC), 256, 0, st>>>(F, n, L, kt, pt, ntab);
    for (int c = 0; c < 3; c++) {
        fn.zm_mult_inv(IC_PSI1, F, Bm, c, L, twn, st);  // the multiplier applied as the z pass loads
        fn.yi(Bm, A, twn, st);
        fn.xi_scaled(A, p1 + c * M, twn, 1.0 / (double)M, st);
    }
    if (order >= 2) {
        void *twe = ws.twe;
        cx<double> *Ae = (cx<double> *)ws.Ae, *Be = (cx<double> *)ws.Be;
        double *mu = (double *)ws.mu;
        double *fld[3] = {(double *)ws.fld[0], (double *)ws.fld[1], (double *)ws.fld[2]};
        // the nine A_ik: into the kept fields (A11, A22, A33, then A12, A21, A13, A31, A23, A32), or
        // through the scratch fields
        double *fa[9];
        for (int j = 0; j < 9; j++) fa[j] = keep ? (double *)ws.fk[j] : fld[j < 3 ? j : (j - 3) % 2];
        const double pad_s = std::pow((double)ext / n, 3), crop_s = std::pow((double)n / ext, 3);
        // A_ik = irfft_ext(i k_k pad(psi1_i)); psi1_i spectrum = -(i k_i) phi
        auto deriv = [&](int i, int k, double *out) {
            fe.zm_deriv_pad(F, Be, n, L, i, k, pad_s, twe, st);
            fe.yi(Be, Ae, twe, st);
            fe.xi_scaled(Ae, out, twe, 1.0 / (double)ME, st);
        };
        // mu = A11 A22 + A11 A33 + A22 A33 - A12 A21 - A13 A31 - A23 A32 (evaluated left to right)
        // the last product also applies the order-2 factor ((3-2)/2 - 2) / ((2 + 3/2)(2 - 1)) = -3/7
        const double f2 = ((3.0 - 2) / 2 - 2.0) / ((2 + 1.5) * 1.0);
        const int pairs[3][2][2] = {{{0, 1}, {1, 0}}, {{0, 2}, {2, 0}}, {{1, 2}, {2, 1}}};
        if (keep) {
            // all nine fields first, then mu in one pass
            IcFields9 A9;
            for (int d = 0; d < 3; d++) deriv(d, d, fa[d]);
            for (int q = 0; q < 3; q++) {
                deriv(pairs[q][0][0], pairs[q][0][1], fa[3 + 2 * q]);
                deriv(pairs[q][1][0], pairs[q][1][1], fa[4 + 2 * q]);
            }
            for (int j = 0; j < 9; j++) A9.f[j] = fa[j];
            k_ic_mu9<<<g1(ME), 256, 0, st>>>(mu, A9, ME, f2);
        } else {
            deriv(0, 0, fa[0]);
            deriv(1, 1, fa[1]);
            deriv(2, 2, fa[2]);
            k_ic_mu<<<g1(ME), 256, 0, st>>>(mu, fa[0], fa[1], ME, 0);
            k_ic_mu<<<g1(ME), 256, 0, st>>>(mu, fa[0], fa[2], ME, 1);
            k_ic_mu<<<g1(ME), 256, 0, st>>>(mu, fa[1], fa[2], ME, 1);
            for (int q = 0; q < 3; q++) {
                deriv(pairs[q][0][0], pairs[q][0][1], fa[3 + 2 * q]);
                deriv(pairs[q][1][0], pairs[q][1][1], fa[4 + 2 * q]);
                k_ic_mu<<<g1(ME), 256, 0, st>>>(mu, fa[3 + 2 * q], fa[4 + 2 * q], ME, q < 2 ? 2 : 3, f2);
            }
        }
        fwd(fe, mu, Ae, Be, twe);
        k_ic_crop<<<g1(HC), 256, 0, st>>>(Be, A, n, ext, crop_s);  // fL in A
        for (int c = 0; c < 3; c++) {
            fn.zm_mult_inv(IC_PSI2, A, Bm, c, L, twn, st);
            fn.yi(Bm, F, twn, st);  // F is free now
            fn.xi_scaled(F, p2 + c * M, twn, 1.0 / (double)M, st);
        }
    }
    if (prec == 64)
        k_ic_combine<double><<<g1(3 * M), 256, 0, st>>>(p1, p2, (double *)psi_out, (double *)pi_out, 3 * M, D, order);
    else
        k_ic_combine<float><<<g1(3 * M), 256, 0, st>>>(p1, p2, (float *)psi_out, (float *)pi_out, 3 * M, D, order);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err == hipSuccess && keep) {
        ws.fk_valid = true;
        ws.fk_seeded = !white;
        ws.fk_seed = seed;
        ws.fk_n = n;
        ws.fk_order = order;
        ws.fk_L = L;
        ws.fk_kt.assign(kt_h, kt_h + ntab);
        ws.fk_pt.assign(pt_h, pt_h + ntab);
    }
done:
    (void)hipStreamSynchronize(st);
    return (int)err;
}

// Reverse mode of ics_lpt_device: (Psibar, Pibar) of the ICs (device, precision prec) -> the white-noise
// cotangent wbar (n^3 FP64, host or device).  The forward fields are recomputed (phi(k), the nine A_ik on
// the ext grid, three at a time); each linear step is transposed with the adjoint rules above, and the 2LPT
// source mu is a quadratic form whose cotangent per field is mubar times its partner in mu.
int ics_lpt_vjp_device(IcWork &ws, int n, double L, int order, const double *white, bool white_on_device,
                       uint32_t seed, const double *kt_h, const double *pt_h, int ntab, double D, int prec,
                       const void *psibar, const void *pibar, double *wbar, bool wbar_on_device, hipStream_t st) {
    FftOps fn{}, fe{};
    const int ext = 3 * n / 2;
    if (!fft_dispatch<double>(n, fn)) return (int)hipErrorInvalidValue;
    if (order >= 2 && !fft_dispatch<double>(ext, fe)) return (int)hipErrorInvalidValue;
    const size_t M = (size_t)n * n * n, HC = (size_t)n * (n / 2 + 1) * n, ME = (size_t)ext * ext * ext,
                 HE = (size_t)ext * (ext / 2 + 1) * ext;
    hipError_t err = ic_work_ensure(ws, n, order, ntab);
    if (err != hipSuccess) return (int)err;
    void *twn = ws.twn;
    cx<double> *A = (cx<double> *)ws.A, *Bm = (cx<double> *)ws.Bm, *F = (cx<double> *)ws.F,
               *Fb = (cx<double> *)ws.Fb;
    double *w = (double *)ws.w, *p1b = (double *)ws.p1, *p2b = (double *)ws.p2;
    double *kt = (double *)ws.kt, *pt = (double *)ws.pt;
    err = hipMemcpyAsync(kt, kt_h, ntab * 8, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipMemcpyAsync(pt, pt_h, ntab * 8, hipMemcpyHostToDevice, st);
    const KMult<double> none{};
    auto fwd = [&](FftOps &f, const double *real, cx<double> *a, cx<double> *b, void *tw) {
        f.xf(real, false, a, tw, 1.0, st);
        f.yf(a, b, tw, st);
        f.zm(ZM_NONE, true, false, b, b, tw, &none, st);
    };
    auto inv = [&](FftOps &f, cx<double> *b, cx<double> *a, double *real, void *tw) {  // destroys b
        f.zm(ZM_NONE, false, true, b, b, tw, &none, st);
        f.yi(b, a, tw, st);
        f.xi(a, real, tw, st);
    };
    // the same with the output multiplied by sc as it is stored (a normalisation pass saved)
    auto inv_s = [&](FftOps &f, cx<double> *b, cx<double> *a, double *real, void *tw, double sc) {
        f.zm(ZM_NONE, false, true, b, b, tw, &none, st);
        f.yi(b, a, tw, st);
        f.xi_scaled(a, real, tw, sc, st);
    };
    // The A_ik of the last forward serve when they were computed from the same inputs: n, L, order, the
    // same P(k) table values, and the same seed or white-noise values (compared bit for bit on the device)
    bool hit = order >= 2 && ws.fk_valid && ws.fk_n == n && ws.fk_order == order && ws.fk_L == L &&
               ws.fk_seeded == !white && (white || ws.fk_seed == seed) && ws.fk_kt.size() == (size_t)ntab &&
               std::equal(kt_h, kt_h + ntab, ws.fk_kt.begin()) && std::equal(pt_h, pt_h + ntab, ws.fk_pt.begin());
    if (err == hipSuccess && (!hit || white)) err = ic_white(ws, fn, n, white, white_on_device, seed, st);
    if (err == hipSuccess && hit && white) {
        unsigned differs = 1;
        err = hipMemsetAsync(ws.flag, 0, sizeof(unsigned), st);
        if (err == hipSuccess) {
            k_ic_differs<<<g1(M), 256, 0, st>>>((const unsigned long long *)w, (const unsigned long long *)ws.wkey, M,
                                                (unsigned *)ws.flag);
            err = hipMemcpyAsync(&differs, ws.flag, sizeof(unsigned), hipMemcpyDeviceToHost, st);
        }
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        hit = err == hipSuccess && !differs;
    }
    if (err != hipSuccess) goto done;
    // forward phi(k) (only the recomputed A_ik need it), then the state cotangents split into psi1bar, psi2bar
    if (!hit) {
        fwd(fn, w, A, F, twn);
        k_ic_scale<<<g1(HC), 256, 0, st>>>(F, n, L, kt, pt, ntab);
    }
    if (prec == 64)
        k_ic_combine_adj<double><<<g1(3 * M), 256, 0, st>>>((const double *)psibar, (const double *)pibar, p1b, p2b,
                                                          3 * M, D, order);
    else
        k_ic_combine_adj<float><<<g1(3 * M), 256, 0, st>>>((const float *)psibar, (const float *)pibar, p1b, p2b,
                                                         3 * M, D, order);
    (void)hipMemsetAsync(Fb, 0, HC * 16, st);
    // psi1_c = c2r(-(i k_c) phi) / M:  phibar += (i k_c) w r2c(psi1bar_c) / M
    for (int c = 0; c < 3; c++) {
        // the weights, the multiplier and the accumulation applied as the z pass stores
        fn.xf(p1b + c * M, false, A, twn, 1.0, st);
        fn.yf(A, Bm, twn, st);
        fn.zm_fwd_mult_acc(IC_DERIV, Bm, Fb, c, L, 1.0 / (double)M, twn, st);
    }
    if (order >= 2) {
        void *twe = ws.twe;
        cx<double> *Ae = (cx<double> *)ws.Ae, *Be = (cx<double> *)ws.Be;
        double *mub = (double *)ws.mu;
        double *fld[3] = {(double *)ws.fld[0], (double *)ws.fld[1], (double *)ws.fld[2]};
        cx<double> *fLb = (cx<double> *)ws.fLb;
        const double pad_s = std::pow((double)ext / n, 3), crop_s = std::pow((double)n / ext, 3);
        // psi2_c = c2r((-1/k^2)(i k_c) fL) / M:  fLbar = sum_c (1/k^2)(i k_c) w r2c(psi2bar_c) / M
        (void)hipMemsetAsync(fLb, 0, HC * 16, st);
        for (int c = 0; c < 3; c++) {
            fn.xf(p2b + c * M, false, A, twn, 1.0, st);
            fn.yf(A, Bm, twn, st);
            fn.zm_fwd_mult_acc(IC_PSI2_ADJ, Bm, fLb, c, L, 1.0 / (double)M, twn, st);
        }
        // fL = crop(r2c_ext(-3/7 mu)) crop_s:  mubar = -3/7 c2r_ext(pad(fLbar) crop_s / w)
        k_ic_pad<<<g1(HE), 256, 0, st>>>(fLb, Be, n, ext, crop_s);
        k_ic_halfw<<<g1(HE), 256, 0, st>>>(Be, ext, 1, 1.0);
        inv_s(fe, Be, Ae, mub, twe, ((3.0 - 2) / 2 - 2.0) / ((2 + 1.5) * 1.0));
        // A_ik = c2r_ext((i k_k) pad(-(i k_i) phi) pad_s) / ME
        auto deriv = [&](int i, int k, double *out) {
            fe.zm_deriv_pad(F, Be, n, L, i, k, pad_s, twe, st);
            fe.yi(Be, Ae, twe, st);
            fe.xi_scaled(Ae, out, twe, 1.0 / (double)ME, st);
        };
        // Abar_ik -> phibar: w r2c_ext(Abar) / ME, conj(i k_k), crop pad_s, conj(-(i k_i))
        // with the field Abar_ik = -(mubar a) (neg) or mubar (a + b) formed by the transform as it loads it
        auto deriv_adj = [&](int i, int k, const double *a, const double *b, int neg) {
            fe.xf_partner(mub, a, b, neg, Ae, twe, st);
            fe.yf(Ae, Be, twe, st);
            fe.zm_deriv_adj(Be, Fb, n, L, i, k, 1.0 / (double)ME, pad_s, twe, st);
        };
        // field j of the forward's order (A11, A22, A33, A12, A21, A13, A31, A23, A32): kept, or recomputed
        // into a scratch field
        auto field = [&](int j, int i, int k, double *scratch) -> const double * {
            if (hit) return (const double *)ws.fk[j];
            deriv(i, k, scratch);
            return scratch;
        };
        // mu = A11 A22 + A11 A33 + A22 A33 - A12 A21 - A13 A31 - A23 A32
        const double *dg[3];
        for (int d = 0; d < 3; d++) dg[d] = field(d, d, d, fld[d]);
        for (int d = 0; d < 3; d++) deriv_adj(d, d, dg[(d + 1) % 3], dg[(d + 2) % 3], 0);
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int q = 0; q < 3; q++) {
            const int i = pairs[q][0], k = pairs[q][1];
            const double *aik = field(3 + 2 * q, i, k, fld[0]), *aki = field(4 + 2 * q, k, i, fld[1]);
            deriv_adj(i, k, aki, nullptr, 1);
            deriv_adj(k, i, aik, nullptr, 1);
        }
    }
    // phi = r2c(w) s(k):  wbar = c2r(phibar s / w)
    k_ic_scale<<<g1(HC), 256, 0, st>>>(Fb, n, L, kt, pt, ntab);
    k_ic_halfw<<<g1(HC), 256, 0, st>>>(Fb, n, 1, 1.0);
    inv(fn, Fb, A, w, twn);
    err = hipGetLastError();
    if (err == hipSuccess)
        err = hipMemcpyAsync(wbar, w, M * 8, wbar_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
done:
    (void)hipStreamSynchronize(st);
    return (int)err;
}

}  // namespace dj
