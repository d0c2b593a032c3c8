// Power spectrum of a deposited density on the device: the accuracy measurement of the PM path.
// Follows the study estimator (studies/eft_shell_crossing/src/measure.jl:403-447) and the
// normalisation of NR/src/analysis/power_spectrum.jl:25-79:
//   delta = mesh * N^3 / n^3 - 1,  P(k) = V <w |delta_k|^2 / |W(k)|^2>_bin / N^6
// with mode weights w = 1 on the two planes the half spectrum keeps unpaired (here kx = 0 and
// kx = N/2) and 2 elsewhere, |W(k)|^2 = prod_a sinc(k_a dx / 2)^(2 W) the mass-assignment window, and
// linear bins b = [kf (b + 1/2), kf (b + 3/2)).  The spectrum is the library's x-halved B layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "dj_internal.h"
#include "fft_passes.h"

namespace dj {

template <typename T>
__global__ __launch_bounds__(256) void k_pk_bin(const cx<T> *__restrict__ B, int N, double L, int W, int nbins,
                                                double scale, double *__restrict__ sum_p, double *__restrict__ sum_w,
                                                unsigned long long *__restrict__ cnt) {
    extern __shared__ double sh[];
    double *sp = sh, *sw = sh + nbins;
    unsigned long long *sc = (unsigned long long *)(sh + 2 * nbins);
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) { sp[b] = 0; sw[b] = 0; sc[b] = 0; }
    __syncthreads();
    const int H = N / 2 + 1;
    const size_t tot = (size_t)N * H * N;
    const double kf = 2 * M_PI / L, dx = L / N;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        int kz = (int)(e % N);
        size_t t = e / N;
        int kx = (int)(t % H), ky = (int)(t / H);
        double k3[3] = {kx * kf, (ky < N / 2 ? ky : ky - N) * kf, (kz < N / 2 ? kz : kz - N) * kf};
        double k = sqrt(k3[0] * k3[0] + k3[1] * k3[1] + k3[2] * k3[2]);
        int b = (int)floor(k / kf - 0.5);
        if (k == 0 || b < 0 || b >= nbins) continue;
        double win = 1;
        if (W > 0)
            for (int a = 0; a < 3; a++) {
                double u = k3[a] * dx / 2, s = u == 0 ? 1.0 : sin(u) / u;
                win *= pow(s, 2 * W);
            }
        cx<T> v = B[e];
        double p = ((double)v.x * v.x + (double)v.y * v.y) * scale;
        double w = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
        atomicAdd(&sp[b], w * p / win);
        atomicAdd(&sw[b], w);
        atomicAdd(&sc[b], 1ull);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
        if (sc[b]) {
            atomicAdd(&sum_p[b], sp[b]);
            atomicAdd(&sum_w[b], sw[b]);
            atomicAdd(&cnt[b], sc[b]);
        }
    }
}

template <typename T>
void pk_bin(const void *B, int N, double L, int W, int nbins, double scale, double *sum_p, double *sum_w,
            unsigned long long *cnt, hipStream_t st) {
    size_t shm = (size_t)nbins * (2 * sizeof(double) + sizeof(unsigned long long));
    k_pk_bin<T><<<2048, 256, shm, st>>>((const cx<T> *)B, N, L, W, nbins, scale, sum_p, sum_w, cnt);
}

template void pk_bin<float>(const void *, int, double, int, int, double, double *, double *, unsigned long long *,
                            hipStream_t);
template void pk_bin<double>(const void *, int, double, int, int, double, double *, double *,
                             unsigned long long *, hipStream_t);


// ------------------------------------------------------------------ binned spectra over given edges
// The binning of evaluate_power_spectrum / evaluate_cross_power_spectrum (NR/src/analysis/power_spectrum.jl:
// 25-123) and of the study estimator (studies/eft_shell_crossing/src/measure.jl:403-447), on the library's
// x-halved spectrum.  The reference halves the z axis: its modes are (i, j, l), i, j < N, l <= N/2, with
// signed wavenumbers (i <= N/2 ? i : i - N) kf.  One thread takes one of those modes, forms |k| exactly as
// the reference does, sqrt((kx kx + ky ky) + kz kz) in FP64, so that every mode lands in the reference's
// bin, and reads the value from the library's storage: (x = i, y = j, z = l) when i <= N/2, else the
// conjugate-symmetric (N - i, -j, -l) (the value |F|^2, or Re conj(F1) F2, is the same).  Per bin it sums
// |k|, w v, w and the count in LDS, then once per block into sums (4 nbins doubles).
//   wmode 0: w = 1/2 on the l = 0 and l = N/2 planes, 1 elsewhere (the auto spectrum and the study
//            estimator, whose 1 / 2 weights have the same ratio); wmode 1: w = 1 (the cross spectrum)
//   W > 0:   v is divided by |W(k)|^2 = prod_a sinc(k_a dx / 2)^(2 W) (mass-assignment window)
template <typename T>
__global__ __launch_bounds__(256) void k_pk_edges(const cx<T> *__restrict__ F1, const cx<T> *__restrict__ F2, int N,
                                                  double kf, double dx, int W, int wmode, double scale, PkEdges ed,
                                                  double *__restrict__ sums) {
    __shared__ double se[PK_MAX_BINS + 1];
    __shared__ double sk[PK_MAX_BINS], sp[PK_MAX_BINS], sw[PK_MAX_BINS], sc[PK_MAX_BINS];
    const int nb = ed.nbins;
    for (int b = threadIdx.x; b <= nb; b += blockDim.x) se[b] = ed.e[b];
    for (int b = threadIdx.x; b < nb; b += blockDim.x) { sk[b] = 0; sp[b] = 0; sw[b] = 0; sc[b] = 0; }
    __syncthreads();
    const int H = N / 2 + 1;
    const size_t tot = (size_t)N * N * H;
    const double lo = se[0], hi = se[nb];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(e % H);
        const size_t t = e / H;
        const int i = (int)(t % N), j = (int)(t / N);
        const double kx = (double)(i <= N / 2 ? i : i - N) * kf, ky = (double)(j <= N / 2 ? j : j - N) * kf,
                     kz = (double)l * kf;
        const double k = sqrt(kx * kx + ky * ky + kz * kz);
        if (!(k > 0) || k < lo || !(k < hi)) continue;
        int a = 0, b = nb;  // se[a] <= k < se[b]
        while (b - a > 1) {
            const int m = (a + b) >> 1;
            if (se[m] <= k) a = m; else b = m;
        }
        size_t idx;
        if (i <= N / 2) idx = (size_t)l + (size_t)N * ((size_t)i + (size_t)H * j);
        else idx = (size_t)((N - l) % N) + (size_t)N * ((size_t)(N - i) + (size_t)H * ((N - j) % N));
        const cx<T> u = F1[idx];
        double v;
        if (F2) {
            const cx<T> w2 = F2[idx];
            v = ((double)u.x * w2.x + (double)u.y * w2.y) * scale;
        } else {
            v = ((double)u.x * u.x + (double)u.y * u.y) * scale;
        }
        if (W > 0) {
            const double k3[3] = {kx, ky, kz};
            double win = 1;
            for (int ax = 0; ax < 3; ax++) {
                const double s0 = k3[ax] * dx / 2, s = s0 == 0 ? 1.0 : sin(s0) / s0;
                win *= pow(s, 2 * W);
            }
            v /= win;
        }
        const double w = wmode == 1 ? 1.0 : ((l == 0 || l == N / 2) ? 0.5 : 1.0);
        atomicAdd(&sk[a], k);
        atomicAdd(&sp[a], w * v);
        atomicAdd(&sw[a], w);
        atomicAdd(&sc[a], 1.0);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        if (sc[b] > 0) {
            atomicAdd(&sums[b], sk[b]);
            atomicAdd(&sums[nb + b], sp[b]);
            atomicAdd(&sums[2 * nb + b], sw[b]);
            atomicAdd(&sums[3 * nb + b], sc[b]);
        }
    }
}

// Per bin: k_out = mean |k| of its modes (the geometric centre sqrt(e_b e_b+1) of an empty bin, or
// always when centre_only), pk_out = mult sum(w v) / sum(w) (0 when empty), nmodes_out = count / ncomp.
__global__ void k_pk_finalize(const double *__restrict__ sums, PkEdges ed, double mult, int ncomp, int centre_only,
                              double *__restrict__ k_out, double *__restrict__ pk_out, long long *__restrict__ nm_out) {
    const int nb = ed.nbins;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double cnt = sums[3 * nb + b], sw = sums[2 * nb + b];
        if (k_out) k_out[b] = (cnt > 0 && !centre_only) ? sums[b] / cnt : sqrt(ed.e[b] * ed.e[b + 1]);
        if (pk_out) pk_out[b] = cnt > 0 ? mult * sums[nb + b] / sw : 0.0;
        if (nm_out) nm_out[b] = (long long)(cnt / ncomp);
    }
}

__global__ void k_fill_zero_d(double *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

template <typename T>
void pk_edges(const void *F1, const void *F2, int N, double L, int W, int wmode, double scale, const PkEdges &ed,
              double *sums, hipStream_t st) {
    const size_t tot = (size_t)N * N * (N / 2 + 1);
    const int blocks = (int)std::min<size_t>(4096, (tot + 255) / 256);
    k_pk_edges<T><<<blocks, 256, 0, st>>>((const cx<T> *)F1, (const cx<T> *)F2, N, 2 * M_PI / L, L / N, W, wmode,
                                          scale, ed, sums);
}
template void pk_edges<float>(const void *, const void *, int, double, int, int, double, const PkEdges &, double *,
                              hipStream_t);
template void pk_edges<double>(const void *, const void *, int, double, int, int, double, const PkEdges &, double *,
                               hipStream_t);

void pk_finalize(const double *sums, const PkEdges &ed, double mult, int ncomp, bool centre_only, double *k_out,
                 double *pk_out, int64_t *nm_out, hipStream_t st) {
    k_pk_finalize<<<1, 256, 0, st>>>(sums, ed, mult, ncomp, centre_only ? 1 : 0, k_out, pk_out, (long long *)nm_out);
}

void zero_doubles(double *p, int n, hipStream_t st) { k_fill_zero_d<<<(n + 255) / 256, 256, 0, st>>>(p, n); }

// The scratch of dj_field_power_spectrum_noalloc: the bin sums (4 PK_MAX_BINS doubles) in a
// file-scope device array, one per device, held until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns on that device (pk_scratch).
static __device__ __align__(256) double scratch[4 * PK_MAX_BINS];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};
}  // namespace

hipError_t pk_scratch_init(int dev) {
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[dev] = (double *)p;
    if (cur != dev) (void)hipSetDevice(cur);
    return e;
}

double *pk_scratch(int dev) {
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

}  // namespace dj
