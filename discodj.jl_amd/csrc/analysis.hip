// Power spectrum of a deposited density on the device: the accuracy measurement of the PM path.
// Follows the study estimator (studies/eft_shell_crossing/src/measure.jl:403-447) and the
// normalisation of NR/src/analysis/power_spectrum.jl:25-79:
//   delta = mesh * N^3 / n^3 - 1,  P(k) = V <w |delta_k|^2 / |W(k)|^2>_bin / N^6
// with mode weights w = 1 on the two planes the half spectrum keeps unpaired (here kx = 0 and
// kx = N/2) and 2 elsewhere, |W(k)|^2 = prod_a sinc(k_a dx / 2)^(2 W) the mass-assignment window, and
// linear bins b = [kf (b + 1/2), kf (b + 3/2)).  The spectrum is the library's x-halved B layout.
#include <hip/hip_runtime.h>

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

}  // namespace dj
