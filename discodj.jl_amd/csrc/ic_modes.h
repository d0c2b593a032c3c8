// Mode indexing of the IC pipeline's spectra (x-halved B layout B[kz + N (kx + H ky)]) and the 2LPT
// derivative spectra in the form the FP64 z pass loads or stores them (fused into that pass, so the ext
// spectrum between the multipliers and the z transform is not written).
#pragma once

#include <hip/hip_runtime.h>

#include "dj_complex.h"

namespace dj {

struct BIdx {
    int kx, ky, kz;
};
__device__ __forceinline__ BIdx bidx(size_t e, int N) {
    const int H = N / 2 + 1;
    BIdx r;
    r.kz = (int)(e % N);
    size_t t = e / N;
    r.kx = (int)(t % H);
    r.ky = (int)(t / H);
    return r;
}
__device__ __forceinline__ int sgn_full(int m, int N) { return m < N / 2 ? m : m - N; }

// i k_c along a physical axis with the Nyquist entry zeroed (nlpt_core.jl:49-67)
__device__ __forceinline__ double kdir(const BIdx &b, int c, int N, double dk) {
    if (c == 0) return b.kx == N / 2 ? 0.0 : b.kx * dk;
    int m = c == 1 ? b.ky : b.kz;
    return m == N / 2 ? 0.0 : sgn_full(m, N) * dk;
}

enum { IC_PSI1 = 0, IC_DERIV = 1, IC_PSI2 = 2, IC_PSI2_ADJ = 3 };

// mult(k) v at entry e:  PSI1: -(i k_c) ;  DERIV: i k_c ;  PSI2: (-1/k^2) (i k_c) ;  PSI2_ADJ: the complex
// conjugate of PSI2's multiplier, (1/k^2) (i k_c)  (the adjoint of a complex multiplier is its conjugate)
template <int MODE>
__device__ __forceinline__ cx<double> ic_mult_at(size_t e, cx<double> v, int N, double L, int c) {
    const BIdx b = bidx(e, N);
    const double dk = 2 * M_PI / L;
    const double g = kdir(b, c, N, dk);
    cx<double> r = {-g * v.y, g * v.x};  // (i g) v
    if constexpr (MODE == IC_PSI1) r = {-r.x, -r.y};
    if constexpr (MODE == IC_PSI2 || MODE == IC_PSI2_ADJ) {
        const double kx = b.kx * dk, ky = sgn_full(b.ky, N) * dk, kz = sgn_full(b.kz, N) * dk;
        const double k2 = (kx * kx + ky * ky) + kz * kz;
        const double il = -1.0 / (e == 0 ? 1.0 : k2);
        r = scale(r, MODE == IC_PSI2 ? il : -il);
    }
    return r;
}
// the multiplier applied to in[e] as a z pass loads it
template <int MODE> struct MultSrc {
    const cx<double> *in;
    int N, c;
    double L;
    __device__ cx<double> operator()(size_t e) const { return ic_mult_at<MODE>(e, in[e], N, L, c); }
};
// a z pass's output v at e: the half-spectrum weight times s (1 at kx = 0 and N/2, else 2), the multiplier,
// then added into acc[e]
template <int MODE> struct MultAccDst {
    cx<double> *acc;
    int N, c;
    double L, s;
    __device__ void operator()(size_t e, cx<double> v) const {
        const int H = N / 2 + 1, kx = (int)((e / N) % H);
        const double w = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
        const cx<double> r = ic_mult_at<MODE>(e, scale(v, s * w), N, L, c);
        acc[e] = {acc[e].x + r.x, acc[e].y + r.y};
    }
};

// Entry e of one A_ik spectrum on the ext grid from phi (F, n grid): the steps of the PSI1 multiplier
// (n grid, axis i), the pad (scale s) and the DERIV multiplier (ext grid, axis k), in their order
struct DerivPadSrc {
    const cx<double> *F;
    int n, ext, i, k;
    double L, s;
    __device__ cx<double> operator()(size_t e) const {
        const int Hn = n / 2 + 1, h = n / 2;
        const BIdx b = bidx(e, ext);
        auto src = [&](int m) { return m < h ? m : (m > ext - h ? m - ext + n : -1); };
        const int sx = b.kx < h ? b.kx : -1, sy = src(b.ky), sz = src(b.kz);
        const double dk = 2 * M_PI / L;
        cx<double> v = {0, 0};
        if (sx >= 0 && sy >= 0 && sz >= 0) {
            const BIdx bn = {sx, sy, sz};
            const double g = kdir(bn, i, n, dk);
            const cx<double> x = F[(size_t)sz + (size_t)n * (sx + (size_t)Hn * sy)];
            cx<double> r = {-g * x.y, g * x.x};
            r = {-r.x, -r.y};
            v = scale(r, s);
        }
        const double g2 = kdir(b, k, ext, dk);
        return {-g2 * v.y, g2 * v.x};
    }
};

// The transpose for the r2c of an A_ik cotangent, applied to entry e (value x, ext grid) as the z pass
// stores it: half-spectrum weight times sme, the PSI1 multiplier (ext grid, axis k), the crop (scale s),
// the DERIV multiplier (n grid, axis i), added to the n-grid entry of Fb it crops to (each n-grid entry
// has one ext entry; ext entries the crop drops are not used)
struct DerivAdjDst {
    cx<double> *Fb;
    int n, ext, i, k;
    double L, sme, s;
    __device__ void operator()(size_t e, cx<double> x0) const {
        const int Hn = n / 2 + 1, h = n / 2;
        const BIdx be = bidx(e, ext);
        auto dst = [&](int q) { return q < h ? q : (q > n ? q - ext + n : -1); };
        const int nx = be.kx < h ? be.kx : -1, ny = dst(be.ky), nz = dst(be.kz);
        if (nx < 0 || ny < 0 || nz < 0) return;
        const double dk = 2 * M_PI / L;
        const double w = (be.kx == 0 || be.kx == ext / 2) ? 1.0 : 2.0;
        const cx<double> x = scale(x0, sme * w);
        const double g = kdir(be, k, ext, dk);
        cx<double> r = {-g * x.y, g * x.x};
        r = {-r.x, -r.y};
        const cx<double> v = scale(r, s);
        const BIdx bn = {nx, ny, nz};
        const double g2 = kdir(bn, i, n, dk);
        const cx<double> t = {-g2 * v.y, g2 * v.x};
        const size_t en = (size_t)nz + (size_t)n * (nx + (size_t)Hn * ny);
        Fb[en] = {Fb[en].x + t.x, Fb[en].y + t.y};
    }
};

}  // namespace dj
