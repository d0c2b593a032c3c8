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
// the same for a grid size known at compile time (every index of these grids is below 2^32)
template <int N> __device__ __forceinline__ BIdx bidx_c(size_t e) {
    constexpr unsigned H = N / 2 + 1;
    const unsigned u = (unsigned)e, t = u / (unsigned)N;
    BIdx r;
    r.kz = (int)(u - t * (unsigned)N);
    r.ky = (int)(t / H);
    r.kx = (int)(t - (unsigned)r.ky * H);
    return r;
}

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
__device__ __forceinline__ cx<double> ic_mult_b(const BIdx &b, bool dc, cx<double> v, int N, double L, int c) {
    const double dk = 2 * M_PI / L;
    const double g = kdir(b, c, N, dk);
    cx<double> r = {-g * v.y, g * v.x};  // (i g) v
    if constexpr (MODE == IC_PSI1) r = {-r.x, -r.y};
    if constexpr (MODE == IC_PSI2 || MODE == IC_PSI2_ADJ) {
        const double kx = b.kx * dk, ky = sgn_full(b.ky, N) * dk, kz = sgn_full(b.kz, N) * dk;
        const double k2 = (kx * kx + ky * ky) + kz * kz;
        const double il = -1.0 / (dc ? 1.0 : k2);
        r = scale(r, MODE == IC_PSI2 ? il : -il);
    }
    return r;
}
template <int MODE>
__device__ __forceinline__ cx<double> ic_mult_at(size_t e, cx<double> v, int N, double L, int c) {
    return ic_mult_b<MODE>(bidx(e, N), e == 0, v, N, L, c);
}
// the multiplier applied to in[e] as a z pass loads it (grid N)
template <int MODE, int N> struct MultSrc {
    const cx<double> *in;
    int c;
    double L;
    __device__ cx<double> operator()(size_t e) const { return ic_mult_b<MODE>(bidx_c<N>(e), e == 0, in[e], N, L, c); }
};
// a z pass's output v at e: the half-spectrum weight times s (1 at kx = 0 and N/2, else 2), the multiplier,
// then added into acc[e]
template <int MODE, int N> struct MultAccDst {
    cx<double> *acc;
    int c;
    double L, s;
    __device__ void operator()(size_t e, cx<double> v) const {
        const BIdx b = bidx_c<N>(e);
        const double w = (b.kx == 0 || b.kx == N / 2) ? 1.0 : 2.0;
        const cx<double> r = ic_mult_b<MODE>(b, e == 0, scale(v, s * w), N, L, c);
        acc[e] = {acc[e].x + r.x, acc[e].y + r.y};
    }
};

// Entry e of one A_ik spectrum on the ext grid from phi (F, n grid): the steps of the PSI1 multiplier
// (n grid, axis i), the pad (scale s) and the DERIV multiplier (ext grid, axis k), in their order
template <int EXT> struct DerivPadSrc {
    static constexpr int ext = EXT, n = 2 * EXT / 3;
    const cx<double> *F;
    int i, k;
    double L, s;
    __device__ cx<double> operator()(size_t e) const {
        constexpr int Hn = n / 2 + 1, h = n / 2;
        const BIdx b = bidx_c<EXT>(e);
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
template <int EXT> struct DerivAdjDst {
    static constexpr int ext = EXT, n = 2 * EXT / 3;
    cx<double> *Fb;
    int i, k;
    double L, sme, s;
    __device__ void operator()(size_t e, cx<double> x0) const {
        constexpr int Hn = n / 2 + 1, h = n / 2;
        const BIdx be = bidx_c<EXT>(e);
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
