// Complex arithmetic and small in-register DFTs shared by the FFT kernels.
// Compiles for the device (gfx950) and, for the CPU emulation of the kernels' index logic, for
// the host.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DJ_HD __host__ __device__ __forceinline__
#else
#define DJ_HD inline
#endif

namespace dj {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <typename T> DJ_HD cx<T> mk(T a, T b) { return cx<T>{a, b}; }

// Accesses of the transforms' contiguous (read-once, write-once) sides, with non-temporal hints in
// the HIP build so that these lines do not displace the transposed chunks that neighbouring blocks
// still have to merge in L2; plain accesses in the host emulation.
template <typename T> DJ_HD cx<T> ld_stream(const cx<T> *p) {
#if defined(__HIPCC__)
    typedef T V __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load((const V *)p);
    return cx<T>{v.x, v.y};
#else
    return *p;
#endif
}
template <typename T> DJ_HD void st_stream(cx<T> *p, cx<T> a) {
#if defined(__HIPCC__)
    typedef T V __attribute__((ext_vector_type(2)));
    V v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, (V *)p);
#else
    *p = a;
#endif
}
template <typename T> DJ_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> DJ_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> DJ_HD cx<T> scale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T> DJ_HD cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// a * b with fused multiply-adds (twiddle products inside the transforms)
template <typename T> DJ_HD cx<T> cmul(cx<T> a, cx<T> b) {
    return {std::fma(a.x, b.x, -(a.y * b.y)), std::fma(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
template <typename T> DJ_HD cx<T> cmulc(cx<T> a, cx<T> b) {
    return {std::fma(a.x, b.x, a.y * b.y), std::fma(a.y, b.x, -(a.x * b.y))};
}
// multiply by DIR * i   (DIR = -1: -i, DIR = +1: +i)
template <int DIR, typename T> DJ_HD cx<T> mul_i(cx<T> a) {
    return DIR < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
}

// exp(DIR * 2 pi i m / 16) for m = 0..9 (cos, sin of multiples of pi/8)
template <typename T, int DIR> DJ_HD cx<T> w16(int m) {
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    const double cs[10][2] = {{1, 0}, {c1, s1}, {h, h}, {s1, c1}, {0, 1},
                              {-s1, c1}, {-h, h}, {-c1, s1}, {-1, 0}, {-c1, -s1}};
    return {(T)cs[m][0], (T)(DIR * cs[m][1])};
}

template <int DIR, typename T> DJ_HD void dft2(cx<T> *a) {
    cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
}

template <int DIR, typename T> DJ_HD void dft3(cx<T> *a) {
    const T c = (T)-0.5, s = (T)(DIR * 0.86602540378443864676);
    cx<T> t1 = a[1] + a[2], t2 = a[1] - a[2];
    cx<T> m = {std::fma(c, t1.x, a[0].x), std::fma(c, t1.y, a[0].y)};
    cx<T> r = {-s * t2.y, s * t2.x};  // i * s * (a1 - a2)
    a[0] = a[0] + t1;
    a[1] = m + r;
    a[2] = m - r;
}

template <int DIR, typename T> DJ_HD void dft4(cx<T> *a) {
    cx<T> t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = mul_i<DIR>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
}

// 8 = 4 x 2: n = 2 n1 + n2, k = k1 + 4 k2
template <int DIR, typename T> DJ_HD void dft8(cx<T> *a) {
    cx<T> y[2][4];
#pragma unroll
    for (int n2 = 0; n2 < 2; n2++) {
#pragma unroll
        for (int n1 = 0; n1 < 4; n1++) y[n2][n1] = a[2 * n1 + n2];
        dft4<DIR>(y[n2]);
    }
#pragma unroll
    for (int k1 = 1; k1 < 4; k1++) y[1][k1] = cmul(y[1][k1], w16<T, DIR>(2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        a[k1] = y[0][k1] + y[1][k1];
        a[k1 + 4] = y[0][k1] - y[1][k1];
    }
}

// 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
template <int DIR, typename T> DJ_HD void dft16(cx<T> *a) {
    cx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
#pragma unroll
        for (int n1 = 0; n1 < 4; n1++) y[n2][n1] = a[4 * n1 + n2];
        dft4<DIR>(y[n2]);
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; n2++)
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++) y[n2][k1] = cmul(y[n2][k1], w16<T, DIR>(n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        cx<T> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<DIR>(z);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) a[k1 + 4 * k2] = z[k2];
    }
}

template <int R, int DIR, typename T> DJ_HD void dft(cx<T> *a) {
    if constexpr (R == 2) dft2<DIR>(a);
    else if constexpr (R == 3) dft3<DIR>(a);
    else if constexpr (R == 4) dft4<DIR>(a);
    else if constexpr (R == 8) dft8<DIR>(a);
    else if constexpr (R == 16) dft16<DIR>(a);
    else static_assert(R == 2, "unsupported radix");
}

}  // namespace dj
