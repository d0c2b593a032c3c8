// Sheet resampling (n_resample > 1): n_resample^3 - 1 extra point sets per interlacing pass, each the
// particle sheet displaced by a Lagrangian offset d (grid cells), deposited into the same density
// (NR/src/nbody/pm.jl:189-234 and 430-464).
//
//   psi_c = mod(Xs_c - q_c + L/2, L) - L/2                     (wrap, per component)
//   Fourier: npsi_c = brfft(rfft(psi_c) * px py pz / n^3)      (p_a = exp(2 pi i m_a d_a / n))
//   linear:  npsi_c = periodic trilinear interpolation of psi_c at (i + dx, j + dy, k + dz)
//   Xr_c = mod(q_c + npsi_c + d_c L / n + L, L), deposited like the particles
//
// The reference halves the slowest axis (z) of its real transforms, and once the phase factors break
// the Hermitian symmetry of the half spectrum the inverse depends on that choice.  The library's FFT
// passes halve the contiguous axis, so the Fourier path works on the transposed field psiT[z + n (y + n x)]:
// its halved axis is the reference's z, with the half-axis frequencies m / n (Nyquist +1/2), and the
// inverse is then the reference's.
#include <hip/hip_runtime.h>

#include "dj_complex.h"
#include "dj_internal.h"

namespace dj {

template <typename T> __device__ __forceinline__ T rs_jmod(T x, T L) {
    T r = fmod(x, L);
    if (r == (T)0) return copysign(r, L);
    if ((r > (T)0) != (L > (T)0)) return r + L;
    return r;
}

// psi_c of particle (i, j, k) at X_s = (psi + q) + s_c, into the natural or the transposed layout
template <typename T, bool TRANSPOSED>
__global__ __launch_bounds__(256) void k_rs_wrap(const T *__restrict__ psi, const T *__restrict__ q1, int n, int c, T sc,
                                                 T L, T *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, p = (size_t)i + N * ((size_t)j + N * k);
    const T qc = q1[c == 0 ? i : c == 1 ? j : k];
    const T xs = (psi[p + (size_t)c * N * N * N] + qc) + sc;
    const T v = rs_jmod(xs - qc + L / (T)2, L) - L / (T)2;
    out[TRANSPOSED ? (size_t)k + N * ((size_t)j + N * i) : p] = v;
}

// the phase of offset d on the transposed field's half spectrum, library layout B[w + n (h + H v)]:
// w = the reference's x (full axis), h = its z (half axis), v = its y (full axis)
template <typename T>
__global__ __launch_bounds__(256) void k_rs_phase(const cx<T> *__restrict__ in, cx<T> *__restrict__ out, int n,
                                                  const cx<T> *__restrict__ px, const cx<T> *__restrict__ py,
                                                  const cx<T> *__restrict__ pz, T scale) {
    const int H = n / 2 + 1;
    const size_t tot = (size_t)n * H * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(e % n);
        const size_t r = e / n;
        const int h = (int)(r % H), v = (int)(r / H);
        // psi_hat * (px py pz scale), the factor built left to right as the reference does
        cx<T> f = cmul(cmul(px[w], py[v]), pz[h]);
        f = {f.x * scale, f.y * scale};
        out[e] = cmul(in[e], f);
    }
}

// periodic trilinear interpolation (pm.jl:220-231, the same sum order)
template <typename T>
__global__ __launch_bounds__(256) void k_rs_trilinear(const T *__restrict__ f, int n, T dx, T dy, T dz, T *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n;
    const int i1 = i + 1 == n ? 0 : i + 1, j1 = j + 1 == n ? 0 : j + 1, k1 = k + 1 == n ? 0 : k + 1;
    auto F = [&](int a, int b, int cc) { return f[(size_t)a + N * ((size_t)b + N * cc)]; };
    const T ex = (T)1 - dx, ey = (T)1 - dy, ez = (T)1 - dz;
    out[(size_t)i + N * ((size_t)j + N * k)] =
        F(i, j, k) * ex * ey * ez + F(i1, j, k) * dx * ey * ez + F(i, j1, k) * ex * dy * ez + F(i, j, k1) * ex * ey * dz +
        F(i1, j1, k) * dx * dy * ez + F(i1, j, k1) * dx * ey * dz + F(i, j1, k1) * ex * dy * dz +
        F(i1, j1, k1) * dx * dy * dz;
}

// Xr_c = mod(q_c + npsi_c + shift + L, L); npsi in the natural or the transposed layout.  Written at
// point (nr i + o0, nr j + o1, nr k + o2) of the (nr n)^3 lattice of all point sets of a pass (offset
// (o0, o1, o2) / nr cells: the sets interleave into a finer lattice; nr = 1, o = 0 is the plain layout)
template <typename T, bool TRANSPOSED>
__global__ __launch_bounds__(256) void k_rs_place(const T *__restrict__ npsi, const T *__restrict__ q1, int n, int c,
                                                  T shift, T L, T *__restrict__ xr, int nr, int o0, int o1, int o2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, p = (size_t)i + N * ((size_t)j + N * k), NR = (size_t)nr * n;
    const T qc = q1[c == 0 ? i : c == 1 ? j : k];
    const T np_ = npsi[TRANSPOSED ? (size_t)k + N * ((size_t)j + N * i) : p];
    const size_t pr = ((size_t)nr * i + o0) + NR * (((size_t)nr * j + o1) + NR * ((size_t)nr * k + o2));
    xr[pr + (size_t)c * NR * NR * NR] = rs_jmod(((qc + np_) + shift) + L, L);
}

// the particles' own positions of the pass, X_c + s_c = (psi_c + q_c) + s_c (the deposit's arithmetic), at
// points (nr i, nr j, nr k) of the point-set lattice
template <typename T>
__global__ __launch_bounds__(256) void k_rs_shifted(const T *__restrict__ psi, const T *__restrict__ q1, int n, T s0,
                                                    T s1, T s2, int nr, T *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, p = (size_t)i + N * ((size_t)j + N * k), NR = (size_t)nr * n, NR3 = NR * NR * NR;
    const size_t pr = (size_t)nr * i + NR * ((size_t)nr * j + NR * ((size_t)nr * k));
    out[pr] = (psi[p] + q1[i]) + s0;
    out[pr + NR3] = (psi[p + N * N * N] + q1[j]) + s1;
    out[pr + 2 * NR3] = (psi[p + 2 * N * N * N] + q1[k]) + s2;
}

// 1-D assignment weights and base cell (pm.jl:49-55), as in the particle kernels
template <int W, typename T> __device__ __forceinline__ T rs_wk(T d) {
    if constexpr (W == 2) {
        const T a = (T)1 - d;
        return d > (T)1 ? (T)0 : a;
    } else if constexpr (W == 3) {
        const T t = (T)1.5 - d;
        const T outer = (T)0.5 * (t * t), inner = (T)0.75 - d * d;
        return d > (T)1.5 ? (T)0 : (d > (T)0.5 ? outer : inner);
    } else {
        const T t = (T)2 - d;
        const T outer = t * t * t / (T)6;
        const T inner = ((T)4 - (T)6 * (d * d) + (T)3 * (d * d * d)) / (T)6;
        return d > (T)2 ? (T)0 : (d > (T)1 ? outer : inner);
    }
}

// unit-mass deposit of the positions X (3, n^3) with device atomics, added to a mesh that the tile
// deposit has already written (float atomics, or int64 fixed point when FIXED)
template <typename T, int W, bool FIXED>
__global__ __launch_bounds__(256) void k_rs_scatter(const T *__restrict__ X, int n, void *mesh, int res, T L,
                                                    double fscale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, p = (size_t)i + N * ((size_t)j + N * k), np = N * N * N;
    T w[3][W];
    int cell[3][W];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T g = rs_jmod(X[p + a * np], L) / L * (T)res;
        const int b = (W % 2 == 0) ? (int)floor(g) : (int)rint(g);
        const int lo = W == 2 ? 0 : -1;
#pragma unroll
        for (int o = 0; o < W; o++) {
            const int bo = b + lo + o;
            w[a][o] = rs_wk<W>(fabs(g - (T)bo));
            int cc = bo % res;
            cell[a][o] = cc < 0 ? cc + res : cc;
        }
    }
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++) {
            const size_t row = (size_t)res * ((size_t)cell[1][oy] + (size_t)res * cell[2][oz]);
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const T wv = w[0][ox] * w[1][oy] * w[2][oz];
                const size_t cidx = row + cell[0][ox];
                if constexpr (FIXED) {
                    atomicAdd((unsigned long long *)mesh + cidx, (unsigned long long)__double2ll_rn((double)wv * fscale));
                } else {
                    unsafeAtomicAdd((T *)mesh + cidx, wv);
                }
            }
        }
}

// ------------------------------------------------------------------------ reverse mode
// The resampled points enter the density only: the cotangent of their positions (the gather of grad W
// times the density cotangent at each point, on the point lattice) goes back to Psi through the
// transposed interpolation.  d Xr / d npsi and d psi_c / d Psi_c are the identity (the mods are locally
// translations).  Linear: the transpose of the trilinear shift scatters a point's cotangent to its 8
// corners, written here as a gather over the 8 points that use a lattice site.  Fourier: npsi = c2r(P s
// r2c(psi)) has the transpose c2r(conj(P) s r2c(npsibar)) (the half-axis weights of c2r and r2c cancel).

// the cotangent of point set (o0, o1, o2), component c, from the point lattice into the natural or the
// transposed n^3 layout
template <typename T, bool TRANSPOSED>
__global__ __launch_bounds__(256) void k_rs_extract(const T *__restrict__ pl, int n, int nr, int o0, int o1, int o2,
                                                    int c, T *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, NR = (size_t)nr * n;
    const size_t pr = ((size_t)nr * i + o0) + NR * (((size_t)nr * j + o1) + NR * ((size_t)nr * k + o2));
    out[TRANSPOSED ? (size_t)k + N * ((size_t)j + N * i) : (size_t)i + N * ((size_t)j + N * k)] =
        pl[pr + (size_t)c * NR * NR * NR];
}

// hb (+)= in * conj((px py) pz scale): the forward's phase factor, conjugated
template <typename T, bool ACC>
__global__ __launch_bounds__(256) void k_rs_phase_adj(const cx<T> *__restrict__ in, cx<T> *__restrict__ hb, int n,
                                                      const cx<T> *__restrict__ px, const cx<T> *__restrict__ py,
                                                      const cx<T> *__restrict__ pz, T scale) {
    const int H = n / 2 + 1;
    const size_t tot = (size_t)n * H * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(e % n);
        const size_t r = e / n;
        const int h = (int)(r % H), v = (int)(r / H);
        cx<T> f = cmul(cmul(px[w], py[v]), pz[h]);
        f = {f.x * scale, -(f.y * scale)};
        const cx<T> x = cmul(in[e], f);
        hb[e] = ACC ? hb[e] + x : x;
    }
}

// psibar_c (natural layout) += v (transposed layout)
template <typename T>
__global__ __launch_bounds__(256) void k_rs_add_transposed(const T *__restrict__ v, int n, T *__restrict__ psibar_c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n;
    psibar_c[(size_t)i + N * ((size_t)j + N * k)] += v[(size_t)k + N * ((size_t)j + N * i)];
}

// psibar_c(i, j, k) += sum_{a, b, e in {0, 1}} wx(a) wy(b) wz(e) pbar(i - a, j - b, k - e), the transpose of
// k_rs_trilinear (w(0) = 1 - d, w(1) = d per axis), with pbar the point set's cotangent on the lattice
template <typename T>
__global__ __launch_bounds__(256) void k_rs_trilinear_adj(const T *__restrict__ pl, int n, int nr, int o0, int o1, int o2,
                                                          int c, T dx, T dy, T dz, T *__restrict__ psibar_c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= n) return;
    const size_t N = (size_t)n, NR = (size_t)nr * n;
    const T *pc = pl + (size_t)c * NR * NR * NR;
    const int im = i == 0 ? n - 1 : i - 1, jm = j == 0 ? n - 1 : j - 1, km = k == 0 ? n - 1 : k - 1;
    auto P = [&](int a, int b, int e) {
        return pc[((size_t)nr * a + o0) + NR * (((size_t)nr * b + o1) + NR * ((size_t)nr * e + o2))];
    };
    const T ex = (T)1 - dx, ey = (T)1 - dy, ez = (T)1 - dz;
    const T v = P(i, j, k) * ex * ey * ez + P(im, j, k) * dx * ey * ez + P(i, jm, k) * ex * dy * ez +
                P(i, j, km) * ex * ey * dz + P(im, jm, k) * dx * dy * ez + P(im, j, km) * dx * ey * dz +
                P(i, jm, km) * ex * dy * dz + P(im, jm, km) * dx * dy * dz;
    psibar_c[(size_t)i + N * ((size_t)j + N * k)] += v;
}

template <typename T>
struct RsLaunch {
    static dim3 pgrid(int n) { return dim3((n + 255) / 256, n, n); }
    static void wrap(const void *psi, const void *q1, int n, int c, double sc, double L, bool transposed, void *out,
                     hipStream_t st) {
        if (transposed)
            k_rs_wrap<T, true><<<pgrid(n), 256, 0, st>>>((const T *)psi, (const T *)q1, n, c, (T)sc, (T)L, (T *)out);
        else
            k_rs_wrap<T, false><<<pgrid(n), 256, 0, st>>>((const T *)psi, (const T *)q1, n, c, (T)sc, (T)L, (T *)out);
    }
    static void phase(const void *in, void *out, int n, const void *px, const void *py, const void *pz, double scale,
                      hipStream_t st) {
        k_rs_phase<T><<<2048, 256, 0, st>>>((const cx<T> *)in, (cx<T> *)out, n, (const cx<T> *)px, (const cx<T> *)py,
                                            (const cx<T> *)pz, (T)scale);
    }
    static void trilinear(const void *f, int n, const double *d, void *out, hipStream_t st) {
        k_rs_trilinear<T><<<pgrid(n), 256, 0, st>>>((const T *)f, n, (T)d[0], (T)d[1], (T)d[2], (T *)out);
    }
    static void place(const void *npsi, const void *q1, int n, int c, double shift, double L, bool transposed, void *xr,
                      int nr, const int *o, hipStream_t st) {
        if (transposed)
            k_rs_place<T, true><<<pgrid(n), 256, 0, st>>>((const T *)npsi, (const T *)q1, n, c, (T)shift, (T)L, (T *)xr,
                                                          nr, o[0], o[1], o[2]);
        else
            k_rs_place<T, false><<<pgrid(n), 256, 0, st>>>((const T *)npsi, (const T *)q1, n, c, (T)shift, (T)L, (T *)xr,
                                                           nr, o[0], o[1], o[2]);
    }
    static void extract(const void *pl, int n, int nr, const int *o, int c, bool transposed, void *out,
                        hipStream_t st) {
        if (transposed)
            k_rs_extract<T, true><<<pgrid(n), 256, 0, st>>>((const T *)pl, n, nr, o[0], o[1], o[2], c, (T *)out);
        else
            k_rs_extract<T, false><<<pgrid(n), 256, 0, st>>>((const T *)pl, n, nr, o[0], o[1], o[2], c, (T *)out);
    }
    static void phase_adj(const void *in, void *hb, int n, const void *px, const void *py, const void *pz, double scale,
                          bool acc, hipStream_t st) {
        if (acc)
            k_rs_phase_adj<T, true><<<2048, 256, 0, st>>>((const cx<T> *)in, (cx<T> *)hb, n, (const cx<T> *)px,
                                                          (const cx<T> *)py, (const cx<T> *)pz, (T)scale);
        else
            k_rs_phase_adj<T, false><<<2048, 256, 0, st>>>((const cx<T> *)in, (cx<T> *)hb, n, (const cx<T> *)px,
                                                           (const cx<T> *)py, (const cx<T> *)pz, (T)scale);
    }
    static void add_transposed(const void *v, int n, void *psibar_c, hipStream_t st) {
        k_rs_add_transposed<T><<<pgrid(n), 256, 0, st>>>((const T *)v, n, (T *)psibar_c);
    }
    static void trilinear_adj(const void *pl, int n, int nr, const int *o, int c, const double *d, void *psibar_c,
                              hipStream_t st) {
        k_rs_trilinear_adj<T><<<pgrid(n), 256, 0, st>>>((const T *)pl, n, nr, o[0], o[1], o[2], c, (T)d[0], (T)d[1],
                                                        (T)d[2], (T *)psibar_c);
    }
    static void shifted(const void *psi, const void *q1, int n, const double *sh, int nr, void *out, hipStream_t st) {
        k_rs_shifted<T><<<pgrid(n), 256, 0, st>>>((const T *)psi, (const T *)q1, n, (T)sh[0], (T)sh[1], (T)sh[2], nr,
                                                  (T *)out);
    }
    template <int W>
    static void scatter_w(const void *X, int n, void *mesh, int res, double L, bool fixed, double fscale, hipStream_t st) {
        if (fixed) k_rs_scatter<T, W, true><<<pgrid(n), 256, 0, st>>>((const T *)X, n, mesh, res, (T)L, fscale);
        else k_rs_scatter<T, W, false><<<pgrid(n), 256, 0, st>>>((const T *)X, n, mesh, res, (T)L, fscale);
    }
    static void scatter(const void *X, int n, int W, void *mesh, int res, double L, bool fixed, double fscale,
                        hipStream_t st) {
        if (W == 2) scatter_w<2>(X, n, mesh, res, L, fixed, fscale, st);
        else if (W == 3) scatter_w<3>(X, n, mesh, res, L, fixed, fscale, st);
        else scatter_w<4>(X, n, mesh, res, L, fixed, fscale, st);
    }
};

template <typename T> void resample_ops(ResampleOps &ops) {
    ops.extract = RsLaunch<T>::extract;
    ops.phase_adj = RsLaunch<T>::phase_adj;
    ops.add_transposed = RsLaunch<T>::add_transposed;
    ops.trilinear_adj = RsLaunch<T>::trilinear_adj;
    ops.wrap = RsLaunch<T>::wrap;
    ops.phase = RsLaunch<T>::phase;
    ops.trilinear = RsLaunch<T>::trilinear;
    ops.place = RsLaunch<T>::place;
    ops.shifted = RsLaunch<T>::shifted;
    ops.scatter = RsLaunch<T>::scatter;
}

template void resample_ops<float>(ResampleOps &);
template void resample_ops<double>(ResampleOps &);

}  // namespace dj
