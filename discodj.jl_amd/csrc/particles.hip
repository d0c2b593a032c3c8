// Particle-side kernels of the PM step: mass assignment, force gather, kick and drift.
//
// Particles are the reference's column-major (n,n,n,3) arrays: particle p = i + n (j + n k) has
// Lagrangian position (q[i], q[j], q[k]) with q[i] = i L/n (FP64, rounded to T), and its
// component c at p + c n^3.  X = (psi + q) + shift is formed in T exactly as the reference forms
// X = psi + q and then Xs = X + T(s) (NR/src/nbody/pm.jl:512, 516).
//
// Cell assignment and weights restate NR/src/nbody/pm.jl:49-74 operation for operation (the
// library is built with -ffp-contract=off): g = (mod(x, L) / L) * res with Julia's float mod,
// base = floor(g) for even orders and round-half-even(g) for odd orders, weights of |g - cell|.
// The FP32 cell indices and weights therefore equal those of the reference's FP32 path bit for bit.
#include <hip/hip_runtime.h>

#include "dj_internal.h"

namespace dj {

template <typename T> __device__ __forceinline__ T jmod(T x, T L) {
    T r = fmod(x, L);
    if (r == (T)0) return copysign(r, L);
    if ((r > (T)0) != (L > (T)0)) return r + L;
    return r;
}

template <int W, typename T> __device__ __forceinline__ T wk(T d) {
    if constexpr (W == 2) {
        return d > (T)1 ? (T)0 : (T)1 - d;
    } else if constexpr (W == 3) {
        if (d > (T)1.5) return (T)0;
        if (d > (T)0.5) { T t = (T)1.5 - d; return (T)0.5 * (t * t); }
        return (T)0.75 - d * d;
    } else {
        if (d > (T)2) return (T)0;
        if (d > (T)1) { T t = (T)2 - d; return t * t * t / (T)6; }
        return ((T)4 - (T)6 * (d * d) + (T)3 * (d * d * d)) / (T)6;
    }
}

template <int W> __device__ __forceinline__ constexpr int wlo() { return W == 2 ? 0 : -1; }

template <int W, typename T> __device__ __forceinline__ int wbase(T g) {
    return (W % 2 == 0) ? (int)floor(g) : (int)rint(g);
}

__device__ __forceinline__ int imod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

// per-axis stencil: cells and weights of the W points, plus the FD neighbourhood (3 each side)
template <typename T, int W, int HALO>
struct Axis {
    T w[W];
    int nb[W + 2 * HALO];  // nb[t] = mod(base + lo - HALO + t, res)
    __device__ __forceinline__ void make(T x, int res, T L) {
        T g = jmod(x, L) / L * (T)res;
        int b = wbase<W>(g);
#pragma unroll
        for (int o = 0; o < W; o++) w[o] = wk<W>(fabs(g - (T)(b + wlo<W>() + o)));
        int c0 = imod(b + wlo<W>() - HALO, res);
#pragma unroll
        for (int t = 0; t < W + 2 * HALO; t++) {
            int c = c0 + t;
            nb[t] = c >= res ? c - res : c;
        }
    }
};

struct PIdx {
    int i, j, k;
    size_t p, np;
    bool ok;
};
__device__ __forceinline__ PIdx pidx(int n) {
    PIdx r;
    r.i = blockIdx.x * blockDim.x + threadIdx.x;
    r.j = blockIdx.y;
    r.k = blockIdx.z;
    r.ok = r.i < n;
    r.np = (size_t)n * n * n;
    r.p = (size_t)r.i + (size_t)n * ((size_t)r.j + (size_t)n * r.k);
    return r;
}

template <typename T, int W, bool FIXED>
__global__ __launch_bounds__(256) void k_scatter(const T *__restrict__ psi, const T *__restrict__ q1, int n,
                                                 T sx, T sy, T sz, void *mesh, int res, T L, double fscale) {
    PIdx P = pidx(n);
    if (!P.ok) return;
    T x = (psi[P.p] + q1[P.i]) + sx;
    T y = (psi[P.p + P.np] + q1[P.j]) + sy;
    T z = (psi[P.p + 2 * P.np] + q1[P.k]) + sz;
    Axis<T, W, 0> ax, ay, az;
    ax.make(x, res, L);
    ay.make(y, res, L);
    az.make(z, res, L);
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++) {
            size_t row = (size_t)res * ((size_t)ay.nb[oy] + (size_t)res * az.nb[oz]);
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                T w = ax.w[ox] * ay.w[oy] * az.w[oz];
                size_t c = row + ax.nb[ox];
                if constexpr (FIXED) {
                    long long v = __double2ll_rn((double)w * fscale);
                    atomicAdd((unsigned long long *)mesh + c, (unsigned long long)v);
                } else {
                    unsafeAtomicAdd((T *)mesh + c, w);
                }
            }
        }
}

// -D_a phi at a cell, order 2/4/6 central differences (pm.jl:330-342); nb holds the cell's
// neighbourhood along the axis with the cell at index t0
template <typename T, int ORDER>
__device__ __forceinline__ T fdiff(const T *__restrict__ phi, size_t base, size_t stride, const int *nb, int t0,
                                   T c1, T c2, T c3) {
    T d1 = phi[base + stride * nb[t0 + 1]] - phi[base + stride * nb[t0 - 1]];
    if constexpr (ORDER == 2) return -(c1 * d1);
    T d2 = phi[base + stride * nb[t0 + 2]] - phi[base + stride * nb[t0 - 2]];
    if constexpr (ORDER == 4) return -(c1 * d1 + c2 * d2);
    T d3 = phi[base + stride * nb[t0 + 3]] - phi[base + stride * nb[t0 - 3]];
    return -(c1 * d1 + c2 * d2 + c3 * d3);
}

template <typename T, int W, int ORDER, int MODE>
__global__ __launch_bounds__(256) void k_gather_fd(const T *__restrict__ psi_in, T *__restrict__ psi_out,
                                                   T *__restrict__ mom, T *__restrict__ acc,
                                                   const T *__restrict__ q1, int n, T sx, T sy, T sz,
                                                   const T *__restrict__ phi, int res, T L, T c1, T c2, T c3,
                                                   T alpha, T beta, T dd2, T dd1n, int has_next, int ns) {
    PIdx P = pidx(n);
    if (!P.ok) return;
    T px = psi_in[P.p], py = psi_in[P.p + P.np], pz = psi_in[P.p + 2 * P.np];
    T x = (px + q1[P.i]) + sx;
    T y = (py + q1[P.j]) + sy;
    T z = (pz + q1[P.k]) + sz;
    constexpr int HALO = ORDER / 2;
    Axis<T, W, HALO> ax, ay, az;
    ax.make(x, res, L);
    ay.make(y, res, L);
    az.make(z, res, L);
    const size_t R = (size_t)res, RR = R * R;
    T s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                T w = ax.w[ox] * ay.w[oy] * az.w[oz];
                const int tx = ox + HALO, ty = oy + HALO, tz = oz + HALO;
                size_t cx = ax.nb[tx], cy = ay.nb[ty], cz = az.nb[tz];
                T fx = fdiff<T, ORDER>(phi, R * cy + RR * cz, 1, ax.nb, tx, c1, c2, c3);
                T fy = fdiff<T, ORDER>(phi, cx + RR * cz, R, ay.nb, ty, c1, c2, c3);
                T fz = fdiff<T, ORDER>(phi, cx + R * cy, RR, az.nb, tz, c1, c2, c3);
                s1 += fx * w;
                s2 += fy * w;
                s3 += fz * w;
            }
    if constexpr (MODE == GATHER_ACC) {
        acc[P.p] += s1;
        acc[P.p + P.np] += s2;
        acc[P.p + 2 * P.np] += s3;
    } else if constexpr (MODE == GATHER_ASSIGN) {
        acc[P.p] = s1;
        acc[P.p + P.np] = s2;
        acc[P.p + 2 * P.np] = s3;
    } else {
        T a[3] = {s1, s2, s3}, ps[3] = {px, py, pz};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            size_t q = P.p + c * P.np;
            T ac = a[c];
            if (ns > 1) ac = (acc[q] + ac) / (T)ns;
            T m = alpha * mom[q] + beta * ac;
            T s = ps[c] + dd2 * m;
            if (has_next) s = s + dd1n * m;
            mom[q] = m;
            psi_out[q] = s;
        }
    }
}

template <typename T, int W>
__global__ __launch_bounds__(256) void k_gather_comp(const T *__restrict__ psi, T *__restrict__ acc,
                                                     const T *__restrict__ q1, int n, T sx, T sy, T sz,
                                                     const T *__restrict__ f, int res, T L, int comp) {
    PIdx P = pidx(n);
    if (!P.ok) return;
    T x = (psi[P.p] + q1[P.i]) + sx;
    T y = (psi[P.p + P.np] + q1[P.j]) + sy;
    T z = (psi[P.p + 2 * P.np] + q1[P.k]) + sz;
    Axis<T, W, 0> ax, ay, az;
    ax.make(x, res, L);
    ay.make(y, res, L);
    az.make(z, res, L);
    const size_t R = (size_t)res;
    T s = 0;
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                T w = ax.w[ox] * ay.w[oy] * az.w[oz];
                s += f[ax.nb[ox] + R * (ay.nb[oy] + R * az.nb[oz])] * w;
            }
    acc[P.p + comp * P.np] += s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_drift(T *__restrict__ psi, const T *__restrict__ mom, int64_t n3, T dd) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) psi[i] = psi[i] + dd * mom[i];
}

// out = psi + dd mom (out may be psi), the same arithmetic as k_drift; with CK, psi and mom are also
// copied to ckp and ckm as they are read (the checkpoint of a step, without copy passes of its own)
template <typename T, bool CK>
__global__ __launch_bounds__(256) void k_drift_to(T *out, const T *psi, const T *__restrict__ mom, T *__restrict__ ckp,
                                                  T *__restrict__ ckm, int64_t n3, T dd) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const T p = psi[i], m = mom[i];
    if constexpr (CK) {
        ckp[i] = p;
        ckm[i] = m;
    }
    out[i] = p + dd * m;
}

template <typename T>
__global__ __launch_bounds__(256) void k_kick_drift(T *__restrict__ psi, T *__restrict__ mom,
                                                    const T *__restrict__ acc, int64_t n3, T alpha, T beta,
                                                    T dd2, T dd1n, int has_next, int ns) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    T a = acc[i];
    if (ns > 1) a = a / (T)ns;
    T m = alpha * mom[i] + beta * a;
    T s = psi[i] + dd2 * m;
    if (has_next) s = s + dd1n * m;
    mom[i] = m;
    psi[i] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale(T *__restrict__ x, int64_t n, T s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] * s;
}

__global__ __launch_bounds__(256) void k_zero(uint4 *__restrict__ x, size_t n16, uint8_t *tail, size_t ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) x[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// ------------------------------------------------------------------------------------ launchers
static dim3 pgrid(int n) { return dim3((n + 255) / 256, n, n); }
static unsigned lgrid(int64_t n) { return (unsigned)((n + 255) / 256); }

template <typename T, int W>
static void scatter_w(const PartArgs &a, bool fixed, int frac_bits, void *mesh, hipStream_t st) {
    double fs = ldexp(1.0, frac_bits);
    const T *psi = (const T *)a.psi, *q1 = (const T *)a.q1;
    T sx = (T)a.shift[0], sy = (T)a.shift[1], sz = (T)a.shift[2], L = (T)a.L;
    if (fixed)
        k_scatter<T, W, true><<<pgrid(a.n), 256, 0, st>>>(psi, q1, a.n, sx, sy, sz, mesh, a.res, L, fs);
    else
        k_scatter<T, W, false><<<pgrid(a.n), 256, 0, st>>>(psi, q1, a.n, sx, sy, sz, mesh, a.res, L, fs);
}

template <typename T>
static void scatter(const PartArgs &a, int W, bool fixed, int frac_bits, void *mesh, hipStream_t st) {
    if (W == 2) scatter_w<T, 2>(a, fixed, frac_bits, mesh, st);
    else if (W == 3) scatter_w<T, 3>(a, fixed, frac_bits, mesh, st);
    else scatter_w<T, 4>(a, fixed, frac_bits, mesh, st);
}

template <typename T, int W, int ORDER>
static void gather_fd_wo(const PartArgs &a, const void *phi, const double *c, int mode, const KickArgs &k,
                         void *psi_out, hipStream_t st) {
    const T *psi = (const T *)a.psi, *q1 = (const T *)a.q1;
    T sx = (T)a.shift[0], sy = (T)a.shift[1], sz = (T)a.shift[2], L = (T)a.L;
    T c1 = (T)c[0], c2 = (T)c[1], c3 = (T)c[2];
#define GF(M)                                                                                               \
    k_gather_fd<T, W, ORDER, M><<<pgrid(a.n), 256, 0, st>>>(                                                 \
        psi, (T *)psi_out, (T *)a.mom, (T *)a.acc, q1, a.n, sx, sy, sz, (const T *)phi, a.res, L, c1, c2, c3, \
        (T)k.alpha, (T)k.beta, (T)k.dd2, (T)k.dd1_next, k.has_next, k.nshifts)
    if (mode == GATHER_ACC) GF(GATHER_ACC);
    else if (mode == GATHER_ASSIGN) GF(GATHER_ASSIGN);
    else GF(GATHER_KICK);
#undef GF
}

template <typename T, int W>
static void gather_fd_w(const PartArgs &a, int order, const void *phi, const double *c, int mode,
                        const KickArgs &k, void *psi_out, hipStream_t st) {
    if (order == 2) gather_fd_wo<T, W, 2>(a, phi, c, mode, k, psi_out, st);
    else if (order == 4) gather_fd_wo<T, W, 4>(a, phi, c, mode, k, psi_out, st);
    else gather_fd_wo<T, W, 6>(a, phi, c, mode, k, psi_out, st);
}

template <typename T>
static void gather_fd(const PartArgs &a, int W, int order, const void *phi, const double *c, int mode,
                      const KickArgs &k, void *psi_out, hipStream_t st) {
    if (W == 2) gather_fd_w<T, 2>(a, order, phi, c, mode, k, psi_out, st);
    else if (W == 3) gather_fd_w<T, 3>(a, order, phi, c, mode, k, psi_out, st);
    else gather_fd_w<T, 4>(a, order, phi, c, mode, k, psi_out, st);
}

template <typename T>
static void gather_comp(const PartArgs &a, int W, const void *field, int comp, hipStream_t st) {
    const T *psi = (const T *)a.psi, *q1 = (const T *)a.q1;
    T sx = (T)a.shift[0], sy = (T)a.shift[1], sz = (T)a.shift[2], L = (T)a.L;
#define GC(w) \
    k_gather_comp<T, w><<<pgrid(a.n), 256, 0, st>>>(psi, (T *)a.acc, q1, a.n, sx, sy, sz, (const T *)field, a.res, L, comp)
    if (W == 2) GC(2);
    else if (W == 3) GC(3);
    else GC(4);
#undef GC
}

template <typename T>
static void drift(void *psi, const void *mom, int64_t n3, double dd, hipStream_t st) {
    k_drift<T><<<lgrid(n3), 256, 0, st>>>((T *)psi, (const T *)mom, n3, (T)dd);
}

template <typename T>
static void drift_to(void *out, const void *psi, const void *mom, void *ck_psi, void *ck_mom, int64_t n3, double dd,
                     hipStream_t st) {
    if (ck_psi && ck_mom)
        k_drift_to<T, true><<<lgrid(n3), 256, 0, st>>>((T *)out, (const T *)psi, (const T *)mom, (T *)ck_psi,
                                                      (T *)ck_mom, n3, (T)dd);
    else
        k_drift_to<T, false><<<lgrid(n3), 256, 0, st>>>((T *)out, (const T *)psi, (const T *)mom, nullptr, nullptr,
                                                       n3, (T)dd);
}

template <typename T>
static void kick_drift(void *psi, void *mom, const void *acc, int64_t n3, const KickArgs &k, hipStream_t st) {
    k_kick_drift<T><<<lgrid(n3), 256, 0, st>>>((T *)psi, (T *)mom, (const T *)acc, n3, (T)k.alpha, (T)k.beta,
                                               (T)k.dd2, (T)k.dd1_next, k.has_next, k.nshifts);
}

template <typename T>
static void scale(void *x, int64_t n, double s, hipStream_t st) {
    k_scale<T><<<lgrid(n), 256, 0, st>>>((T *)x, n, (T)s);
}

static void zero(void *x, size_t bytes, hipStream_t st) {
    size_t n16 = bytes / 16, tail = bytes - n16 * 16;
    unsigned blocks = (unsigned)((n16 + 255) / 256);
    if (blocks > 65536u * 4) blocks = 65536u * 4;
    if (blocks == 0) blocks = 1;
    k_zero<<<blocks, 256, 0, st>>>((uint4 *)x, n16, (uint8_t *)x + n16 * 16, tail);
}

template <typename T> void part_ops(PartOps &ops) {
    ops.scatter = scatter<T>;
    ops.gather_fd = gather_fd<T>;
    ops.gather_comp = gather_comp<T>;
    ops.drift = drift<T>;
    ops.drift_to = drift_to<T>;
    ops.kick_drift = kick_drift<T>;
    ops.scale = scale<T>;
    ops.zero = zero;
}

template void part_ops<float>(PartOps &);
template void part_ops<double>(PartOps &);

}  // namespace dj
