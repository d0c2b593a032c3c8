// Adjoint (reverse-mode) kernels of the PM force and the DKD step.  The reference has no adjoint of
// pm_acceleration! or run_nbody; the conventions follow its hand-written CIC rrule (the VJP of a deposit
// gathers the mesh cotangent with the weight gradients, NR/src/field/sheet_deposit.jl:75-94, 118-132),
// here for CIC, TSC and PCS.
//
// One interlacing pass of the force at X = q + psi + s is
//   rho = S(X) (unit-mass deposit),  phi = A rho  (A = c2r K_phi r2c),  F_a = -D_a phi (FD stencil),
//   acc = G(X) F  (the gather with the deposit's weights)
// and for a cotangent abar of acc its VJP is
//   Fbar_a   = G(X)^T abar_a                   vector deposit of abar with the weights      (scatter3)
//   phibar   = sum_a D_a Fbar_a                (D_a^T = -D_a: the stencil is antisymmetric)   (divergence)
//   rhobar   = A phibar                        (K_phi is real and even in k: A^T = A)
//   psibar  += sum_c grad W(X - c) (rhobar_c + sum_a abar_a F_a(c))                           (gather_grad)
// With the spectral gradient (grad_order 0) F_a = B_a rho, B_a = c2r (-g_a K_phi) r2c with g_a = i k_a (its
// Nyquist zeroed): -g_a K_phi is imaginary and odd in k, so B_a^T = -B_a and rhobar = -sum_a B_a Fbar_a, and
// the weight-gradient gather reads the three recomputed force meshes (gather_grad_f).
// The weights' slopes dW/dg follow the same branches as the weights; g = mod(X, L) / L N.
#include <hip/hip_runtime.h>

#include "dj_internal.h"

namespace dj {

template <typename T> __device__ __forceinline__ T adj_jmod(T x, T L) {
    if (x >= (T)0 && x < L) return x + (T)0;
    T r = fmod(x, L);
    if (r == (T)0) return copysign(r, L);
    if ((r > (T)0) != (L > (T)0)) return r + L;
    return r;
}

// 1-D weight w and slope dw/dg at signed offset d = g - cell (pm.jl:49-55 for the weights)
template <int W, typename T> __device__ __forceinline__ void wslope(T d, T &w, T &dw) {
    const T a = fabs(d), sg = d >= (T)0 ? (T)1 : (T)-1;
    if constexpr (W == 2) {
        w = a > (T)1 ? (T)0 : (T)1 - a;
        dw = a > (T)1 ? (T)0 : -sg;
    } else if constexpr (W == 3) {
        const T t = (T)1.5 - a;
        if (a > (T)1.5) { w = 0; dw = 0; }
        else if (a > (T)0.5) { w = (T)0.5 * (t * t); dw = -t * sg; }
        else { w = (T)0.75 - a * a; dw = (T)-2 * d; }
    } else {
        const T t = (T)2 - a;
        if (a > (T)2) { w = 0; dw = 0; }
        else if (a > (T)1) { w = t * t * t / (T)6; dw = (T)-0.5 * (t * t) * sg; }
        else { w = ((T)4 - (T)6 * (a * a) + (T)3 * (a * a * a)) / (T)6; dw = (T)-2 * d + (T)1.5 * d * a; }
    }
}

// Particle of a thread: the grid is (ceil(n / blockDim.x), n, n) over Lagrangian (i, j, k), so that no
// thread divides its index (the 64-bit divisions and remainders took most of these kernels' instructions)
struct AdjPart {
    int ii[3];
    size_t p;
};
__device__ __forceinline__ bool adj_particle(int n, AdjPart &q) {
    q.ii[0] = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    q.ii[1] = (int)blockIdx.y;
    q.ii[2] = (int)blockIdx.z;
    q.p = (size_t)q.ii[0] + (size_t)n * ((size_t)q.ii[1] + (size_t)n * q.ii[2]);
    return q.ii[0] < n;
}

// per-axis stencil of particle q: cells (wrapped: b + lo + o lies in [-1, N + 2]), weights, slopes
template <typename T, int W>
__device__ __forceinline__ void adj_stencil(const T *psi, const T *q1, size_t n3, const AdjPart &q, int N, T L,
                                            const T (&sh)[3], int (&cell)[3][W], T (&w)[3][W], T (&dw)[3][W]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T X = (psi[q.p + a * n3] + q1[q.ii[a]]) + sh[a];
        const T g = adj_jmod(X, L) / L * (T)N;
        const int b = (W % 2 == 0) ? (int)floor(g) : (int)rint(g);
        const int lo = W == 2 ? 0 : -1;
#pragma unroll
        for (int o = 0; o < W; o++) {
            const int bo = b + lo + o;
            wslope<W>(g - (T)bo, w[a][o], dw[a][o]);
            cell[a][o] = bo < 0 ? bo + N : (bo >= N ? bo - N : bo);
        }
    }
}

// Fbar_a += sum_p W(X_p - c) abar_p,a s   (three meshes, device atomics)
template <typename T, int W>
__global__ __launch_bounds__(256) void k_adj_scatter3(const T *__restrict__ psi, const T *__restrict__ q1, int n, int N,
                                                      T L, T sx, T sy, T sz, const T *__restrict__ abar, T s,
                                                      T *__restrict__ fbar) {
    const size_t n3 = (size_t)n * n * n, M = (size_t)N * N * N;
    AdjPart q;
    if (!adj_particle(n, q)) return;
    const size_t p = q.p;
    const T sh[3] = {sx, sy, sz};
    int cell[3][W];
    T w[3][W], dw[3][W];
    adj_stencil<T, W>(psi, q1, n3, q, N, L, sh, cell, w, dw);
    const T v0 = abar[p] * s, v1 = abar[p + n3] * s, v2 = abar[p + 2 * n3] * s;
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++) {
            const size_t row = (size_t)N * ((size_t)cell[1][oy] + (size_t)N * cell[2][oz]);
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const T wv = w[0][ox] * w[1][oy] * w[2][oz];
                const size_t c = row + cell[0][ox];
                unsafeAtomicAdd(fbar + c, wv * v0);
                unsafeAtomicAdd(fbar + M + c, wv * v1);
                unsafeAtomicAdd(fbar + 2 * M + c, wv * v2);
            }
        }
}

// phibar(c) = sum_a sum_t c_t (Fbar_a(c + t e_a) - Fbar_a(c - t e_a))
template <typename T>
__global__ __launch_bounds__(256) void k_adj_div(const T *__restrict__ fbar, int N, T c1, T c2, T c3,
                                                 T *__restrict__ phibar) {
    const size_t M = (size_t)N * N * N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M) return;
    const int x = (int)(e % N), y = (int)((e / N) % N), z = (int)(e / ((size_t)N * N));
    const T ct[3] = {c1, c2, c3};
    auto at = [&](int a, int xx, int yy, int zz) {
        return fbar[(size_t)a * M + (size_t)xx + (size_t)N * ((size_t)yy + (size_t)N * zz)];
    };
    auto wr = [&](int v) { return v < 0 ? v + N : (v >= N ? v - N : v); };
    T r = 0;
#pragma unroll
    for (int t = 1; t <= 3; t++) {
        if (ct[t - 1] == (T)0) continue;
        const T d = (at(0, wr(x + t), y, z) - at(0, wr(x - t), y, z)) + (at(1, x, wr(y + t), z) - at(1, x, wr(y - t), z)) +
                    (at(2, x, y, wr(z + t)) - at(2, x, y, wr(z - t)));
        r += ct[t - 1] * d;
    }
    phibar[e] = r;
}

// A cell's four values for the weight-gradient gather: rhobar and the three force components, one
// vector load per stencil cell instead of four scalar ones
template <typename T> struct alignas(4 * sizeof(T)) Cell4 { T r, f0, f1, f2; };

// The three force meshes of the FD gradient, F_a(c) = -(c1 (phi(c + e_a) - phi(c - e_a)) + c2 (...) + c3 (...)),
// summed in k_adj_gather's order and packed with rhobar into V[c] = {rhobar, F_0, F_1, F_2}, so that
// k_adj_gather_v over them gives k_adj_gather's values bit for bit while each particle reads W^3 cells of
// four values instead of W^3 (1 + 6 H) mesh values (FD order 2 H)
template <typename T>
__global__ __launch_bounds__(256) void k_fd_force3_packed(const T *__restrict__ phi, const T *__restrict__ rhobar,
                                                          int N, T c1, T c2, T c3, Cell4<T> *__restrict__ V) {
    const size_t M = (size_t)N * N * N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M) return;
    const int x = (int)(e % N), y = (int)((e / N) % N), z = (int)(e / ((size_t)N * N));
    const T ct[3] = {c1, c2, c3};
    auto wr = [&](int v) { return v < 0 ? v + N : (v >= N ? v - N : v); };
    auto P = [&](int xx, int yy, int zz) { return phi[(size_t)xx + (size_t)N * ((size_t)yy + (size_t)N * zz)]; };
    T f0 = 0, f1 = 0, f2 = 0;
    for (int t = 1; t <= 3; t++) {
        if (ct[t - 1] == (T)0) continue;
        f0 += ct[t - 1] * (P(wr(x + t), y, z) - P(wr(x - t), y, z));
        f1 += ct[t - 1] * (P(x, wr(y + t), z) - P(x, wr(y - t), z));
        f2 += ct[t - 1] * (P(x, y, wr(z + t)) - P(x, y, wr(z - t)));
    }
    V[e] = Cell4<T>{rhobar[e], -f0, -f1, -f2};
}

// psibar_p += N/L sum_c grad W(X_p - c) (rhobar_c + sum_a abar_p,a s F_a(c)),
// F_a(c) = -(c1 (phi(c + e_a) - phi(c - e_a)) + c2 (...) + c3 (...));  rhobar or phi may be null
template <typename T, int W>
__global__ __launch_bounds__(256) void k_adj_gather(const T *__restrict__ psi, const T *__restrict__ q1, int n, int N,
                                                    T L, T sx, T sy, T sz, const T *__restrict__ rhobar,
                                                    const T *__restrict__ phi, const T *__restrict__ abar, T s, T c1,
                                                    T c2, T c3, T *__restrict__ psibar) {
    const size_t n3 = (size_t)n * n * n;
    AdjPart q;
    if (!adj_particle(n, q)) return;
    const size_t p = q.p;
    const T sh[3] = {sx, sy, sz};
    int cell[3][W];
    T w[3][W], dw[3][W];
    adj_stencil<T, W>(psi, q1, n3, q, N, L, sh, cell, w, dw);
    T ab[3] = {0, 0, 0};
    if (phi) {
#pragma unroll
        for (int a = 0; a < 3; a++) ab[a] = abar[p + a * n3] * s;
    }
    const T ct[3] = {c1, c2, c3};
    auto wr = [&](int v) { return v < 0 ? v + N : (v >= N ? v - N : v); };
    auto P = [&](int xx, int yy, int zz) { return phi[(size_t)xx + (size_t)N * ((size_t)yy + (size_t)N * zz)]; };
    T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const int x = cell[0][ox], y = cell[1][oy], z = cell[2][oz];
                T v = rhobar ? rhobar[(size_t)x + (size_t)N * ((size_t)y + (size_t)N * z)] : (T)0;
                if (phi) {
                    T f0 = 0, f1 = 0, f2 = 0;
                    for (int t = 1; t <= 3; t++) {
                        if (ct[t - 1] == (T)0) continue;
                        f0 += ct[t - 1] * (P(wr(x + t), y, z) - P(wr(x - t), y, z));
                        f1 += ct[t - 1] * (P(x, wr(y + t), z) - P(x, wr(y - t), z));
                        f2 += ct[t - 1] * (P(x, y, wr(z + t)) - P(x, y, wr(z - t)));
                    }
                    v -= ab[0] * f0 + ab[1] * f1 + ab[2] * f2;
                }
                g0 += dw[0][ox] * w[1][oy] * w[2][oz] * v;
                g1 += w[0][ox] * dw[1][oy] * w[2][oz] * v;
                g2 += w[0][ox] * w[1][oy] * dw[2][oz] * v;
            }
    const T k = (T)N / L;
    psibar[p] += k * g0;
    psibar[p + n3] += k * g1;
    psibar[p + 2 * n3] += k * g2;
}

// The same for the spectral (ik) gradient, whose forward gathers the three force meshes F (3 N^3):
// psibar_p += N/L sum_c grad W(X_p - c) (rhobar_c + sum_a abar_p,a s F_a(c))
template <typename T, int W>
__global__ __launch_bounds__(256) void k_adj_gather_f(const T *__restrict__ psi, const T *__restrict__ q1, int n, int N,
                                                      T L, T sx, T sy, T sz, const T *__restrict__ rhobar,
                                                      const T *__restrict__ F, const T *__restrict__ abar, T s,
                                                      T *__restrict__ psibar) {
    const size_t n3 = (size_t)n * n * n, M = (size_t)N * N * N;
    AdjPart q;
    if (!adj_particle(n, q)) return;
    const size_t p = q.p;
    const T sh[3] = {sx, sy, sz};
    int cell[3][W];
    T w[3][W], dw[3][W];
    adj_stencil<T, W>(psi, q1, n3, q, N, L, sh, cell, w, dw);
    const T ab[3] = {abar[p] * s, abar[p + n3] * s, abar[p + 2 * n3] * s};
    const T *F1 = F + M, *F2 = F + 2 * M;
    // the W^2 rows of a z plane of the stencil: all their mesh values are read before any is used
    // (cell offsets < N^3 <= 2^30)
    T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
    for (int oz = 0; oz < W; oz++) {
        T vr[W][W], vf0[W][W], vf1[W][W], vf2[W][W];
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const unsigned c = (unsigned)cell[0][ox] + (unsigned)N * ((unsigned)cell[1][oy] + (unsigned)N * (unsigned)cell[2][oz]);
                vr[oy][ox] = rhobar[c];
                vf0[oy][ox] = F[c];
                vf1[oy][ox] = F1[c];
                vf2[oy][ox] = F2[c];
            }
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const T v = vr[oy][ox] + (ab[0] * vf0[oy][ox] + ab[1] * vf1[oy][ox] + ab[2] * vf2[oy][ox]);
                g0 += dw[0][ox] * w[1][oy] * w[2][oz] * v;
                g1 += w[0][ox] * dw[1][oy] * w[2][oz] * v;
                g2 += w[0][ox] * w[1][oy] * dw[2][oz] * v;
            }
    }
    const T k = (T)N / L;
    psibar[p] += k * g0;
    psibar[p + n3] += k * g1;
    psibar[p + 2 * n3] += k * g2;
}

// k_adj_gather_f over the packed cells of k_fd_force3_packed (the same values, the same sums).  TAIL: the
// step's reverse ends here for each particle, with k_adj_drift's update of pibar and, when kick, the
// previous step's k_adj_kick (abar of that step into abar, pibar), each with their kernels' arithmetic
template <typename T, int W, bool TAIL>
__global__ __launch_bounds__(256) void k_adj_gather_v(const T *__restrict__ psi, const T *__restrict__ q1, int n, int N,
                                                      T L, T sx, T sy, T sz, const Cell4<T> *__restrict__ V, T *abar,
                                                      T s, T *__restrict__ psibar, T *__restrict__ pibar, T dd1,
                                                      int kick, T alpha, T beta, T dd2) {
    const size_t n3 = (size_t)n * n * n;
    AdjPart q;
    if (!adj_particle(n, q)) return;
    const size_t p = q.p;
    const T sh[3] = {sx, sy, sz};
    int cell[3][W];
    T w[3][W], dw[3][W];
    adj_stencil<T, W>(psi, q1, n3, q, N, L, sh, cell, w, dw);
    const T ab[3] = {abar[p] * s, abar[p + n3] * s, abar[p + 2 * n3] * s};
    // a z plane's W^2 cells are read before any is used (cell offsets < N^3 <= 2^30)
    T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
    for (int oz = 0; oz < W; oz++) {
        Cell4<T> vc[W][W];
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++)
                vc[oy][ox] = V[(unsigned)cell[0][ox] +
                               (unsigned)N * ((unsigned)cell[1][oy] + (unsigned)N * (unsigned)cell[2][oz])];
#pragma unroll
        for (int oy = 0; oy < W; oy++)
#pragma unroll
            for (int ox = 0; ox < W; ox++) {
                const Cell4<T> &u = vc[oy][ox];
                const T v = u.r + (ab[0] * u.f0 + ab[1] * u.f1 + ab[2] * u.f2);
                g0 += dw[0][ox] * w[1][oy] * w[2][oz] * v;
                g1 += w[0][ox] * dw[1][oy] * w[2][oz] * v;
                g2 += w[0][ox] * w[1][oy] * dw[2][oz] * v;
            }
    }
    const T k = (T)N / L;
    if constexpr (!TAIL) {
        psibar[p] += k * g0;
        psibar[p + n3] += k * g1;
        psibar[p + 2 * n3] += k * g2;
    } else {
        const T g[3] = {g0, g1, g2};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const size_t i = p + c * n3;
            const T ps = psibar[i] + k * g[c];
            psibar[i] = ps;
            T pb = pibar[i] + dd1 * ps;
            if (kick) {
                const T pt = pb + dd2 * ps;
                abar[i] = beta * pt;
                pb = alpha * pt;
            }
            pibar[i] = pb;
        }
    }
}

// y += s x
template <typename T>
__global__ __launch_bounds__(256) void k_adj_axpy(T *__restrict__ y, const T *__restrict__ x, size_t n, T s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + s * x[i];
}

// backward of the kick and second drift of a DKD step (run_nbody.jl:156-158):
//   Pibar_t = Pibar + dd2 Psibar;  abar = beta Pibar_t;  Pibar = alpha Pibar_t
template <typename T>
__global__ __launch_bounds__(256) void k_adj_kick(const T *__restrict__ psibar, T *__restrict__ pibar,
                                                  T *__restrict__ abar, size_t n, T alpha, T beta, T dd2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T pt = pibar[i] + dd2 * psibar[i];
    abar[i] = beta * pt;
    pibar[i] = alpha * pt;
}

// backward of the first drift: Pibar += dd1 Psibar
template <typename T>
__global__ __launch_bounds__(256) void k_adj_drift(const T *__restrict__ psibar, T *__restrict__ pibar, size_t n, T dd1) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pibar[i] = pibar[i] + dd1 * psibar[i];
}

// The density loss L = sum (delta - target)^2, delta = mesh / rho - 1, and its cotangent mesh_bar =
// (2 / rho)(delta - target) rounded to T (the host path's operations, in its order).  Each of the
// LOSS_BLOCKS blocks sums its grid-stride share of r^2 in a fixed order into part[block]; k_loss_total
// adds the block sums in a fixed order, so the loss is the same in every call.
constexpr int LOSS_BLOCKS = 1024;
template <typename T>
__global__ __launch_bounds__(256) void k_adj_mse(const T *__restrict__ mesh, const double *__restrict__ target,
                                                 size_t n, double rho, T *__restrict__ mesh_bar,
                                                 double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)LOSS_BLOCKS * 256) {
        const double delta = (double)mesh[i] / rho - 1.0;
        const double r = target ? delta - target[i] : delta;
        mesh_bar[i] = (T)((2.0 / rho) * r);
        acc += r * r;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_loss_total(double *part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < LOSS_BLOCKS; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[LOSS_BLOCKS] = red[0];
}

static unsigned agrid(size_t n) { return (unsigned)((n + 255) / 256); }
// particle kernels: blocks of up to 256 threads along i, one block row per (j, k)
static unsigned pblock(int n) { return n >= 256 ? 256u : (unsigned)(((n + 63) / 64) * 64); }
static dim3 pgrid(int n) { return dim3((unsigned)((n + pblock(n) - 1) / pblock(n)), (unsigned)n, (unsigned)n); }

template <typename T>
struct AdjLaunch {
    static void scatter3(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *abar,
                         double s, int W, void *fbar, hipStream_t st) {
#define ADJ_S(WW)                                                                                                   \
    k_adj_scatter3<T, WW><<<pgrid(n), pblock(n), 0, st>>>((const T *)psi, (const T *)q1, n, N, (T)L, (T)sh[0], (T)sh[1], \
                                                     (T)sh[2], (const T *)abar, (T)s, (T *)fbar)
        if (W == 2) ADJ_S(2);
        else if (W == 3) ADJ_S(3);
        else ADJ_S(4);
#undef ADJ_S
    }
    static void divergence(const void *fbar, int N, const double *c, void *phibar, hipStream_t st) {
        k_adj_div<T><<<agrid((size_t)N * N * N), 256, 0, st>>>((const T *)fbar, N, (T)c[0], (T)c[1], (T)c[2], (T *)phibar);
    }
    static void gather_grad(const void *psi, const void *q1, int n, int N, double L, const double *sh,
                            const void *rhobar, const void *phi, const void *abar, double s, const double *c, int W,
                            void *psibar, hipStream_t st) {
#define ADJ_G(WW)                                                                                                   \
    k_adj_gather<T, WW><<<pgrid(n), pblock(n), 0, st>>>((const T *)psi, (const T *)q1, n, N, (T)L, (T)sh[0], (T)sh[1],   \
                                                   (T)sh[2], (const T *)rhobar, (const T *)phi, (const T *)abar, (T)s, \
                                                   (T)c[0], (T)c[1], (T)c[2], (T *)psibar)
        if (W == 2) ADJ_G(2);
        else if (W == 3) ADJ_G(3);
        else ADJ_G(4);
#undef ADJ_G
    }
    static void gather_grad_f(const void *psi, const void *q1, int n, int N, double L, const double *sh,
                              const void *rhobar, const void *F, const void *abar, double s, int W, void *psibar,
                              hipStream_t st) {
#define ADJ_F(WW)                                                                                                    \
    k_adj_gather_f<T, WW><<<pgrid(n), pblock(n), 0, st>>>((const T *)psi, (const T *)q1, n, N, (T)L, (T)sh[0], (T)sh[1], \
                                                     (T)sh[2], (const T *)rhobar, (const T *)F, (const T *)abar, (T)s, \
                                                     (T *)psibar)
        if (W == 2) ADJ_F(2);
        else if (W == 3) ADJ_F(3);
        else ADJ_F(4);
#undef ADJ_F
    }
    static void gather_grad_v(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *V,
                              void *abar, double s, int W, void *psibar, const AdjTail *tl, hipStream_t st) {
        const AdjTail t0{};
        const AdjTail &t = tl ? *tl : t0;
#define ADJ_V(WW, TL)                                                                                                  \
    k_adj_gather_v<T, WW, TL><<<pgrid(n), pblock(n), 0, st>>>(                                                         \
        (const T *)psi, (const T *)q1, n, N, (T)L, (T)sh[0], (T)sh[1], (T)sh[2], (const Cell4<T> *)V, (T *)abar, (T)s, \
        (T *)psibar, (T *)t.pibar, (T)t.dd1, t.kick, (T)t.alpha, (T)t.beta, (T)t.dd2)
        if (tl) {
            if (W == 2) ADJ_V(2, true);
            else if (W == 3) ADJ_V(3, true);
            else ADJ_V(4, true);
        } else {
            if (W == 2) ADJ_V(2, false);
            else if (W == 3) ADJ_V(3, false);
            else ADJ_V(4, false);
        }
#undef ADJ_V
    }
    static void fd_force3_packed(const void *phi, const void *rhobar, int N, const double *c, void *V,
                                 hipStream_t st) {
        k_fd_force3_packed<T><<<agrid((size_t)N * N * N), 256, 0, st>>>((const T *)phi, (const T *)rhobar, N, (T)c[0],
                                                                        (T)c[1], (T)c[2], (Cell4<T> *)V);
    }
    static void axpy(void *y, const void *x, size_t n, double s, hipStream_t st) {
        k_adj_axpy<T><<<agrid(n), 256, 0, st>>>((T *)y, (const T *)x, n, (T)s);
    }
    static void kick(const void *psibar, void *pibar, void *abar, size_t n, double alpha, double beta, double dd2,
                     hipStream_t st) {
        k_adj_kick<T><<<agrid(n), 256, 0, st>>>((const T *)psibar, (T *)pibar, (T *)abar, n, (T)alpha, (T)beta, (T)dd2);
    }
    static void drift(const void *psibar, void *pibar, size_t n, double dd1, hipStream_t st) {
        k_adj_drift<T><<<agrid(n), 256, 0, st>>>((const T *)psibar, (T *)pibar, n, (T)dd1);
    }
    static void mse(const void *mesh, const double *target, size_t n, double rho, void *mesh_bar, double *part,
                    hipStream_t st) {
        k_adj_mse<T><<<LOSS_BLOCKS, 256, 0, st>>>((const T *)mesh, target, n, rho, (T *)mesh_bar, part);
        k_loss_total<<<1, 256, 0, st>>>(part);
    }
};

template <typename T> void adjoint_ops(AdjointOps &ops) {
    ops.scatter3 = AdjLaunch<T>::scatter3;
    ops.divergence = AdjLaunch<T>::divergence;
    ops.gather_grad = AdjLaunch<T>::gather_grad;
    ops.gather_grad_f = AdjLaunch<T>::gather_grad_f;
    ops.axpy = AdjLaunch<T>::axpy;
    ops.kick = AdjLaunch<T>::kick;
    ops.drift = AdjLaunch<T>::drift;
    ops.mse = AdjLaunch<T>::mse;
    ops.fd_force3_packed = AdjLaunch<T>::fd_force3_packed;
    ops.gather_grad_v = AdjLaunch<T>::gather_grad_v;
}

template void adjoint_ops<float>(AdjointOps &);
template void adjoint_ops<double>(AdjointOps &);

}  // namespace dj
