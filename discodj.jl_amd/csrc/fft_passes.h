// The five passes of the PM Poisson solve on an N^3 mesh, as barrier-separated phases.
//
// HBM layouts (complex half spectrum, x halved: H = N/2 + 1):
//   mesh  real   R[x + N (y + N z)]                the density / potential (column-major N^3)
//   A     cx     A[y + N (kx + H z)]               after the x pass: y lines contiguous
//   B     cx     B[z + N (kx + H ky)]              after the y pass: z lines contiguous
// Pass sequence of one Poisson solve (each pass reads and writes HBM once):
//   XF  r2c along x      mesh -> A   (lines of B consecutive y; stores transposed, B-wide chunks)
//   YF  c2c along y      A -> B      (lines of B consecutive z; stores transposed)
//   ZM  c2c along z, multiplier, inverse along z, in place on B (no transpose)
//   YI  inverse along y  B -> A
//   XI  c2r along x      A -> mesh   (Re of the kx = 0 and kx = N/2 planes, as c2r does)
// The x axis is halved (the reference halves its slowest axis): every multiplier of the PM force
// is even in k, so the choice of halved axis only changes round-off.
//
// Phases are __host__ __device__ so that tests can run a whole workgroup on the host; the
// __global__ wrappers in fft.hip call them with __syncthreads() between phases.
#pragma once

#include <type_traits>

#include "fft_engine.h"

namespace dj {

// k-space multiplier of one Poisson solve, from 1-D tables built on the host in FP64 and rounded
// to T (the reference builds them in T as well): K_phi = invlap * mak / (rho nsets N^3), DC = 0
template <typename T>
struct KMult {
    const T *k2f, *k2h;        // (k^2) per full / half axis index      (lap_order == 0)
    const T *tf, *th;          // FD Laplace symbol per full / half index (lap_order != 0)
    const T *mf, *mh;          // deconvolution window^-2W, or null
    const cx<T> *gf, *gh;      // gradient kernels (ik path) full / half
    T scale;                   // 1 / (rho nsets N^3), rho = n^3/N^3, built in FP64 and rounded once
    int lap_order;
    // optional isotropic transfer function g(|k|) of the kernel (a fitted correction): kg_n values
    // on a uniform grid |k| = i / kg_inv_dk, linear in between, the last value beyond; null = none
    const T *kg;
    T kg_inv_dk;
    int kg_n;

    DJ_HD T kphi(int kx, int ky, int kz) const {
        if (kx == 0 && ky == 0 && kz == 0) return (T)0;
        T il;
        if (lap_order == 0) {
            T k2 = k2h[kx] + k2f[ky] + k2f[kz];
            il = (T)-1 / k2;
        } else {
            T inv = th[kx] + tf[ky] + tf[kz];
            il = (T)1 / inv;
        }
        T v = il;
        if (mf) v = il * (mh[kx] * mf[ky] * mf[kz]);
        if (kg) {
            const T u = sqrt(k2h[kx] + k2f[ky] + k2f[kz]) * kg_inv_dk;
            const int i = (int)u;
            if (i >= kg_n - 1) {
                v = v * kg[kg_n - 1];
            } else {
                const T f = u - (T)i;
                v = v * (kg[i] + f * (kg[i + 1] - kg[i]));
            }
        }
        return v * scale;
    }
};

enum { ZM_POT = 0, ZM_GRADX = 1, ZM_GRADY = 2, ZM_GRADZ = 3, ZM_NONE = 4 };

// Wave-local tile copy: wave w moves elements [w E, (w + 1) E) of a workgroup's tile (E = the wave's
// lines times the line length), 64 consecutive ones per step, all reads before the first write.
template <typename V, int E, typename SRC, typename DST>
DJ_HD void copy_wave(int wave, int lane, const SRC &src, const DST &dst) {
    constexpr int CNT = (E + 63) / 64;
    V v[CNT];
#pragma unroll
    for (int c = 0; c < CNT; c++) {
        const int i = lane + 64 * c;
        if (E % 64 == 0 || i < E) v[c] = src(wave * E + i);
    }
#pragma unroll
    for (int c = 0; c < CNT; c++) {
        const int i = lane + 64 * c;
        if (E % 64 == 0 || i < E) dst(wave * E + i, v[c]);
    }
}

// Tile copy in two phases: a thread issues all of its TOTAL / NT reads (src(i)) before its first
// write (dst(i, v)), so the memory latencies overlap instead of adding up per element.
template <typename V, int TOTAL, int NT, typename SRC, typename DST>
DJ_HD void copy_batched(int tid, const SRC &src, const DST &dst) {
    constexpr int CNT = (TOTAL + NT - 1) / NT;
    V v[CNT];
#pragma unroll
    for (int c = 0; c < CNT; c++) {
        const int i = tid + c * NT;
        if (TOTAL % NT == 0 || i < TOTAL) v[c] = src(i);
    }
#pragma unroll
    for (int c = 0; c < CNT; c++) {
        const int i = tid + c * NT;
        if (TOTAL % NT == 0 || i < TOTAL) dst(i, v[c]);
    }
}

template <typename T, int N, int B, int NT>
struct Passes {
    static constexpr int NC = N / 2;      // complex length of the packed real x lines
    static constexpr int H = N / 2 + 1;
    using EX = LineFFT<T, NC, B, NT>;     // x pass engine (packed real lines)
    using EF = LineFFT<T, N, B, NT>;      // y and z pass engine
    static_assert(N % B == 0, "B must divide N");

    // ------------------------------------------------------------------ XF: mesh -> A
    // block (yb, z): lines y = yb*B + l.  IN is T (float mesh) or int64 (fixed-point mesh).
    template <typename IN>
    DJ_HD static void xf_load(const IN *mesh, cx<T> *lds, int yb, int z, int tid, double fixed_scale) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_batched<cx<T>, B * NC, NT>(
            tid,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                const size_t g = base + (size_t)N * l + 2 * m;
                if constexpr (sizeof(IN) == 8 && !std::is_floating_point<IN>::value)
                    return cx<T>{(T)((double)mesh[g] * fixed_scale), (T)((double)mesh[g + 1] * fixed_scale)};
                else
                    return cx<T>{(T)mesh[g], (T)mesh[g + 1]};
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                lds[EX::idx(l, m)] = v;
            });
    }
    // the same loads from a computed real field: src(g) returns the pair of reals (g, g + 1) as a complex
    template <typename SRC>
    DJ_HD static void xf_load_src(const SRC &src, cx<T> *lds, int yb, int z, int tid) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_batched<cx<T>, B * NC, NT>(
            tid,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                return src(base + (size_t)N * l + 2 * m);
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                lds[EX::idx(l, m)] = v;
            });
    }
    template <typename SRC>
    DJ_HD static void xf_load_src_w(const SRC &src, cx<T> *lds, int yb, int z, int tid) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_wave<cx<T>, EX::LW * NC>(
            tid >> 6, tid & 63,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                return src(base + (size_t)N * l + 2 * m);
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                lds[EX::idx(l, m)] = v;
            });
    }
    // X[k] = E[k] + W_N^k O[k], E = (Z[k] + conj Z[NC-k]) / 2, O = (Z[k] - conj Z[NC-k]) / 2i
    DJ_HD static void xf_store(const cx<T> *lds, const cx<T> *twN, cx<T> *A, int yb, int z, int tid) {
        copy_batched<cx<T>, B * H, NT>(
            tid,
            [&](int i) {
                const int kx = i / B, l = i - kx * B;
                const cx<T> z1 = lds[EX::idx(l, kx == NC ? 0 : kx)];
                const cx<T> z2 = conj(lds[EX::idx(l, kx == 0 ? 0 : NC - kx)]);
                const cx<T> e = scale(z1 + z2, (T)0.5);
                const cx<T> d = z1 - z2;
                const cx<T> o = {d.y * (T)0.5, -d.x * (T)0.5};
                return e + cmul(o, twN[kx]);
            },
            [&](int i, cx<T> x) {
                const int kx = i / B, l = i - kx * B;
                A[(size_t)(yb * B + l) + (size_t)N * (kx + (size_t)H * z)] = x;
            });
    }

    // ------------------------------------------------------------------ XI: A -> mesh
    // The load fused with the packed-real pre-processing of c2r: work item (k, l), k <= NC/2, reads
    // X[k] and X[NC-k] of line l and writes Z[k] = (X[k] + conj X[NC-k]) + i (X[k] - conj X[NC-k]) W_N^-k
    // and Z[NC-k] likewise; the imaginary parts of X[0] and X[NC] are dropped (c2r of a Hermitian line).
    // All reads of a thread are issued before its first LDS write.
    static constexpr int XI_CNT = (B * (NC / 2 + 1) + NT - 1) / NT;  // work items per thread
    DJ_HD static void xi_fetch(const cx<T> *A, int yb, int z, int tid, cx<T> (&va)[XI_CNT], cx<T> (&vb)[XI_CNT]) {
        constexpr int HP = NC / 2 + 1, TOTAL = B * HP;
#pragma unroll
        for (int c = 0; c < XI_CNT; c++) {
            const int i = tid + c * NT;
            if (TOTAL % NT == 0 || i < TOTAL) {
                const int k = i / B, l = i - k * B;
                const size_t row = (size_t)(yb * B + l) + (size_t)N * (size_t)H * z;
                va[c] = A[row + (size_t)N * k];
                vb[c] = A[row + (size_t)N * (NC - k)];
            }
        }
    }
    DJ_HD static void xi_pre(cx<T> *lds, const cx<T> *twN, int tid, const cx<T> (&va)[XI_CNT],
                             const cx<T> (&vb)[XI_CNT]) {
        constexpr int HP = NC / 2 + 1, TOTAL = B * HP;
#pragma unroll
        for (int c = 0; c < XI_CNT; c++) {
            const int i = tid + c * NT;
            if (!(TOTAL % NT == 0 || i < TOTAL)) continue;
            const int k = i / B, l = i - k * B;
            cx<T> a = va[c], b = vb[c];
            if (k == 0) { a.y = 0; b.y = 0; }
            const cx<T> e = a + conj(b);
            const cx<T> o = cmulc(a - conj(b), twN[k]);
            lds[EX::idx(l, k)] = {e.x - o.y, e.y + o.x};
            if (k != 0 && 2 * k != NC) {
                const cx<T> e2 = b + conj(a);
                const cx<T> o2 = cmulc(b - conj(a), twN[NC - k]);
                lds[EX::idx(l, NC - k)] = {e2.x - o2.y, e2.y + o2.x};
            }
        }
    }
    // all reads of a thread are issued before its first LDS write
    DJ_HD static void xi_load_pre(const cx<T> *A, cx<T> *lds, const cx<T> *twN, int yb, int z, int tid) {
        cx<T> va[XI_CNT], vb[XI_CNT];
        xi_fetch(A, yb, z, tid, va, vb);
        xi_pre(lds, twN, tid, va, vb);
    }
    // SC: each value multiplied by s as it is stored (a normalisation that would otherwise be a pass of its own)
    template <bool SC = false>
    DJ_HD static void xi_store(const cx<T> *lds, T *mesh, int yb, int z, int tid, T s = (T)1) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_batched<cx<T>, B * NC, NT>(
            tid,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                return lds[EX::idx(l, m)];
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                const size_t g = base + (size_t)N * l + 2 * m;
                mesh[g] = SC ? v.x * s : v.x;
                mesh[g + 1] = SC ? v.y * s : v.y;
            });
    }

    // ------------------------------------------------------------------ YF: A -> B, YI: B -> A
    // block (zb, kx): lines z = zb*B + l
    DJ_HD static size_t a_at(int y, int kx, int z) { return (size_t)y + (size_t)N * (kx + (size_t)H * z); }
    DJ_HD static size_t b_at(int z, int kx, int ky) { return (size_t)z + (size_t)N * (kx + (size_t)H * ky); }
    DJ_HD static void y_load_a(const cx<T> *A, cx<T> *lds, int zb, int kx, int tid) {
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int l = i / N, y = i - l * N; return A[a_at(y, kx, zb * B + l)]; },
            [&](int i, cx<T> v) { const int l = i / N, y = i - l * N; lds[EF::idx(l, y)] = v; });
    }
    DJ_HD static void y_store_b(const cx<T> *lds, cx<T> *Bm, int zb, int kx, int tid) {
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int ky = i / B, l = i - ky * B; return lds[EF::idx(l, ky)]; },
            [&](int i, cx<T> v) { const int ky = i / B, l = i - ky * B; Bm[b_at(zb * B + l, kx, ky)] = v; });
    }
    DJ_HD static void y_load_b(const cx<T> *Bm, cx<T> *lds, int zb, int kx, int tid) {
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int ky = i / B, l = i - ky * B; return Bm[b_at(zb * B + l, kx, ky)]; },
            [&](int i, cx<T> v) { const int ky = i / B, l = i - ky * B; lds[EF::idx(l, ky)] = v; });
    }
    DJ_HD static void y_store_a(const cx<T> *lds, cx<T> *A, int zb, int kx, int tid) {
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int l = i / N, y = i - l * N; return lds[EF::idx(l, y)]; },
            [&](int i, cx<T> v) { const int l = i / N, y = i - l * N; A[a_at(y, kx, zb * B + l)] = v; });
    }

    // ------------------------------------------------------------------ ZM: on B lines
    // block b: line ids L = b*B + l, kx = L % H, ky = L / H
    DJ_HD static void z_load(const cx<T> *Bm, cx<T> *lds, int b, int tid) {
        const cx<T> *src = Bm + (size_t)N * B * b;
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { return src[i]; },
            [&](int i, cx<T> v) { const int l = i / N, zz = i - l * N; lds[EF::idx(l, zz)] = v; });
    }
    // multiplier of element i (line l = i / N, kz = i % N) of block b's tile
    template <int MODE>
    DJ_HD static void z_mult_one(cx<T> *lds, const KMult<T> &K, int b, int i) {
        const int l = i / N, kz = i - l * N;
        const int L = b * B + l, ky = L / H, kx = L - ky * H;
        cx<T> v = lds[EF::idx(l, kz)];
        const T kp = K.kphi(kx, ky, kz);
        if constexpr (MODE == ZM_POT) {
            v = scale(v, kp);
        } else if constexpr (MODE == ZM_GRADX || MODE == ZM_GRADY || MODE == ZM_GRADZ) {
            const cx<T> g = MODE == ZM_GRADX ? K.gh[kx] : MODE == ZM_GRADY ? K.gf[ky] : K.gf[kz];
            const cx<T> gk = scale(g, kp);
            const cx<T> r = {gk.x * v.x - gk.y * v.y, gk.x * v.y + gk.y * v.x};
            v = {-r.x, -r.y};
        }
        lds[EF::idx(l, kz)] = v;
    }
    template <int MODE>
    DJ_HD static void z_mult(cx<T> *lds, const KMult<T> &K, int b, int tid) {
        // fixed trip count, unrolled: the multiplier-table reads of all iterations are in flight together
        constexpr int CNT = (B * N + NT - 1) / NT;
#pragma unroll
        for (int c = 0; c < CNT; c++) {
            const int i = tid + c * NT;
            if ((B * N) % NT == 0 || i < B * N) z_mult_one<MODE>(lds, K, b, i);
        }
    }
    // the z lines of block b from / to functors of the B-layout entry index (src(e), dst(e, v))
    template <typename SRC>
    DJ_HD static void z_load_src(const SRC &src, cx<T> *lds, int b, int tid) {
        const size_t base = (size_t)N * B * b;
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { return src(base + i); },
            [&](int i, cx<T> v) { const int l = i / N, zz = i - l * N; lds[EF::idx(l, zz)] = v; });
    }
    template <typename DST>
    DJ_HD static void z_store_dst(const cx<T> *lds, const DST &dst, int b, int tid) {
        const size_t base = (size_t)N * B * b;
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int l = i / N, zz = i - l * N; return lds[EF::idx(l, zz)]; },
            [&](int i, cx<T> v) { dst(base + i, v); });
    }
    template <typename SRC>
    DJ_HD static void z_load_src_w(const SRC &src, cx<T> *lds, int b, int tid) {
        const size_t base = (size_t)N * B * b;
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63, [&](int i) { return src(base + i); },
            [&](int i, cx<T> v) { const int l = i / N, zz = i - l * N; lds[EF::idx(l, zz)] = v; });
    }
    template <typename DST>
    DJ_HD static void z_store_dst_w(const cx<T> *lds, const DST &dst, int b, int tid) {
        const size_t base = (size_t)N * B * b;
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63, [&](int i) { const int l = i / N, zz = i - l * N; return lds[EF::idx(l, zz)]; },
            [&](int i, cx<T> v) { dst(base + i, v); });
    }
    DJ_HD static void z_store(const cx<T> *lds, cx<T> *Bm, int b, int tid) {
        cx<T> *dst = Bm + (size_t)N * B * b;
        copy_batched<cx<T>, B * N, NT>(
            tid, [&](int i) { const int l = i / N, zz = i - l * N; return lds[EF::idx(l, zz)]; },
            [&](int i, cx<T> v) { dst[i] = v; });
    }

    // ------------------------------------------------------------------ wave-local variants
    // The same element maps as above for the line-contiguous sides of each pass, split so that
    // every wave touches only the lines it transforms (run_wave); tid = 64 wave + lane.
    template <typename IN>
    DJ_HD static void xf_load_w(const IN *mesh, cx<T> *lds, int yb, int z, int tid, double fixed_scale) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_wave<cx<T>, EX::LW * NC>(
            tid >> 6, tid & 63,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                const size_t g = base + (size_t)N * l + 2 * m;
                if constexpr (sizeof(IN) == 8 && !std::is_floating_point<IN>::value)
                    return cx<T>{(T)((double)mesh[g] * fixed_scale), (T)((double)mesh[g + 1] * fixed_scale)};
                else if constexpr (std::is_same<IN, T>::value)
                    return ld_stream((const cx<T> *)(mesh + g));  // g even: a pair of reals
                else
                    return cx<T>{(T)mesh[g], (T)mesh[g + 1]};
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                lds[EX::idx(l, m)] = v;
            });
    }
    template <bool SC = false>
    DJ_HD static void xi_store_w(const cx<T> *lds, T *mesh, int yb, int z, int tid, T s = (T)1) {
        const size_t base = (size_t)N * N * z + (size_t)N * yb * B;
        copy_wave<cx<T>, EX::LW * NC>(
            tid >> 6, tid & 63,
            [&](int i) {
                const int l = i / NC, m = i - l * NC;
                return lds[EX::idx(l, m)];
            },
            [&](int i, cx<T> v) {
                const int l = i / NC, m = i - l * NC;
                const size_t g = base + (size_t)N * l + 2 * m;
                st_stream((cx<T> *)(mesh + g), SC ? cx<T>{v.x * s, v.y * s} : v);  // g even: a pair of reals
            });
    }
    DJ_HD static void y_load_a_w(const cx<T> *A, cx<T> *lds, int zb, int kx, int tid) {
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63,
            [&](int i) { const int l = i / N, y = i - l * N; return A[a_at(y, kx, zb * B + l)]; },
            [&](int i, cx<T> v) { const int l = i / N, y = i - l * N; lds[EF::idx(l, y)] = v; });
    }
    DJ_HD static void y_store_a_w(const cx<T> *lds, cx<T> *A, int zb, int kx, int tid) {
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63, [&](int i) { const int l = i / N, y = i - l * N; return lds[EF::idx(l, y)]; },
            [&](int i, cx<T> v) { const int l = i / N, y = i - l * N; st_stream(&A[a_at(y, kx, zb * B + l)], v); });
    }
    DJ_HD static void z_load_w(const cx<T> *Bm, cx<T> *lds, int b, int tid) {
        const cx<T> *src = Bm + (size_t)N * B * b;
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63, [&](int i) { return src[i]; },
            [&](int i, cx<T> v) { const int l = i / N, zz = i - l * N; lds[EF::idx(l, zz)] = v; });
    }
    DJ_HD static void z_store_w(const cx<T> *lds, cx<T> *Bm, int b, int tid) {
        cx<T> *dst = Bm + (size_t)N * B * b;
        copy_wave<cx<T>, EF::LW * N>(
            tid >> 6, tid & 63, [&](int i) { const int l = i / N, zz = i - l * N; return lds[EF::idx(l, zz)]; },
            [&](int i, cx<T> v) { dst[i] = v; });
    }
    template <int MODE>
    DJ_HD static void z_mult_w(cx<T> *lds, const KMult<T> &K, int b, int tid) {
        constexpr int E = EF::LW * N, CNT = (E + 63) / 64;
        const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
        for (int c = 0; c < CNT; c++) {
            const int il = lane + 64 * c;
            if (!(E % 64 == 0 || il < E)) continue;
            z_mult_one<MODE>(lds, K, b, wave * E + il);
        }
    }

    static constexpr int XBLOCKS() { return (N / B) * N; }
    static constexpr int YBLOCKS() { return (N / B) * H; }
    static constexpr int ZBLOCKS() { return (N * H) / B; }
};

}  // namespace dj
