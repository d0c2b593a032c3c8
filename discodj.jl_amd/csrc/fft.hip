// Device entry of the FFT passes (fft_passes.h) and their launchers for every supported mesh size.
#include <hip/hip_runtime.h>

#include "dj_internal.h"
#include "fft_passes.h"

namespace dj {

template <typename T, int N> struct FftShape {
    // lines per workgroup: about 64 KiB of LDS for the y/z tiles, at most 16, dividing N
    static constexpr int target = 65536 / (N * (int)sizeof(cx<T>));
    static constexpr int B = (N % 16 == 0 && target >= 16) ? 16
                           : (N % 8 == 0 && target >= 8) ? 8
                           : (N % 4 == 0 && target >= 4) ? 4
                           : (N % 2 == 0 && target >= 2) ? 2 : 1;
    // x passes (half-length packed lines, about 32 KiB of LDS) run 256 threads; the y/z passes run
    // 512 once a radix-16 pass has at least that many butterflies, so that the two workgroups their
    // LDS tile allows per CU still hold 16 waves
    static constexpr int NT = 256;
    static constexpr int NTY = (B * N / 16 >= 512) ? 512 : 256;
    using P = Passes<T, N, B, NT>;
    using PY = Passes<T, N, B, NTY>;
    // the N-entry twiddle table moves into LDS when two workgroups of the y/z tile plus table
    // still fit a CU (160 KiB); otherwise the passes read it through the caches
    static constexpr bool TWL = (PY::EF::LDS_ELEMS + N) * sizeof(cx<T>) <= 80 * 1024;
};

// The hardware hands workgroup b to XCD b % 8.  Logical block (b % 8) (nb / 8) + b / 8 puts runs of
// consecutive logical blocks on one XCD, where they run close in time: neighbouring line groups
// whose transposed loads and stores share 128-byte lines then meet in that XCD's L2 instead of
// fetching each line twice from HBM.
__device__ __forceinline__ int xcd_block(int b, int nb) { return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3); }

// the twiddle table for a workgroup: a copy in LDS, or the global table
template <typename T, int N, bool TWL>
__device__ __forceinline__ const cx<T> *stage_twiddles(const cx<T> *__restrict__ tw, cx<T> *twl, int tid, int nt) {
    if constexpr (TWL) {
        for (int i = tid; i < N; i += nt) twl[i] = tw[i];
        return twl;  // the caller's barrier after the tile load covers this copy
    } else {
        return tw;
    }
}

// Each kernel runs its line FFTs wave by wave (run_wave) when the engine's pass shapes allow it,
// with workgroup barriers only around the twiddle table and the transposed (strided) side of a pass;
// otherwise every FFT pass ends with a workgroup barrier (run).
template <typename T, int N, typename IN>
__global__ __launch_bounds__(256) void k_xf(const IN *__restrict__ mesh, cx<T> *__restrict__ A,
                                            const cx<T> *__restrict__ tw, double fixed_scale) {
    using S = FftShape<T, N>;
    using P = typename S::P;
    __shared__ cx<T> lds[P::EX::LDS_ELEMS];
    const int lb = xcd_block(blockIdx.x, gridDim.x), yb = lb % (N / S::B), z = lb / (N / S::B);
    if constexpr (P::EX::wave_ok()) {
        P::xf_load_w(mesh, lds, yb, z, threadIdx.x, fixed_scale);
        P::EX::wave_sync();
        P::EX::template run_wave<-1, 2>(lds, tw, threadIdx.x);
        __syncthreads();
    } else {
        P::xf_load(mesh, lds, yb, z, threadIdx.x, fixed_scale);
        __syncthreads();
        P::EX::template run<-1, 2>(lds, tw, threadIdx.x);
    }
    P::xf_store(lds, tw, A, yb, z, threadIdx.x);
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_xi(const cx<T> *__restrict__ A, T *__restrict__ mesh,
                                            const cx<T> *__restrict__ tw) {
    using S = FftShape<T, N>;
    using P = typename S::P;
    __shared__ cx<T> lds[P::EX::LDS_ELEMS];
    const int lb = xcd_block(blockIdx.x, gridDim.x), yb = lb % (N / S::B), z = lb / (N / S::B);
    P::xi_load_pre(A, lds, tw, yb, z, threadIdx.x);
    __syncthreads();
    if constexpr (P::EX::wave_ok()) {
        P::EX::template run_wave<+1, 2>(lds, tw, threadIdx.x);
        P::xi_store_w(lds, mesh, yb, z, threadIdx.x);
    } else {
        P::EX::template run<+1, 2>(lds, tw, threadIdx.x);
        P::xi_store(lds, mesh, yb, z, threadIdx.x);
    }
}

template <typename T, int N>
__global__ __launch_bounds__((FftShape<T, N>::NTY)) void k_yf(const cx<T> *__restrict__ A, cx<T> *__restrict__ Bm,
                                                             const cx<T> *__restrict__ tw) {
    using S = FftShape<T, N>;
    using P = typename S::PY;
    __shared__ cx<T> lds[P::EF::LDS_ELEMS];
    __shared__ cx<T> twl[S::TWL ? N : 1];
    const int lb = xcd_block(blockIdx.x, gridDim.x), zb = lb % (N / S::B), kx = lb / (N / S::B);
    const cx<T> *t = stage_twiddles<T, N, S::TWL>(tw, twl, threadIdx.x, S::NTY);
    if constexpr (P::EF::wave_ok()) {
        P::y_load_a_w(A, lds, zb, kx, threadIdx.x);
        __syncthreads();
        P::EF::template run_wave<-1, 1>(lds, t, threadIdx.x);
        __syncthreads();
    } else {
        P::y_load_a(A, lds, zb, kx, threadIdx.x);
        __syncthreads();
        P::EF::template run<-1, 1>(lds, t, threadIdx.x);
    }
    P::y_store_b(lds, Bm, zb, kx, threadIdx.x);
}

template <typename T, int N>
__global__ __launch_bounds__((FftShape<T, N>::NTY)) void k_yi(const cx<T> *__restrict__ Bm, cx<T> *__restrict__ A,
                                                             const cx<T> *__restrict__ tw) {
    using S = FftShape<T, N>;
    using P = typename S::PY;
    __shared__ cx<T> lds[P::EF::LDS_ELEMS];
    __shared__ cx<T> twl[S::TWL ? N : 1];
    const int lb = xcd_block(blockIdx.x, gridDim.x), zb = lb % (N / S::B), kx = lb / (N / S::B);
    const cx<T> *t = stage_twiddles<T, N, S::TWL>(tw, twl, threadIdx.x, S::NTY);
    P::y_load_b(Bm, lds, zb, kx, threadIdx.x);
    __syncthreads();
    if constexpr (P::EF::wave_ok()) {
        P::EF::template run_wave<+1, 1>(lds, t, threadIdx.x);
        P::y_store_a_w(lds, A, zb, kx, threadIdx.x);
    } else {
        P::EF::template run<+1, 1>(lds, t, threadIdx.x);
        P::y_store_a(lds, A, zb, kx, threadIdx.x);
    }
}

// z lines of B: optional forward transform, multiplier, optional inverse; in may equal out
template <typename T, int N, int MODE, bool FWD, bool INV>
__global__ __launch_bounds__((FftShape<T, N>::NTY)) void k_zm(const cx<T> *in, cx<T> *out, const cx<T> *__restrict__ tw,
                                                             KMult<T> K) {
    using S = FftShape<T, N>;
    using P = typename S::PY;
    __shared__ cx<T> lds[P::EF::LDS_ELEMS];
    __shared__ cx<T> twl[S::TWL ? N : 1];
    const int b = blockIdx.x;
    const cx<T> *t = stage_twiddles<T, N, S::TWL>(tw, twl, threadIdx.x, S::NTY);
    if constexpr (P::EF::wave_ok()) {
        P::z_load_w(in, lds, b, threadIdx.x);
        __syncthreads();  // the twiddle table; from here on every wave works on its own lines
        if constexpr (FWD) P::EF::template run_wave<-1, 1>(lds, t, threadIdx.x);
        if constexpr (MODE != ZM_NONE) {
            P::template z_mult_w<MODE>(lds, K, b, threadIdx.x);
            P::EF::wave_sync();
        }
        if constexpr (INV) P::EF::template run_wave<+1, 1>(lds, t, threadIdx.x);
        P::z_store_w(lds, out, b, threadIdx.x);
    } else {
        P::z_load(in, lds, b, threadIdx.x);
        __syncthreads();
        if constexpr (FWD) P::EF::template run<-1, 1>(lds, t, threadIdx.x);
        if constexpr (MODE != ZM_NONE) {
            P::template z_mult<MODE>(lds, K, b, threadIdx.x);
            __syncthreads();
        }
        if constexpr (INV) P::EF::template run<+1, 1>(lds, t, threadIdx.x);
        P::z_store(lds, out, b, threadIdx.x);
    }
}

template <typename T, int N>
struct FftN {
    using S = FftShape<T, N>;
    using P = typename S::P;
    static void xf(const void *mesh, bool fixed, void *A, const void *tw, double fscale, hipStream_t st) {
        if (fixed)
            k_xf<T, N, long long><<<P::XBLOCKS(), 256, 0, st>>>((const long long *)mesh, (cx<T> *)A,
                                                               (const cx<T> *)tw, fscale);
        else
            k_xf<T, N, T><<<P::XBLOCKS(), 256, 0, st>>>((const T *)mesh, (cx<T> *)A, (const cx<T> *)tw, fscale);
    }
    static void xi(const void *A, void *mesh, const void *tw, hipStream_t st) {
        k_xi<T, N><<<P::XBLOCKS(), 256, 0, st>>>((const cx<T> *)A, (T *)mesh, (const cx<T> *)tw);
    }
    static void yf(const void *A, void *Bm, const void *tw, hipStream_t st) {
        k_yf<T, N><<<P::YBLOCKS(), S::NTY, 0, st>>>((const cx<T> *)A, (cx<T> *)Bm, (const cx<T> *)tw);
    }
    static void yi(const void *Bm, void *A, const void *tw, hipStream_t st) {
        k_yi<T, N><<<P::YBLOCKS(), S::NTY, 0, st>>>((const cx<T> *)Bm, (cx<T> *)A, (const cx<T> *)tw);
    }
    static void zm(int mode, bool fwd, bool inv, const void *in, void *out, const void *tw, const void *Kp,
                   hipStream_t st) {
        const KMult<T> &K = *(const KMult<T> *)Kp;
        const cx<T> *i = (const cx<T> *)in;
        cx<T> *o = (cx<T> *)out;
        const cx<T> *t = (const cx<T> *)tw;
        dim3 g(P::ZBLOCKS()), b(S::NTY);
#define ZM(M, F, I) k_zm<T, N, M, F, I><<<g, b, 0, st>>>(i, o, t, K)
        if (mode == ZM_POT && fwd && inv) ZM(ZM_POT, true, true);
        else if (mode == ZM_NONE && fwd && !inv) ZM(ZM_NONE, true, false);
        else if (mode == ZM_NONE && !fwd && inv) ZM(ZM_NONE, false, true);
        else if (mode == ZM_NONE && fwd && inv) ZM(ZM_NONE, true, true);
        else if (mode == ZM_GRADX && !fwd && inv) ZM(ZM_GRADX, false, true);
        else if (mode == ZM_GRADY && !fwd && inv) ZM(ZM_GRADY, false, true);
        else if (mode == ZM_GRADZ && !fwd && inv) ZM(ZM_GRADZ, false, true);
#undef ZM
    }
};

template <typename T>
bool fft_dispatch(int N, FftOps &ops) {
#define CASE(n)                         \
    case n:                             \
        ops.xf = FftN<T, n>::xf;        \
        ops.xi = FftN<T, n>::xi;        \
        ops.yf = FftN<T, n>::yf;        \
        ops.yi = FftN<T, n>::yi;        \
        ops.zm = FftN<T, n>::zm;        \
        ops.lines_per_block = FftN<T, n>::S::B; \
        return true;
    switch (N) {
        CASE(16) CASE(24) CASE(32) CASE(48) CASE(64) CASE(96) CASE(128) CASE(192) CASE(256) CASE(384)
        CASE(512) CASE(768) CASE(1024)
    }
#undef CASE
    return false;
}

template bool fft_dispatch<float>(int, FftOps &);
template bool fft_dispatch<double>(int, FftOps &);

}  // namespace dj
