// Device entry of the FFT passes (fft_passes.h) and their launchers for every supported mesh size.
#include <hip/hip_runtime.h>

#include "dj_internal.h"
#include "fft_passes.h"
#include "ic_modes.h"

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
    // the c2r x pass at 1024^3 (FP32): 16 lines per workgroup and 512 threads, so that its transposed
    // loads read whole 128-byte lines of a block's own (1.99 against 2.12 ms with 8 lines and 256)
    static constexpr int BXI = (sizeof(T) == 4 && N == 1024) ? 16 : B;
    static constexpr int NTXI = BXI == B ? NT : 512;
    using PXI = Passes<T, N, BXI, NTXI>;
};

// Shape of the persistent z pass (k_zm_pipe): 4 waves of one line-group each, 1024 points per wave
// (4096 / N lines per workgroup), about 43 KiB of LDS with the twiddle table, three workgroups per CU
template <typename T, int N> struct ZPipeShape {
    static constexpr bool ok = sizeof(T) == 4 && (N == 256 || N == 512 || N == 1024);
    static constexpr int B = ok ? 4096 / N : 1, NT = 256;
    using P = Passes<T, N, B, NT>;
    static constexpr int GROUPS_PER_CU = 3;
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
    // FP32: the twiddle table in LDS (read by the transforms and the post-processing; through the caches
    // the pass took 1.92 against 1.62 ms at 1024^3, at three workgroups per CU instead of four)
    constexpr bool TWX = sizeof(T) == 4;
    __shared__ cx<T> twl[TWX ? N : 1];
    const cx<T> *t = stage_twiddles<T, N, TWX>(tw, twl, threadIdx.x, 256);
    const int lb = xcd_block(blockIdx.x, gridDim.x), yb = lb % (N / S::B), z = lb / (N / S::B);
    if constexpr (P::EX::wave_ok()) {
        P::xf_load_w(mesh, lds, yb, z, threadIdx.x, fixed_scale);
        if constexpr (TWX) __syncthreads();
        else P::EX::wave_sync();
        P::EX::template run_wave<-1, 2>(lds, t, threadIdx.x);
        __syncthreads();
    } else {
        P::xf_load(mesh, lds, yb, z, threadIdx.x, fixed_scale);
        __syncthreads();
        P::EX::template run<-1, 2>(lds, t, threadIdx.x);
    }
    P::xf_store(lds, t, A, yb, z, threadIdx.x);
}

// The r2c x pass of a real field computed as it is loaded (src(g): the reals g, g + 1 as a complex)
template <typename T, int N, typename SRC>
__global__ __launch_bounds__(256) void k_xf_src(SRC src, cx<T> *__restrict__ A, const cx<T> *__restrict__ tw) {
    using S = FftShape<T, N>;
    using P = typename S::P;
    __shared__ cx<T> lds[P::EX::LDS_ELEMS];
    constexpr bool TWX = sizeof(T) == 4;
    __shared__ cx<T> twl[TWX ? N : 1];
    const cx<T> *t = stage_twiddles<T, N, TWX>(tw, twl, threadIdx.x, 256);
    const int lb = xcd_block(blockIdx.x, gridDim.x), yb = lb % (N / S::B), z = lb / (N / S::B);
    if constexpr (P::EX::wave_ok()) {
        P::xf_load_src_w(src, lds, yb, z, threadIdx.x);
        if constexpr (TWX) __syncthreads();
        else P::EX::wave_sync();
        P::EX::template run_wave<-1, 2>(lds, t, threadIdx.x);
        __syncthreads();
    } else {
        P::xf_load_src(src, lds, yb, z, threadIdx.x);
        __syncthreads();
        P::EX::template run<-1, 2>(lds, t, threadIdx.x);
    }
    P::xf_store(lds, t, A, yb, z, threadIdx.x);
}

// The 2LPT source's partner product of the IC reverse pass, -(mubar a) (neg) or mubar (a + b) (b non-null),
// as a source of k_xf_src: the values k_ic_partner writes, without the field in memory
struct PartnerSrc {
    const double *mub, *a, *b;
    int neg;
    __device__ double at(size_t i) const {
        const double p = b ? a[i] + b[i] : a[i];
        return neg ? -(mub[i] * p) : mub[i] * p;
    }
    __device__ cx<double> operator()(size_t g) const { return cx<double>{at(g), at(g + 1)}; }
};

// SC: the output multiplied by s as it is stored
template <typename T, int N, bool SC = false>
__global__ __launch_bounds__((FftShape<T, N>::NTXI)) void k_xi(const cx<T> *__restrict__ A, T *__restrict__ mesh,
                                            const cx<T> *__restrict__ tw, T s = (T)1) {
    using S = FftShape<T, N>;
    using P = typename S::PXI;
    __shared__ cx<T> lds[P::EX::LDS_ELEMS];
    // FP32: the twiddle table in LDS, staged while the block's loads are in flight (1.99 against 1.56 ms
    // at 1024^3)
    constexpr bool TWX = sizeof(T) == 4;
    __shared__ cx<T> twl[TWX ? N : 1];
    const int lb = xcd_block(blockIdx.x, gridDim.x), yb = lb % (N / S::BXI), z = lb / (N / S::BXI);
    const cx<T> *t = tw;
    if constexpr (TWX) {
        cx<T> va[P::XI_CNT], vb[P::XI_CNT];
        P::xi_fetch(A, yb, z, threadIdx.x, va, vb);
        t = stage_twiddles<T, N, true>(tw, twl, threadIdx.x, S::NTXI);
        __syncthreads();
        P::xi_pre(lds, t, threadIdx.x, va, vb);
    } else {
        P::xi_load_pre(A, lds, tw, yb, z, threadIdx.x);
    }
    __syncthreads();
    if constexpr (P::EX::wave_ok()) {
        P::EX::template run_wave<+1, 2>(lds, t, threadIdx.x);
        P::template xi_store_w<SC>(lds, mesh, yb, z, threadIdx.x, s);
    } else {
        P::EX::template run<+1, 2>(lds, t, threadIdx.x);
        P::template xi_store<SC>(lds, mesh, yb, z, threadIdx.x, s);
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

// The z pass (forward or inverse) of lines read from src(e) and written to dst(e, v), e the B-layout index:
// the FP64 2LPT derivative spectra formed as the pass loads them, or their cotangents cropped and added
// into the n-grid spectrum as it stores them
template <typename T, int N, bool FWD, typename SRC, typename DST>
__global__ __launch_bounds__((FftShape<T, N>::NTY)) void k_zm_io(SRC src, DST dst, const cx<T> *__restrict__ tw) {
    using S = FftShape<T, N>;
    using P = typename S::PY;
    __shared__ cx<T> lds[P::EF::LDS_ELEMS];
    __shared__ cx<T> twl[S::TWL ? N : 1];
    const int b = blockIdx.x;
    const cx<T> *t = stage_twiddles<T, N, S::TWL>(tw, twl, threadIdx.x, S::NTY);
    if constexpr (P::EF::wave_ok()) {
        P::z_load_src_w(src, lds, b, threadIdx.x);
        __syncthreads();  // the twiddle table; from here on every wave works on its own lines
        P::EF::template run_wave<FWD ? -1 : +1, 1>(lds, t, threadIdx.x);
        P::z_store_dst_w(lds, dst, b, threadIdx.x);
    } else {
        P::z_load_src(src, lds, b, threadIdx.x);
        __syncthreads();
        P::EF::template run<FWD ? -1 : +1, 1>(lds, t, threadIdx.x);
        P::z_store_dst(lds, dst, b, threadIdx.x);
    }
}
template <typename T> struct StoreTo {
    cx<T> *out;
    __device__ void operator()(size_t e, cx<T> v) const { out[e] = v; }
};
template <typename T> struct LoadFrom {
    const cx<T> *in;
    __device__ cx<T> operator()(size_t e) const { return in[e]; }
};

// The z pass as a persistent kernel (three workgroups per CU, block b, b + G, b + 2G, ... for G
// workgroups): each wave holds the next block's lines in registers while it transforms and stores
// the current block's, so a block's loads overlap the previous block's transforms and stores
// instead of waiting for another workgroup on the CU.  The multiplier's kz-dependent table values
// are the same for every block and are read once; the per-line ones (kx, ky) are read before the
// next block's loads are issued, so that waiting for them does not wait for those loads.  The
// arithmetic of the multiplier is KMult::kphi's, in its order.  FP32 on 256^3 to 1024^3 meshes (a
// lane's elements keep one kz per step), the potential multiplier without window deconvolution or
// transfer function.
template <typename T, int N, int MODE, bool FWD, bool INV>
__global__ __launch_bounds__((ZPipeShape<T, N>::NT), (ZPipeShape<T, N>::GROUPS_PER_CU)) void k_zm_pipe(
    const cx<T> *in, cx<T> *out, const cx<T> *__restrict__ tw, KMult<T> K, int nblocks) {
    using S = ZPipeShape<T, N>;
    using P = typename S::P;
    using EF = typename P::EF;
    constexpr int B = S::B, H = N / 2 + 1, LW = EF::LW, E = LW * N, CNT = E / 64;
    static_assert(N % 64 == 0 && E % 64 == 0 && EF::wave_ok(), "lanes keep one kz per step");
    __shared__ cx<T> lds[EF::LDS_ELEMS];
    __shared__ cx<T> twl[N];
    const cx<T> *t = stage_twiddles<T, N, true>(tw, twl, threadIdx.x, S::NT);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool lap0 = K.lap_order == 0;
    // element c of a lane: wave line wave LW + (64 c) / N, kz = (64 c) % N + lane
    T zt[CNT];
    if constexpr (MODE == ZM_POT) {
#pragma unroll
        for (int c = 0; c < CNT; c++) {
            const int kz = (64 * c) % N + lane;
            zt[c] = lap0 ? K.k2f[kz] : K.tf[kz];
        }
    }
    cx<T> pf[CNT];
    auto fetch = [&](int blk) {
        const cx<T> *src = in + (size_t)N * B * blk + (size_t)wave * E;
#pragma unroll
        for (int c = 0; c < CNT; c++) pf[c] = src[lane + 64 * c];
    };
    int b = blockIdx.x;
    if (b < nblocks) fetch(b);
    __syncthreads();  // the twiddle table
    while (b < nblocks) {
#pragma unroll
        for (int c = 0; c < CNT; c++) lds[EF::idx(wave * LW + (64 * c) / N, (64 * c) % N + lane)] = pf[c];
        // per-line multiplier parts: th[kx] + tf[ky] (or k2h[kx] + k2f[ky])
        T lt[LW];
        int lkx[LW], lky[LW];
        if constexpr (MODE == ZM_POT) {
#pragma unroll
            for (int w = 0; w < LW; w++) {
                const int L = b * B + wave * LW + w, ky = L / H, kx = L - ky * H;
                lkx[w] = kx;
                lky[w] = ky;
                lt[w] = lap0 ? K.k2h[kx] + K.k2f[ky] : K.th[kx] + K.tf[ky];
            }
        }
        const int bn = b + (int)gridDim.x;
        if (bn < nblocks) fetch(bn);
        EF::wave_sync();
        if constexpr (FWD) EF::template run_wave<-1, 1>(lds, t, threadIdx.x);
        if constexpr (MODE == ZM_POT) {
#pragma unroll
            for (int c = 0; c < CNT; c++) {
                const int w = (64 * c) / N, kz = (64 * c) % N + lane;
                const int i = EF::idx(wave * LW + w, kz);
                T kp;
                if (lkx[w] == 0 && lky[w] == 0 && kz == 0) {
                    kp = (T)0;
                } else {
                    const T il = lap0 ? (T)-1 / (lt[w] + zt[c]) : (T)1 / (lt[w] + zt[c]);
                    kp = il * K.scale;
                }
                lds[i] = scale(lds[i], kp);
            }
            EF::wave_sync();
        }
        if constexpr (INV) EF::template run_wave<+1, 1>(lds, t, threadIdx.x);
        cx<T> *dst = out + (size_t)N * B * b + (size_t)wave * E;
#pragma unroll
        for (int c = 0; c < CNT; c++) dst[lane + 64 * c] = lds[EF::idx(wave * LW + (64 * c) / N, (64 * c) % N + lane)];
        EF::wave_sync();  // the reads above before the next block's writes
        b = bn;
    }
}

// workgroups of the persistent z pass: per_cu on every CU of the current device, a multiple of 8
static int persistent_groups(int per_cu, int nb) {
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        cus = (hipGetDevice(&dev) == hipSuccess &&
               hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
                  ? n
                  : 256;
    }
    int g = ((per_cu * cus) / 8) * 8;
    if (g < 8) g = 8;
    return g < nb ? g : nb;
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
    static void xf_partner(const double *mub, const double *a, const double *b, int neg, void *A, const void *tw,
                           hipStream_t st) {
        if constexpr (sizeof(T) == 8)
            k_xf_src<T, N, PartnerSrc><<<P::XBLOCKS(), 256, 0, st>>>(PartnerSrc{mub, a, b, neg}, (cx<T> *)A,
                                                                     (const cx<T> *)tw);
    }
    // inverse z pass of the A_ik spectrum (ext grid N) formed from phi (F, n grid) as it is loaded, into Be
    static void zm_deriv_pad(const void *F, void *Be, int n, double L, int i, int k, double s, const void *tw,
                             hipStream_t st) {
        if constexpr (sizeof(T) == 8)
            k_zm_io<T, N, false, DerivPadSrc<N>, StoreTo<T>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                DerivPadSrc<N>{(const cx<double> *)F, i, k, L, s}, StoreTo<T>{(cx<T> *)Be}, (const cx<T> *)tw);
    }
    // forward z pass of Be (ext grid N) whose output is cropped, multiplied and added into Fb (n grid)
    static void zm_deriv_adj(const void *Be, void *Fb, int n, double L, int i, int k, double sme, double s,
                             const void *tw, hipStream_t st) {
        if constexpr (sizeof(T) == 8)
            k_zm_io<T, N, true, LoadFrom<T>, DerivAdjDst<N>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                LoadFrom<T>{(const cx<T> *)Be}, DerivAdjDst<N>{(cx<double> *)Fb, i, k, L, sme, s}, (const cx<T> *)tw);
    }
    // FP64 IC spectra on this grid: the inverse z pass of mult<mode>(in) into out, and the forward z pass of
    // in whose output is weighted (s), multiplied by mult<mode> and added into acc
    static void zm_mult_inv(int mode, const void *in, void *out, int c, double L, const void *tw, hipStream_t st) {
        if constexpr (sizeof(T) == 8) {
            const cx<T> *i = (const cx<T> *)in, *t = (const cx<T> *)tw;
            StoreTo<T> o{(cx<T> *)out};
            if (mode == IC_PSI1)
                k_zm_io<T, N, false, MultSrc<IC_PSI1, N>, StoreTo<T>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                    MultSrc<IC_PSI1, N>{i, c, L}, o, t);
            else
                k_zm_io<T, N, false, MultSrc<IC_PSI2, N>, StoreTo<T>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                    MultSrc<IC_PSI2, N>{i, c, L}, o, t);
        }
    }
    static void zm_fwd_mult_acc(int mode, const void *in, void *acc, int c, double L, double s, const void *tw,
                                hipStream_t st) {
        if constexpr (sizeof(T) == 8) {
            const cx<T> *t = (const cx<T> *)tw;
            LoadFrom<T> i{(const cx<T> *)in};
            if (mode == IC_DERIV)
                k_zm_io<T, N, true, LoadFrom<T>, MultAccDst<IC_DERIV, N>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                    i, MultAccDst<IC_DERIV, N>{(cx<T> *)acc, c, L, s}, t);
            else
                k_zm_io<T, N, true, LoadFrom<T>, MultAccDst<IC_PSI2_ADJ, N>><<<P::ZBLOCKS(), S::NTY, 0, st>>>(
                    i, MultAccDst<IC_PSI2_ADJ, N>{(cx<T> *)acc, c, L, s}, t);
        }
    }
    static void xi(const void *A, void *mesh, const void *tw, hipStream_t st) {
        using PX = typename S::PXI;
        k_xi<T, N><<<PX::XBLOCKS(), S::NTXI, 0, st>>>((const cx<T> *)A, (T *)mesh, (const cx<T> *)tw);
    }
    static void xi_scaled(const void *A, void *mesh, const void *tw, double s, hipStream_t st) {
        using PX = typename S::PXI;
        k_xi<T, N, true><<<PX::XBLOCKS(), S::NTXI, 0, st>>>((const cx<T> *)A, (T *)mesh, (const cx<T> *)tw, (T)s);
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
        if constexpr (ZPipeShape<T, N>::ok) {
            if (mode == ZM_POT && fwd && inv && !K.kg && !K.mf) {
                using ZP = ZPipeShape<T, N>;
                const int nb = ZP::P::ZBLOCKS();
                k_zm_pipe<T, N, ZM_POT, true, true><<<persistent_groups(ZP::GROUPS_PER_CU, nb), ZP::NT, 0, st>>>(i, o, t, K, nb);
                return;
            }
        }
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
        ops.xi_scaled = FftN<T, n>::xi_scaled; \
        ops.xf_partner = FftN<T, n>::xf_partner; \
        ops.zm_deriv_pad = FftN<T, n>::zm_deriv_pad; \
        ops.zm_deriv_adj = FftN<T, n>::zm_deriv_adj; \
        ops.zm_mult_inv = FftN<T, n>::zm_mult_inv; \
        ops.zm_fwd_mult_acc = FftN<T, n>::zm_fwd_mult_acc; \
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
