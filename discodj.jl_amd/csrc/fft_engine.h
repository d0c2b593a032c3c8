// Batched line FFT over an LDS tile: B lines of N complex points, NT threads.
//
// Stockham autosort passes (radix 16/8/4/2, then 3): every pass reads each thread's butterfly
// inputs from LDS into registers, synchronises, applies twiddles and the in-register DFT, and
// writes the outputs back in place, so one LDS tile serves all passes and the result is in natural
// order.  Element e of line l lives at l*LS + e + (e >> PS): one pad slot per 2^PS elements breaks
// the power-of-two strides of the early passes across LDS banks.
//
// Each pass is split into load() and compute_store() so the host can emulate a workgroup (all
// loads, then all stores) and check the index arithmetic without a device.
//
// Two schedules.  run(): the butterflies of a pass are spread over all NT threads and a workgroup
// barrier separates the phases.  run_wave(): wave w of the NT / 64 waves owns lines [w LW, (w+1) LW)
// and every pass maps that wave's butterflies onto its own 64 lanes, so the phases only need the
// wave's own ordering (LDS operations of one wave execute in issue order) and no workgroup barrier;
// waves of a workgroup then run their lines independently.  It applies when every pass has a
// multiple of 64 butterflies per wave (wave_ok()).
#pragma once

#include "dj_complex.h"

namespace dj {

struct RadixPlan {
    int n;
    int r[10];
};

// largest radices first; factors of 3 last
constexpr RadixPlan make_plan(int N) {
    RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    int m = N, threes = 0;
    while (m % 3 == 0) { threes++; m /= 3; }
    while (m % 16 == 0 && m > 0) { p.r[p.n++] = 16; m /= 16; }
    if (m == 8) { p.r[p.n++] = 8; m = 1; }
    if (m == 4) { p.r[p.n++] = 4; m = 1; }
    if (m == 2) { p.r[p.n++] = 2; m = 1; }
    for (int i = 0; i < threes; i++) p.r[p.n++] = 3;
    if (m != 1) p.n = -1;  // unsupported length
    return p;
}

constexpr int plan_stride(RadixPlan p, int pass) {
    int s = 1;
    for (int i = 0; i < pass; i++) s *= p.r[i];
    return s;
}

template <typename T, int N, int B, int NT, int PS = 4>
struct LineFFT {
    static constexpr RadixPlan P = make_plan(N);
    static_assert(P.n > 0, "FFT length must be 2^a 3^b");
    // room for N + 1 points per line (the r2c/c2r edge kernels use index N)
    static constexpr int LS = (N + 1) + ((N + 1) >> PS) + 1;
    static constexpr int LDS_ELEMS = B * LS;

    DJ_HD static int idx(int l, int e) { return l * LS + e + (e >> PS); }

    // twiddle exp(DIR 2 pi i t / N) from a table tw[TWS * t] = exp(-2 pi i TWS t / (TWS N))
    template <int DIR, int TWS>
    DJ_HD static cx<T> twiddle(const cx<T> *tw, int t) {
        cx<T> w = tw[TWS * t];
        return DIR < 0 ? w : conj(w);
    }

    template <int p>
    struct Pass {
        static constexpr int R = P.r[p];
        static constexpr int S = plan_stride(P, p);
        static constexpr int M = N / R;
        static constexpr int NBF = B * M;
        static constexpr int C = (NBF + NT - 1) / NT;
        static constexpr int STEP = N / (S * R);

        DJ_HD static void load(const cx<T> *lds, cx<T> (&reg)[C][R], int tid) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                int u = tid + c * NT;
                if (NBF % NT == 0 || u < NBF) {
                    int l = u / M, j = u - l * M;
#pragma unroll
                    for (int r = 0; r < R; r++) reg[c][r] = lds[idx(l, j + r * M)];
                }
            }
        }

        template <int DIR, int TWS>
        DJ_HD static void compute_store(cx<T> *lds, cx<T> (&reg)[C][R], const cx<T> *tw, int tid) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                int u = tid + c * NT;
                if (NBF % NT == 0 || u < NBF) {
                    int l = u / M, j = u - l * M;
                    int k = j % S;
                    if (S > 1 && k != 0) {
#pragma unroll
                        for (int r = 1; r < R; r++) reg[c][r] = cmul(reg[c][r], twiddle<DIR, TWS>(tw, k * r * STEP));
                    }
                    dft<R, DIR>(reg[c]);
                    int o = (j - k) * R + k;
#pragma unroll
                    for (int q = 0; q < R; q++) lds[idx(l, o + q * S)] = reg[c][q];
                }
            }
        }
    };

    // ---------------------------------------------------------------- wave-local schedule
    static constexpr int NW = NT / 64;
    static constexpr int LW = (B % NW == 0) ? B / NW : 0;  // lines per wave

    template <int p>
    struct WPass {
        static constexpr int R = P.r[p];
        static constexpr int S = plan_stride(P, p);
        static constexpr int M = N / R;
        static constexpr int CW = LW * M / 64;  // butterflies per lane
        static constexpr int STEP = N / (S * R);

        DJ_HD static void load(const cx<T> *lds, cx<T> (&reg)[CW][R], int wave, int lane) {
#pragma unroll
            for (int c = 0; c < CW; c++) {
                const int u = lane + 64 * c, l = wave * LW + u / M, j = u % M;
#pragma unroll
                for (int r = 0; r < R; r++) reg[c][r] = lds[idx(l, j + r * M)];
            }
        }

        template <int DIR, int TWS>
        DJ_HD static void compute_store(cx<T> *lds, cx<T> (&reg)[CW][R], const cx<T> *tw, int wave, int lane) {
#pragma unroll
            for (int c = 0; c < CW; c++) {
                const int u = lane + 64 * c, l = wave * LW + u / M, j = u % M;
                const int k = j % S;
                if (S > 1 && k != 0) {
#pragma unroll
                    for (int r = 1; r < R; r++) reg[c][r] = cmul(reg[c][r], twiddle<DIR, TWS>(tw, k * r * STEP));
                }
                dft<R, DIR>(reg[c]);
                const int o = (j - k) * R + k;
#pragma unroll
                for (int q = 0; q < R; q++) lds[idx(l, o + q * S)] = reg[c][q];
            }
        }
    };

    template <int p = 0>
    static constexpr bool wave_ok() {
        if constexpr (p >= P.n) return LW > 0;
        else return LW > 0 && (LW * (N / P.r[p])) % 64 == 0 && wave_ok<p + 1>();
    }

#if defined(__HIPCC__)
    // orders the LDS accesses of the lanes of one wave (compiler fence at wave scope)
    __device__ __forceinline__ static void wave_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // All passes of the wave's lines; the caller has made the wave's lines visible to it (wave_sync
    // or a barrier) and calls wave_sync (or a barrier) before other waves read them.
    template <int DIR, int TWS, int p = 0>
    __device__ __forceinline__ static void run_wave(cx<T> *lds, const cx<T> *tw, int tid) {
        if constexpr (p < P.n) {
            using PP = WPass<p>;
            const int wave = tid >> 6, lane = tid & 63;
            cx<T> reg[PP::CW][PP::R];
            PP::load(lds, reg, wave, lane);
            wave_sync();
            PP::template compute_store<DIR, TWS>(lds, reg, tw, wave, lane);
            wave_sync();
            run_wave<DIR, TWS, p + 1>(lds, tw, tid);
        }
    }

    // All passes on the device; the caller synchronises before (tile loaded) and the tile holds
    // the transform in natural order on return (a barrier ends every pass).
    template <int DIR, int TWS, int p = 0>
    __device__ __forceinline__ static void run(cx<T> *lds, const cx<T> *tw, int tid) {
        if constexpr (p < P.n) {
            using PP = Pass<p>;
            cx<T> reg[PP::C][PP::R];
            PP::load(lds, reg, tid);
            __syncthreads();
            PP::template compute_store<DIR, TWS>(lds, reg, tw, tid);
            __syncthreads();
            run<DIR, TWS, p + 1>(lds, tw, tid);
        }
    }
#endif
};

}  // namespace dj
