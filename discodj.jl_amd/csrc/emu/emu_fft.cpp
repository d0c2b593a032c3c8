// Host emulation of the FFT pass kernels: runs every workgroup of XF/YF/ZM/YI/XI on the CPU,
// thread by thread and phase by phase (all threads' LDS loads of a Stockham pass before any of
// its stores, as the barriers order them on the device).  It executes exactly the index and
// arithmetic code of fft_passes.h, so the CPU test suite checks the kernels' layouts against
// numpy without a device.  Built by the tests with g++; not part of the product library.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../fft_passes.h"

using namespace dj;

template <class E, int DIR, int TWS, int p = 0, typename T>
static void emu_engine(std::vector<cx<T>> &lds, const cx<T> *tw) {
    if constexpr (p < E::P.n) {
        using PP = typename E::template Pass<p>;
        static constexpr int THREADS = E::THREADS;
        std::vector<cx<T>> regs((size_t)THREADS * PP::C * PP::R);
        for (int tid = 0; tid < THREADS; tid++)
            PP::load(lds.data(), *reinterpret_cast<cx<T>(*)[PP::C][PP::R]>(&regs[(size_t)tid * PP::C * PP::R]), tid);
        for (int tid = 0; tid < THREADS; tid++)
            PP::template compute_store<DIR, TWS>(
                lds.data(), *reinterpret_cast<cx<T>(*)[PP::C][PP::R]>(&regs[(size_t)tid * PP::C * PP::R]), tw, tid);
        emu_engine<E, DIR, TWS, p + 1>(lds, tw);
    }
}

// the wave-local schedule (run_wave): every lane's loads of a pass, then its stores
template <class E, int DIR, int TWS, int p = 0, typename T>
static void emu_engine_wave(std::vector<cx<T>> &lds, const cx<T> *tw) {
    if constexpr (p < E::P.n) {
        using PP = typename E::template WPass<p>;
        static constexpr int THREADS = E::THREADS;
        std::vector<cx<T>> regs((size_t)THREADS * PP::CW * PP::R);
        for (int tid = 0; tid < THREADS; tid++)
            PP::load(lds.data(), *reinterpret_cast<cx<T>(*)[PP::CW][PP::R]>(&regs[(size_t)tid * PP::CW * PP::R]),
                     tid >> 6, tid & 63);
        for (int tid = 0; tid < THREADS; tid++)
            PP::template compute_store<DIR, TWS>(
                lds.data(), *reinterpret_cast<cx<T>(*)[PP::CW][PP::R]>(&regs[(size_t)tid * PP::CW * PP::R]), tw,
                tid >> 6, tid & 63);
        emu_engine_wave<E, DIR, TWS, p + 1>(lds, tw);
    }
}

template <typename T, int N, int B, int NT, bool WAVE = false>
struct Emu {
    using P = Passes<T, N, B, NT>;
    struct EX : P::EX { static constexpr int THREADS = NT; };
    struct EF : P::EF { static constexpr int THREADS = NT; };
    static constexpr int H = N / 2 + 1;
    static_assert(!WAVE || (P::EX::wave_ok() && P::EF::wave_ok()), "wave schedule needs wave-sized passes");

    static void forward(const T *mesh, cx<T> *A, cx<T> *Bm, const cx<T> *tw, bool do_z) {
        std::vector<cx<T>> lds(EX::LDS_ELEMS + EF::LDS_ELEMS);
        for (int z = 0; z < N; z++)
            for (int yb = 0; yb < N / B; yb++) {
                if constexpr (WAVE) {
                    for (int t = 0; t < NT; t++) P::xf_load_w(mesh, lds.data(), yb, z, t, 0.0);
                    emu_engine_wave<EX, -1, 2>(lds, tw);
                } else {
                    for (int t = 0; t < NT; t++) P::xf_load(mesh, lds.data(), yb, z, t, 0.0);
                    emu_engine<EX, -1, 2>(lds, tw);
                }
                for (int t = 0; t < NT; t++) P::xf_store(lds.data(), tw, A, yb, z, t);
            }
        for (int kx = 0; kx < H; kx++)
            for (int zb = 0; zb < N / B; zb++) {
                if constexpr (WAVE) {
                    for (int t = 0; t < NT; t++) P::y_load_a_w(A, lds.data(), zb, kx, t);
                    emu_engine_wave<EF, -1, 1>(lds, tw);
                } else {
                    for (int t = 0; t < NT; t++) P::y_load_a(A, lds.data(), zb, kx, t);
                    emu_engine<EF, -1, 1>(lds, tw);
                }
                for (int t = 0; t < NT; t++) P::y_store_b(lds.data(), Bm, zb, kx, t);
            }
        if (!do_z) return;
        for (int b = 0; b < P::ZBLOCKS(); b++) {
            if constexpr (WAVE) {
                for (int t = 0; t < NT; t++) P::z_load_w(Bm, lds.data(), b, t);
                emu_engine_wave<EF, -1, 1>(lds, tw);
                for (int t = 0; t < NT; t++) P::z_store_w(lds.data(), Bm, b, t);
            } else {
                for (int t = 0; t < NT; t++) P::z_load(Bm, lds.data(), b, t);
                emu_engine<EF, -1, 1>(lds, tw);
                for (int t = 0; t < NT; t++) P::z_store(lds.data(), Bm, b, t);
            }
        }
    }

    template <int MODE>
    static void zm(cx<T> *Bm, const cx<T> *tw, const KMult<T> &K, int fwd, int inv) {
        std::vector<cx<T>> lds(EF::LDS_ELEMS);
        for (int b = 0; b < P::ZBLOCKS(); b++) {
            if constexpr (WAVE) {
                for (int t = 0; t < NT; t++) P::z_load_w(Bm, lds.data(), b, t);
                if (fwd) emu_engine_wave<EF, -1, 1>(lds, tw);
                if (MODE != ZM_NONE)
                    for (int t = 0; t < NT; t++) P::template z_mult_w<MODE>(lds.data(), K, b, t);
                if (inv) emu_engine_wave<EF, +1, 1>(lds, tw);
                for (int t = 0; t < NT; t++) P::z_store_w(lds.data(), Bm, b, t);
            } else {
                for (int t = 0; t < NT; t++) P::z_load(Bm, lds.data(), b, t);
                if (fwd) emu_engine<EF, -1, 1>(lds, tw);
                if (MODE != ZM_NONE)
                    for (int t = 0; t < NT; t++) P::template z_mult<MODE>(lds.data(), K, b, t);
                if (inv) emu_engine<EF, +1, 1>(lds, tw);
                for (int t = 0; t < NT; t++) P::z_store(lds.data(), Bm, b, t);
            }
        }
    }

    static void inverse_yx(cx<T> *Bm, cx<T> *A, T *mesh, const cx<T> *tw) {
        std::vector<cx<T>> lds(EX::LDS_ELEMS + EF::LDS_ELEMS);
        for (int kx = 0; kx < H; kx++)
            for (int zb = 0; zb < N / B; zb++) {
                for (int t = 0; t < NT; t++) P::y_load_b(Bm, lds.data(), zb, kx, t);
                if constexpr (WAVE) {
                    emu_engine_wave<EF, +1, 1>(lds, tw);
                    for (int t = 0; t < NT; t++) P::y_store_a_w(lds.data(), A, zb, kx, t);
                } else {
                    emu_engine<EF, +1, 1>(lds, tw);
                    for (int t = 0; t < NT; t++) P::y_store_a(lds.data(), A, zb, kx, t);
                }
            }
        for (int z = 0; z < N; z++)
            for (int yb = 0; yb < N / B; yb++) {
                for (int t = 0; t < NT; t++) P::xi_load_pre(A, lds.data(), tw, yb, z, t);
                if constexpr (WAVE) {
                    emu_engine_wave<EX, +1, 2>(lds, tw);
                    for (int t = 0; t < NT; t++) P::xi_store_w(lds.data(), mesh, yb, z, t);
                } else {
                    emu_engine<EX, +1, 2>(lds, tw);
                    for (int t = 0; t < NT; t++) P::xi_store(lds.data(), mesh, yb, z, t);
                }
            }
    }
};

template <typename T>
static std::vector<cx<T>> twiddles(int N) {
    std::vector<cx<T>> tw(N);
    for (int k = 0; k < N; k++) {
        double a = -2.0 * M_PI * k / N;
        tw[k] = {(T)std::cos(a), (T)std::sin(a)};
    }
    return tw;
}

// op 0: forward spectrum into B layout; op 1: inverse of a B-layout spectrum into mesh;
// op 2: Poisson solve with the multiplier (K tables from the caller), mesh -> mesh
template <typename T, int N, int B, int NT, bool WAVE = false>
static int run(int op, const void *in, void *out, const void *ktab, int lap_order, double rho, double nsets) {
    using E = Emu<T, N, B, NT, WAVE>;
    constexpr int H = N / 2 + 1;
    auto tw = twiddles<T>(N);
    std::vector<cx<T>> A((size_t)N * H * N), Bm((size_t)N * H * N);
    if (op == 0) {
        E::forward((const T *)in, A.data(), (cx<T> *)out, tw.data(), true);
        return 0;
    }
    if (op == 1) {
        std::memcpy(Bm.data(), in, sizeof(cx<T>) * Bm.size());
        KMult<T> K{};
        E::template zm<ZM_NONE>(Bm.data(), tw.data(), K, 0, 1);
        E::inverse_yx(Bm.data(), A.data(), (T *)out, tw.data());
        return 0;
    }
    // ktab: k2f[N], k2h[H], tf[N], th[H]
    const T *kt = (const T *)ktab;
    KMult<T> K{};
    K.k2f = kt; K.k2h = kt + N; K.tf = kt + N + H; K.th = kt + 2 * N + H;
    K.scale = (T)(1.0 / (rho * nsets * (double)N * N * N)); K.lap_order = lap_order;
    E::forward((const T *)in, A.data(), Bm.data(), tw.data(), false);
    E::template zm<ZM_POT>(Bm.data(), tw.data(), K, 1, 1);
    E::inverse_yx(Bm.data(), A.data(), (T *)out, tw.data());
    return 0;
}

// N = -64: the wave-local schedule (one 64-lane wave per workgroup owning 32 lines of 64 points)
extern "C" int emu_fft(int N, int prec, int op, const void *in, void *out, const void *ktab, int lap_order,
                       double rho, double nsets) {
    if (N == -64)
        return prec == 64 ? run<double, 64, 32, 64, true>(op, in, out, ktab, lap_order, rho, nsets)
                          : run<float, 64, 32, 64, true>(op, in, out, ktab, lap_order, rho, nsets);
#define CASE(n, b)                                                                        \
    case n:                                                                               \
        return prec == 64 ? run<double, n, b, 256>(op, in, out, ktab, lap_order, rho, nsets) \
                          : run<float, n, b, 256>(op, in, out, ktab, lap_order, rho, nsets);
    switch (N) {
        CASE(8, 8)
        CASE(12, 4)
        CASE(16, 16)
        CASE(24, 8)
        CASE(32, 16)
        CASE(48, 16)
        CASE(64, 16)
        CASE(96, 8)
    }
#undef CASE
    return -1;
}
