// C ABI of the PM N-body step (include/discodj_hip.h): context set-up and orchestration.
//
// One force evaluation (pm_acceleration!, NR/src/nbody/pm.jl:466-521): the particles are binned by
// mesh tile once, then per interlacing shift
//   tile deposit -> XF -> YF -> ZM(K_phi) -> YI -> XI (potential) -> FD-fused tile gather
// For grad_order 0 (ik gradient) the spectrum is kept after the forward z pass and three inverse
// sequences produce -d_a phi, each gathered into one acceleration component (pm.jl:487-492).
// In dj_run the gather of the last shift also applies the kick, the second drift and, unless a
// callback needs the state in between, the next step's first drift.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <initializer_list>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/discodj_hip.h"
#include "dj_internal.h"
#include "fft_passes.h"

using namespace dj;

namespace {

thread_local std::string g_create_error;

enum KernelId { K_ZERO = 0, K_SCATTER, K_XF, K_YF, K_ZM, K_YI, K_XI, K_GATHER, K_GATHER_KICK, K_KICK, K_DRIFT, K_SORT,
                K_SORT_EXACT, K_RESAMPLE, K_NN, K_GATHER_LAST, K_KICK_LISTED, K_COUNT };
const char *kernel_names[K_COUNT] = {"zero", "scatter", "fft_x_r2c", "fft_y_fwd", "fft_z_mult", "fft_y_inv",
                                     "fft_x_c2r", "gather", "gather_kick", "kick_drift", "drift", "bin",
                                     "bin_recount", "resample", "nn_correction", "gather_last", "kick_listed"};

struct ProfEvent {
    int id;
    hipEvent_t a, b;
};

}  // namespace

// tile lists of one point set: counts, list starts, fill cursors, overflow flag, scan block sums, entries
struct TileLists {
    unsigned *cnt = nullptr, *start = nullptr, *gstart = nullptr, *cur = nullptr, *flag = nullptr;
    unsigned long long *part = nullptr;  // block sums of the tile-list scan
    bool lists = false;   // cur holds the counts of a completed sort (capacities for the next one)
    void *ent = nullptr;
    size_t ent_cap = 0;   // entries
};

struct dj_ctx {
    dj_pm_config cfg{};
    int tsize = 4;  // bytes of the working precision
    int dev = 0;
    hipStream_t st = nullptr;
    FftOps fft{};
    PartOps part{};
    int n = 0, res = 0;
    int64_t np = 0, n3 = 0;
    size_t M = 0, HC = 0;
    void *psi = nullptr, *mom = nullptr, *acc = nullptr, *psi_in = nullptr, *q1 = nullptr;
    void *mesh = nullptr, *meshfix = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *tw = nullptr;
    void *tabs = nullptr;
    void *kgain = nullptr;  // fitted transfer function of the Poisson kernel (dj_set_kgain)
    alignas(16) unsigned char kmult[sizeof(KMult<double>)];
    double shifts[16][3];
    int nshifts = 1;
    double fdc[3] = {0, 0, 0};
    TileOps tiles{};
    TileGeom tg{};
    TileLists tl{};   // the particles' lists
    TileLists tlr{};  // the lists of the sheet-resampling point sets (n_resample > 1, allocated on first use)
    void *rl_pts = nullptr, *rl_q0 = nullptr;  // the point sets' positions (3, (2n)^3), a zero lattice (2n)
    double *d_shifts = nullptr;  // (16 + 2, 3): the interlacing shifts, a free slot (+48), a zero shift (+51)
    // sheet resampling (n_resample > 1): particle-grid FFT ops and buffers
    ResampleOps rs{};
    FftOps fftn{};
    void *rs_tw = nullptr;                      // twiddles of the particle grid, exp(-2 pi i k / n)
    void *rs_psi = nullptr;                     // (3, n^3) wrapped psi (linear) or one component (Fourier)
    void *rs_hat = nullptr;                     // (3) half spectra of psi (Fourier)
    void *rs_A = nullptr, *rs_B = nullptr;      // particle-grid FFT work spectra
    void *rs_xr = nullptr;                      // (3, n^3) positions of one offset's point set
    void *rs_phase = nullptr;                   // px (n), py (n), pz (n/2+1) of the current offset
    int n_offsets = 0;
    double offsets[64][3];
    int ioffsets[64][3];  // the offsets in units of 1 / n_resample cells
    // the pass's point sets (the particles and every offset's) deposited together through tile lists
    // (tlr, rl_pts) when their (n_resample n)^3 points index in 31 bits; device atomics otherwise
    bool rs_tiled = false;
    // reverse mode: work buffers (allocated on first use) and the step checkpoints of dj_run_checkpointed
    AdjointOps adj{};
    void *vb_psi1 = nullptr, *vb_abar = nullptr, *vb_psibar = nullptr, *vb_pibar = nullptr;
    void *vb_fbar = nullptr, *vb_pbar = nullptr;
    void *vb_fik = nullptr;  // the three force meshes (spectral gradient) or the packed cells (FD), per reverse force
    bool vb_resident = false;  // vb_psibar / vb_pibar hold cotangents left resident by dj_vjp_deposit / dj_vjp_run
    double *loss_work = nullptr;  // block sums of the density loss (LOSS_PARTS doubles); the vector deposit's scales
    void *vb_pl = nullptr;  // (3, (n_resample n)^3): the cotangent of the resampled points' positions
    void *ckpt = nullptr;  // (steps, 2, 3 n^3): Psi and Pi before each step
    int ck_steps = 0;
    // each checkpointed step's particle tile lists (the sort of its force), kept for the reverse pass when
    // the device has room; ck_tl_ok[s]: slot s holds step s's lists of the last checkpointed run
    std::vector<TileLists> ck_tl;
    std::vector<char> ck_tl_ok;
    std::vector<double> ck_dd1, ck_alpha, ck_beta, ck_dd2;
    // learned correction: FP32 master parameters, Adam moments, BF16 lane-ordered operands, activations
    NnOps nn{};
    bool nn_on = false, nn_train = false;
    double nn_a = 1.0, nn_oscale = 1.0;
    std::vector<double> nn_sched;
    void *nn_p = nullptr, *nn_g = nullptr, *nn_m = nullptr, *nn_v = nullptr, *nn_mask = nullptr;
    void *nn_wf = nullptr, *nn_wb = nullptr;  // (layers, 27, 64) x 8 bytes, forward / transposed
    void *nn_pf = nullptr, *nn_pb = nullptr;  // (layers, 4, 64) x 8 bytes, packed for two-channel inputs
    void *nn_t[10] = {};                      // z1 z2 z3 h1 h2 h3 and four gradient tensors (N^3 x 16 BF16)
    void *nn_dphi = nullptr, *nn_dens = nullptr, *nn_tgt = nullptr, *nn_sum = nullptr;
    void *nn_wpart = nullptr;  // the weight gradient's per-block partial sums (NnOps::wgrad_scratch_bytes)
    void *acc_list = nullptr;  // (entries, 3): the acceleration between interlacing passes, in tile-list order
    unsigned *own_at = nullptr;  // (Np): each particle's entry in its own tile list (written by the bin's fill)
    size_t acc_list_cap = 0;
    // binned spectra: bin sums and device copies of the outputs (allocated on first use)
    double *pk_work = nullptr;
    // transforms on the particle grid (n != res_pm) for the displacement spectrum (allocated on first use)
    FftOps pg_fft{};
    void *pg_tw = nullptr, *pg_A = nullptr, *pg_B = nullptr;
    // the IC pipeline's device buffers (dj_ics_lpt and its reverse), kept between calls
    IcWork icw{};
    std::string err;
    bool prof = false;
    std::vector<ProfEvent> pending;
    std::vector<hipEvent_t> pool;
    int calls[K_COUNT] = {};
    double ms[K_COUNT] = {};
    double bytes[K_COUNT] = {};
};

namespace {

dj_status fail(dj_ctx *c, dj_status code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

dj_status hip_fail(dj_ctx *c, hipError_t e, const char *where) {
    return fail(c, e == hipErrorOutOfMemory ? DJ_ERR_OUT_OF_MEMORY : DJ_ERR_DEVICE,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                   \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

hipEvent_t take_event(dj_ctx *c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct Prof {
    dj_ctx *c;
    int id;
    hipEvent_t a = nullptr;
    Prof(dj_ctx *c_, int id_) : c(c_), id(id_) {
        if (c->prof) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, c->st);
        }
    }
    ~Prof() {
        if (c->prof && a) {
            hipEvent_t b = take_event(c);
            if (b) {
                (void)hipEventRecord(b, c->st);
                c->pending.push_back({id, a, b});
            }
        }
    }
};

// collect finished event pairs (call after a stream synchronisation)
void drain_prof(dj_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->ms[p.id] += ms;
            c->calls[p.id] += 1;
        }
        c->pool.push_back(p.a);
        c->pool.push_back(p.b);
    }
    c->pending.clear();
}

dj_status check_launch(dj_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, what);
    return DJ_OK;
}

// ------------------------------------------------------------------------ host FP64 multiplier tables
// Restates the 1-D k-space kernels of NR/src/nbody/pm.jl:119-183, evaluated in FP64 and rounded to
// the working precision T.
template <typename T>
struct Tables {
    std::vector<T> k2f, k2h, tf, th, mf, mh;
    std::vector<cx<T>> gf, gh;
};

template <typename T>
void build_tables(const dj_pm_config &cfg, Tables<T> &tb) {
    // Every table is evaluated in FP64 and rounded once to T.  pm.jl:140-146 evaluates the
    // finite-difference Laplacian in T; in FP32 its cosine sums cancel to a relative error of about
    // 3 eps / (k dx)^2 at small k (5e-3 at the fundamental mode of a 1024^3 mesh), which shows up as
    // a percent-level error of the large-scale power after a run.  For T = double nothing changes.
    const int N = cfg.res_pm, H = N / 2 + 1;
    const double L = cfg.boxsize;
    std::vector<double> kf(N), kh(H);
    for (int m = 0; m < N; m++) kf[m] = (double)(m < N / 2 ? m : m - N) * 2 * M_PI / L;
    for (int m = 0; m < H; m++) kh[m] = (double)m * 2 * M_PI / L;
    tb.k2f.resize(N);
    tb.k2h.resize(H);
    for (int m = 0; m < N; m++) tb.k2f[m] = (T)(kf[m] * kf[m]);
    for (int m = 0; m < H; m++) tb.k2h[m] = (T)(kh[m] * kh[m]);
    const double mres = (double)N / L;  // max |k| / pi over the full axis (pm.jl:131,141)
    auto lap = [&](const std::vector<double> &k, std::vector<T> &t) {
        t.resize(k.size());
        for (size_t i = 0; i < k.size(); i++) {
            double c1 = std::cos(1 * k[i] / mres), c2 = std::cos(2 * k[i] / mres), c3 = std::cos(3 * k[i] / mres), v;
            if (cfg.lap_order == 2) v = -2 * (1 - c1) * (mres * mres);
            else if (cfg.lap_order == 4) v = -(c2 - 16 * c1 + 15) / 6 * (mres * mres);
            else v = -(-2 * c3 + 27 * c2 - 270 * c1 + 245) / 90 * (mres * mres);
            t[i] = (T)v;
        }
    };
    if (cfg.lap_order != 0) {
        lap(kf, tb.tf);
        lap(kh, tb.th);
    } else {
        tb.tf.assign(N, 0);
        tb.th.assign(H, 0);
    }
    auto mak = [&](const std::vector<double> &k, std::vector<T> &o) {
        o.resize(k.size());
        for (size_t i = 0; i < k.size(); i++) {
            double x = k[i] / (2 * M_PI) / mres;
            double v = x == 0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x);
            o[i] = (T)std::pow(v, -2 * cfg.worder);
        }
    };
    if (cfg.deconvolve) {
        mak(kf, tb.mf);
        mak(kh, tb.mh);
    }
    tb.gf.resize(N);
    tb.gh.resize(H);
    for (int m = 0; m < N; m++) tb.gf[m] = {(T)0, (T)kf[m]};
    for (int m = 0; m < H; m++) tb.gh[m] = {(T)0, (T)kh[m]};
    tb.gf[N / 2] = {0, 0};
    tb.gh[H - 1] = {0, 0};
}

template <typename T>
dj_status upload_tables(dj_ctx *c) {
    Tables<T> tb;
    build_tables<T>(c->cfg, tb);
    const int N = c->res, H = N / 2 + 1;
    size_t nreal = 2 * (size_t)N + 2 * H + 2 * (size_t)N + 2 * H + 2 * (size_t)N + 2 * H;
    std::vector<T> buf(nreal, (T)0);
    size_t o = 0;
    auto put = [&](const std::vector<T> &v) {
        size_t at = o;
        for (size_t i = 0; i < v.size(); i++) buf[o + i] = v[i];
        o += v.size();
        return at;
    };
    size_t ok2f = put(tb.k2f), ok2h = put(tb.k2h), otf = put(tb.tf), oth = put(tb.th);
    size_t omf = o, omh = o;
    if (c->cfg.deconvolve) { omf = put(tb.mf); omh = put(tb.mh); }
    o = (o + 1) & ~(size_t)1;
    size_t ogf = o;
    for (auto &g : tb.gf) { buf[o++] = g.x; buf[o++] = g.y; }
    size_t ogh = o;
    for (auto &g : tb.gh) { buf[o++] = g.x; buf[o++] = g.y; }
    HIPCHK(c, hipMalloc(&c->tabs, sizeof(T) * o));
    HIPCHK(c, hipMemcpy(c->tabs, buf.data(), sizeof(T) * o, hipMemcpyHostToDevice));
    T *d = (T *)c->tabs;
    KMult<T> K{};
    K.k2f = d + ok2f; K.k2h = d + ok2h; K.tf = d + otf; K.th = d + oth;
    K.mf = c->cfg.deconvolve ? d + omf : nullptr;
    K.mh = c->cfg.deconvolve ? d + omh : nullptr;
    K.gf = (const cx<T> *)(d + ogf);
    K.gh = (const cx<T> *)(d + ogh);
    double nn = (double)c->n, nr = c->cfg.n_resample;
    // 1 / (rho nsets N^3) = 1 / (n^3 nsets)
    K.scale = (T)(1.0 / (nn * nn * nn * (nr * nr * nr)));
    K.lap_order = c->cfg.lap_order;
    std::memcpy(c->kmult, &K, sizeof(K));
    // twiddles exp(-2 pi i k / N), k < N, in FP64 rounded to T
    std::vector<cx<T>> tw(N);
    for (int k = 0; k < N; k++) {
        double a = -2.0 * M_PI * (double)k / (double)N;
        tw[k] = {(T)std::cos(a), (T)std::sin(a)};
    }
    HIPCHK(c, hipMalloc(&c->tw, sizeof(cx<T>) * N));
    HIPCHK(c, hipMemcpy(c->tw, tw.data(), sizeof(cx<T>) * N, hipMemcpyHostToDevice));
    // Lagrangian lattice (pm.jl:186)
    std::vector<T> q1(c->n);
    for (int i = 0; i < c->n; i++) q1[i] = (T)((double)i * (c->cfg.boxsize / c->n));
    HIPCHK(c, hipMalloc(&c->q1, sizeof(T) * c->n));
    HIPCHK(c, hipMemcpy(c->q1, q1.data(), sizeof(T) * c->n, hipMemcpyHostToDevice));
    return DJ_OK;
}

// pm.jl:237-246
int aa_shifts(int aa, double dh, double (*s)[3]) {
    if (aa == 0) { s[0][0] = s[0][1] = s[0][2] = 0; return 1; }
    if (aa == -1) { s[0][0] = s[0][1] = s[0][2] = dh; return 1; }
    if (aa == 1) { s[0][0] = s[0][1] = s[0][2] = 0; s[1][0] = s[1][1] = s[1][2] = dh; return 2; }
    if (aa != 2 && aa != 3) return -1;
    int n = 0;
    double v[2] = {0.0, dh};
    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
        s[n][0] = v[a]; s[n][1] = v[b]; s[n][2] = v[c]; n++;
    }
    if (aa == 3) {
        double q = dh / 2, u[2] = {-q, q};
        for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
            s[n][0] = u[a]; s[n][1] = u[b]; s[n][2] = u[c]; n++;
        }
    }
    return n;
}

void free_lists(TileLists &tl);

void free_all(dj_ctx *c) {
    void *bufs[] = {c->psi, c->mom, c->acc, c->psi_in, c->q1, c->mesh, c->meshfix, c->A, c->B, c->C, c->tw, c->tabs,
                    c->d_shifts, c->rl_pts, c->rl_q0,
                    c->rs_tw, c->rs_psi, c->rs_hat, c->rs_A, c->rs_B, c->rs_xr, c->rs_phase,
                    c->vb_psi1, c->vb_abar, c->vb_psibar, c->vb_pibar, c->vb_fbar, c->vb_pbar, c->vb_fik, c->ckpt,
                    c->nn_p, c->nn_g, c->nn_m, c->nn_v, c->nn_mask, c->nn_wf, c->nn_wb, c->nn_pf, c->nn_pb,
                    c->nn_t[0], c->nn_t[1],
                    c->nn_t[2], c->nn_t[3], c->nn_t[4], c->nn_t[5], c->nn_t[6], c->nn_t[7], c->nn_t[8], c->nn_t[9],
                    c->nn_dphi, c->nn_dens, c->nn_tgt, c->nn_sum, c->nn_wpart, c->acc_list, c->kgain, c->pk_work,
                    c->pg_tw, c->pg_A, c->pg_B, c->loss_work, c->vb_pl, c->own_at};
    // teardown: failures here have no caller to report to
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    ic_work_free(c->icw);
    free_lists(c->tl);
    free_lists(c->tlr);
    for (auto &t : c->ck_tl) free_lists(t);
    for (auto &p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
}

void set_bytes(dj_ctx *c) {
    const double T = c->tsize, M = (double)c->M, HC = (double)c->HC, np = (double)c->np;
    // algorithmic bytes per launch: what the operation must read and write at least, with the
    // particle positions counted as 3 values per particle (the tile lists are this library's choice)
    c->bytes[K_ZERO] = 3 * T * np;
    c->bytes[K_SCATTER] = 3 * T * np + M * (c->cfg.deposit == DJ_DEPOSIT_FIXED64 ? 8 : T);
    c->bytes[K_XF] = M * (c->cfg.deposit == DJ_DEPOSIT_FIXED64 ? 8 : T) + 2 * T * HC;
    c->bytes[K_YF] = 4 * T * HC;
    c->bytes[K_ZM] = 4 * T * HC;
    c->bytes[K_YI] = 4 * T * HC;
    c->bytes[K_XI] = 2 * T * HC + T * M;
    // FD gather of a pass that stores the acceleration: read the positions and phi, write acc
    c->bytes[K_GATHER] = 3 * T * np * 2 + T * M;
    // FD gather of the last pass fused with kick and drift: read psi, mom (and acc when averaging
    // shifts), read phi once, write psi and mom
    c->bytes[K_GATHER_KICK] = 3 * T * np * (c->nshifts > 1 ? 5 : 4) + T * M;
    c->bytes[K_KICK] = 3 * T * np * 5;
    // the last pass's gather when interlacing: read the positions, the hand-over and phi, write the hand-over
    c->bytes[K_GATHER_LAST] = 3 * T * np * 3 + T * M;
    // kick and drifts in particle order: read the own-entry index, the summed force, psi and mom; write psi, mom
    c->bytes[K_KICK_LISTED] = (4.0 + 3 * T * 5) * np;
    c->bytes[K_DRIFT] = 3 * T * np * 3;
    // binning: read psi twice (count, fill), write one entry per (particle, tile)
    c->bytes[K_SORT] = 3 * T * np + (double)c->tiles.entry_bytes * np;
    c->bytes[K_SORT_EXACT] = 2 * 3 * T * np + (double)c->tiles.entry_bytes * np;
    // learned correction of one pass: read the density, write and read three 16-channel BF16 hidden
    // tensors (the residual layers read their input twice), write the potential correction
    c->bytes[K_NN] = (4.0 + 32 + 64 + 64 + 32 + 4) * M;
    // sheet resampling of one pass: read psi once, then per offset read and write the mesh once
    c->bytes[K_RESAMPLE] = 3 * T * np + (double)c->n_offsets * 2 * T * M;
}

double fscale_dep(const dj_pm_config &cfg) { return std::ldexp(1.0, cfg.fixed_frac_bits ? cfg.fixed_frac_bits : 32); }

// ------------------------------------------------------------------------ tile binning
// the tile lists tl of a point set: positions psi + q1 on the lattice n^3 (the particles: Psi and q; the
// sheet-resampling point sets: their positions and a zero lattice)
TileArgs tile_args_of(dj_ctx *c, TileLists &tl, const void *psi, const void *q1, int n, const double *d_shifts,
                      int ns) {
    TileArgs a{};
    a.psi = psi;
    a.mom = c->mom;
    a.acc = c->acc;
    a.q1 = q1;
    a.n = n;
    a.L = c->cfg.boxsize;
    a.shifts = d_shifts;
    a.ns = ns;
    a.g = c->tg;
    a.cnt = tl.cnt;
    a.start = tl.start;
    a.gstart = tl.gstart;
    a.cur = tl.cur;
    a.flag = tl.flag;
    a.scan_part = tl.part;
    a.ent = tl.ent;
    a.ent_cap = tl.ent_cap;
    a.own_at = &tl == &c->tl ? c->own_at : nullptr;
    return a;
}
TileArgs tile_args(dj_ctx *c, const void *psi, const double *d_shifts, int ns) {
    return tile_args_of(c, c->tl, psi, c->q1, c->n, d_shifts, ns);
}

dj_status alloc_lists(dj_ctx *c, TileLists &tl, size_t ent_cap) {
    const TileGeom &g = c->tg;
    HIPCHK(c, hipMalloc(&tl.cnt, sizeof(unsigned) * 2 * g.nt));
    HIPCHK(c, hipMalloc(&tl.cur, sizeof(unsigned) * 2 * g.nt));
    HIPCHK(c, hipMalloc(&tl.start, sizeof(unsigned) * (g.nt + 1)));
    HIPCHK(c, hipMalloc(&tl.gstart, sizeof(unsigned) * g.nt));
    HIPCHK(c, hipMalloc(&tl.flag, sizeof(unsigned)));
    HIPCHK(c, hipMalloc(&tl.part, sizeof(unsigned long long) * (g.nt / 1024 + 2)));
    tl.lists = false;
    tl.ent_cap = ent_cap;
    HIPCHK(c, hipMalloc(&tl.ent, tl.ent_cap * c->tiles.entry_bytes));
    return DJ_OK;
}

void free_lists(TileLists &tl) {
    void *b[] = {tl.cnt, tl.start, tl.gstart, tl.cur, tl.flag, tl.part, tl.ent};
    for (void *p : b)
        if (p) (void)hipFree(p);
    tl = TileLists{};
}

// Sort of the particles by mesh tile for the given shifts (device array of ns x 3).  Tile lists are
// reserved from the previous sort's counts with room to double (particles move less than a tile per step), so
// one fill pass writes them; a list that outgrows its reserve (or a full entry buffer) raises a flag,
// and then the sort is repeated exactly: counting pass, lists of the exact sizes, fill.
dj_status tile_sort_of(dj_ctx *c, TileLists &tl, const void *psi, const void *q1, int n, const double *d_shifts,
                       int ns, int W) {
    TileArgs a = tile_args_of(c, tl, psi, q1, n, d_shifts, ns);
    for (int attempt = 0; attempt < 3; attempt++) {
        const bool exact = !tl.lists || attempt > 0;
        {
            Prof p(c, exact ? K_SORT_EXACT : K_SORT);
            if (exact) {
                c->part.zero(tl.cnt, sizeof(unsigned) * 2 * c->tg.nt, c->st);
                c->tiles.count(a, W, c->st);
                c->tiles.scan(a, tl.cnt, 0, c->st);
            } else {
                c->tiles.scan(a, tl.cur, 1, c->st);
            }
            c->tiles.fill(a, W, c->st);
        }
        unsigned hv[2] = {0, 0};  // overflow flag, entries reserved
        HIPCHK(c, hipMemcpyAsync(&hv[0], tl.flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(&hv[1], tl.start + c->tg.nt, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (hv[0] == 0 && hv[1] <= tl.ent_cap) {
            tl.lists = true;
            // After an exact sort hv[1] is the number of entries E; the next sort reserves 2 E + 128 nt.
            // Grow the buffer now (keeping this sort's lists), so that the next fill does not overflow it
            // and repeat the sort.
            const size_t next = 2 * (size_t)hv[1] + 128 * (size_t)c->tg.nt;
            if (exact && next > tl.ent_cap) {
                const size_t cap = next + next / 8 + 1024;
                void *grown = nullptr;
                HIPCHK(c, hipMalloc(&grown, cap * c->tiles.entry_bytes));
                HIPCHK(c, hipMemcpyAsync(grown, tl.ent, (size_t)hv[1] * c->tiles.entry_bytes,
                                         hipMemcpyDeviceToDevice, c->st));
                HIPCHK(c, hipStreamSynchronize(c->st));
                HIPCHK(c, hipFree(tl.ent));
                tl.ent = grown;
                tl.ent_cap = cap;
            }
            return check_launch(c, "tile binning");
        }
        tl.lists = false;
        if (hv[1] > tl.ent_cap) {
            // grow the entry buffer (exact sizes follow in the next attempt)
            if (tl.ent) HIPCHK(c, hipFree(tl.ent));
            tl.ent = nullptr;
            const size_t cap = (size_t)hv[1] + hv[1] / 8 + 1024;
            HIPCHK(c, hipMalloc(&tl.ent, cap * c->tiles.entry_bytes));
            tl.ent_cap = cap;
            a.ent = tl.ent;
            a.ent_cap = tl.ent_cap;
        }
    }
    return fail(c, DJ_ERR_DEVICE, "tile binning: lists did not fit after an exact recount");
}

dj_status tile_sort(dj_ctx *c, const void *psi, const double *d_shifts, int ns, int W) {
    return tile_sort_of(c, c->tl, psi, c->q1, c->n, d_shifts, ns, W);
}

// ------------------------------------------------------------------------ one force evaluation
enum { FORCE_ACC = 0, FORCE_KICK = 1 };

template <typename T>
dj_status resample_deposit(dj_ctx *c, const void *psi, const double *sh, void *dens, bool fixed, double fscale,
                           void *pts = nullptr);
void nn_forward_dev(dj_ctx *c, const void *dens, double a);

dj_status force(dj_ctx *c, const void *psi, int mode, const KickArgs &kick) {
    const dj_pm_config &cfg = c->cfg;
    const bool fixed = cfg.deposit == DJ_DEPOSIT_FIXED64;
    const int ns = c->nshifts;
    const size_t mesh_bytes = c->M * (fixed ? 8 : c->tsize);
    void *dens = fixed ? c->meshfix : c->mesh;
    const double fscale = std::ldexp(1.0, -(cfg.fixed_frac_bits ? cfg.fixed_frac_bits : 32));
    PartArgs pa{};
    pa.psi = psi;
    pa.mom = c->mom;
    pa.acc = c->acc;
    pa.q1 = c->q1;
    pa.n = c->n;
    pa.res = c->res;
    pa.L = cfg.boxsize;
    const bool ik = cfg.grad_order == 0;
    // the ik path accumulates the components into acc; the FD path assigns it in its first pass
    if (ik) {
        Prof p(c, K_ZERO);
        c->part.zero(c->acc, (size_t)c->n3 * c->tsize, c->st);
    }
    // the point sets of sheet resampling deposited together through their own tile lists (the particles'
    // lists then serve only a finite-difference gather)
    const bool rs_tiled = cfg.n_resample > 1 && c->rs_tiled;
    // with interlacing the last pass adds its force into the list-ordered hand-over and the kick runs in
    // particle order (k_kick_listed), reading the sum through the own-list entry the bin records
    const bool listed_kick = mode == FORCE_KICK && ns > 1 && !ik;
    if (listed_kick && !c->own_at) HIPCHK(c, hipMalloc(&c->own_at, sizeof(unsigned) * (size_t)c->np));
    if (!(rs_tiled && ik)) {
        dj_status rs = tile_sort(c, psi, c->d_shifts, ns, cfg.worder);
        if (rs) return rs;
    }
    (void)mesh_bytes;
    // in a step the passes hand the acceleration over in tile-list order: both walk the same lists, so
    // the hand-over is read and written contiguously instead of at the particles' positions
    const bool lacc = mode == FORCE_KICK && ns > 1 && !ik;
    if (lacc && c->acc_list_cap < c->tl.ent_cap) {
        if (c->acc_list) HIPCHK(c, hipFree(c->acc_list));
        c->acc_list = nullptr;
        HIPCHK(c, hipMalloc(&c->acc_list, 3 * c->tl.ent_cap * c->tsize));
        c->acc_list_cap = c->tl.ent_cap;
    }
    for (int s = 0; s < ns; s++) {
        pa.shift[0] = c->shifts[s][0];
        pa.shift[1] = c->shifts[s][1];
        pa.shift[2] = c->shifts[s][2];
        TileArgs ta = tile_args(c, psi, c->d_shifts, ns);
        if (lacc) {
            ta.acc = c->acc_list;
            ta.acc_by_entry = 1;
        }
        if (rs_tiled) {
            const int nrn = cfg.n_resample * c->n;
            {
                Prof p(c, K_RESAMPLE);
                c->rs.shifted(psi, c->q1, c->n, c->shifts[s], cfg.n_resample, c->rl_pts, c->st);
                dj_status rr = c->tsize == 8
                                   ? resample_deposit<double>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg), c->rl_pts)
                                   : resample_deposit<float>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg), c->rl_pts);
                if (rr) return rr;
                rr = tile_sort_of(c, c->tlr, c->rl_pts, c->rl_q0, nrn, c->d_shifts + 51, 1, cfg.worder);
                if (rr) return rr;
            }
            const double zero3[3] = {0, 0, 0};
            TileArgs tr = tile_args_of(c, c->tlr, c->rl_pts, c->rl_q0, nrn, c->d_shifts + 51, 1);
            Prof p(c, K_SCATTER);
            c->tiles.deposit(tr, cfg.worder, fixed, zero3, dens, fscale_dep(cfg), c->st);
        } else {
            { Prof p(c, K_SCATTER); c->tiles.deposit(ta, cfg.worder, fixed, c->shifts[s], dens, fscale_dep(cfg), c->st); }
            if (cfg.n_resample > 1) {
                Prof p(c, K_RESAMPLE);
                dj_status rr = c->tsize == 8 ? resample_deposit<double>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg))
                                             : resample_deposit<float>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg));
                if (rr) return rr;
            }
        }
        if (c->nn_on) {
            // the learned potential correction of this pass, from the density before the transform
            Prof p(c, K_NN);
            if (c->nn_train) HIPCHK(c, hipMemcpyAsync(c->nn_dens, dens, c->M * 4, hipMemcpyDeviceToDevice, c->st));
            nn_forward_dev(c, dens, c->nn_a);
        }
        { Prof p(c, K_XF); c->fft.xf(dens, fixed, c->A, c->tw, fscale, c->st); }
        { Prof p(c, K_YF); c->fft.yf(c->A, c->B, c->tw, c->st); }
        if (!ik) {
            { Prof p(c, K_ZM); c->fft.zm(ZM_POT, true, true, c->B, c->B, c->tw, c->kmult, c->st); }
            { Prof p(c, K_YI); c->fft.yi(c->B, c->A, c->tw, c->st); }
            { Prof p(c, K_XI); c->fft.xi(c->A, c->mesh, c->tw, c->st); }
            if (c->nn_on) c->nn.axpy(c->mesh, c->nn_dphi, c->M, 1.0, c->st);
            const bool last = s == ns - 1;
            // first pass stores acc, later passes add to it; the last one of a step kicks and drifts (fused
            // into the gather without interlacing, in particle order after it with)
            const int gmode = (mode == FORCE_KICK && last && !listed_kick) ? GATHER_KICK
                                                                           : (s == 0 ? GATHER_ASSIGN : GATHER_ACC);
            {
                Prof p(c, gmode == GATHER_KICK ? K_GATHER_KICK : (listed_kick && last ? K_GATHER_LAST : K_GATHER));
                c->tiles.gather(ta, cfg.worder, cfg.grad_order, c->shifts[s], c->mesh, c->fdc, gmode, kick, c->psi,
                                c->st);
            }
            if (listed_kick && last) {
                Prof p(c, K_KICK_LISTED);
                c->tiles.kick_listed(c->own_at, c->acc_list, c->psi, c->mom, (size_t)c->np, kick, c->st);
            }
        } else {
            { Prof p(c, K_ZM); c->fft.zm(ZM_NONE, true, false, c->B, c->B, c->tw, c->kmult, c->st); }
            for (int ax = 0; ax < 3; ax++) {
                { Prof p(c, K_ZM); c->fft.zm(ZM_GRADX + ax, false, true, c->B, c->C, c->tw, c->kmult, c->st); }
                { Prof p(c, K_YI); c->fft.yi(c->C, c->A, c->tw, c->st); }
                { Prof p(c, K_XI); c->fft.xi(c->A, c->mesh, c->tw, c->st); }
                Prof p(c, K_GATHER);
                c->part.gather_comp(pa, cfg.worder, c->mesh, ax, c->st);
            }
        }
        dj_status r = check_launch(c, "force pass");
        if (r) return r;
    }
    if (ik && mode == FORCE_KICK) {
        Prof p(c, K_KICK);
        c->part.kick_drift(c->psi, c->mom, c->acc, c->n3, kick, c->st);
    } else if (mode == FORCE_ACC && ns > 1) {
        c->part.scale(c->acc, c->n3, 1.0 / ns, c->st);
    }
    return check_launch(c, "force");
}

template <typename T>
dj_status setup_ops(dj_ctx *c) {
    if (!fft_dispatch<T>(c->res, c->fft))
        return fail(c, DJ_ERR_UNSUPPORTED, "res_pm " + std::to_string(c->res) +
                                               " is not a built mesh size (16..1024, 2^a 3^b)");
    part_ops<T>(c->part);
    tile_ops<T>(c->tiles);
    resample_ops<T>(c->rs);
    adjoint_ops<T>(c->adj);
    nn_ops(c->nn);
    return upload_tables<T>(c);
}

// Buffers of sheet resampling and the offsets d = (a, b, c) / n_resample, (0,0,0) excluded, in the
// order of pm.jl:233-234.  Fourier resampling transforms on the particle grid with the library's
// FFT passes (n must be a built FFT size).
template <typename T>
dj_status setup_resampling(dj_ctx *c) {
    const int nr = c->cfg.n_resample, n = c->n;
    c->n_offsets = 0;
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nr; b++)
            for (int d = 0; d < nr; d++) {
                if (a == 0 && b == 0 && d == 0) continue;
                c->offsets[c->n_offsets][0] = (double)a / nr;
                c->offsets[c->n_offsets][1] = (double)b / nr;
                c->offsets[c->n_offsets][2] = (double)d / nr;
                c->ioffsets[c->n_offsets][0] = a;
                c->ioffsets[c->n_offsets][1] = b;
                c->ioffsets[c->n_offsets][2] = d;
                c->n_offsets++;
            }
    const size_t n3 = (size_t)c->np, hc = (size_t)n * n * (n / 2 + 1);
    const double npts = std::pow((double)nr * n, 3);
    c->rs_tiled = npts < 2147483648.0;
    if (c->rs_tiled) {
        const size_t NR = (size_t)nr * n;
        HIPCHK(c, hipMalloc(&c->rl_pts, 3 * NR * NR * NR * sizeof(T)));
        HIPCHK(c, hipMalloc(&c->rl_q0, NR * sizeof(T)));
        HIPCHK(c, hipMemset(c->rl_q0, 0, NR * sizeof(T)));
        dj_status ra = alloc_lists(c, c->tlr, (size_t)(2.5 * npts) + 1024);
        if (ra) return ra;
    }
    HIPCHK(c, hipMalloc(&c->rs_xr, 3 * n3 * sizeof(T)));
    if (c->cfg.resampling == 1) {
        HIPCHK(c, hipMalloc(&c->rs_psi, 4 * n3 * sizeof(T)));  // 3 wrapped components + 1 shifted
        return DJ_OK;
    }
    if (!fft_dispatch<T>(n, c->fftn))
        return fail(c, DJ_ERR_UNSUPPORTED, "Fourier sheet resampling needs n_part in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024}");
    std::vector<cx<T>> tw(n);
    for (int k = 0; k < n; k++) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = {(T)std::cos(a), (T)std::sin(a)};
    }
    HIPCHK(c, hipMalloc(&c->rs_tw, sizeof(cx<T>) * n));
    HIPCHK(c, hipMemcpy(c->rs_tw, tw.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->rs_psi, n3 * sizeof(T)));
    HIPCHK(c, hipMalloc(&c->rs_hat, 3 * hc * sizeof(cx<T>)));
    HIPCHK(c, hipMalloc(&c->rs_A, hc * sizeof(cx<T>)));
    HIPCHK(c, hipMalloc(&c->rs_B, hc * sizeof(cx<T>)));
    // phase factors of every offset, px (n), py (n), pz (n/2+1) each, as pm.jl:452-454 builds them in T:
    // exp(i (T(2 pi) m) d) with m = (m < n/2 ? m : m - n) / n on the full axes, m / n on the half axis
    const int H = n / 2 + 1, per = 2 * n + H;
    std::vector<cx<T>> ph((size_t)per * c->n_offsets);
    const T tp = (T)(2 * M_PI);
    for (int o = 0; o < c->n_offsets; o++) {
        cx<T> *p = ph.data() + (size_t)o * per;
        for (int a = 0; a < 2; a++)
            for (int m = 0; m < n; m++) {
                const T mf = (T)((double)(m < n / 2 ? m : m - n) / n), th = tp * mf * (T)c->offsets[o][a];
                p[a * n + m] = {(T)std::cos(th), (T)std::sin(th)};
            }
        for (int m = 0; m < H; m++) {
            const T mh = (T)((double)m / n), th = tp * mh * (T)c->offsets[o][2];
            p[2 * n + m] = {(T)std::cos(th), (T)std::sin(th)};
        }
    }
    HIPCHK(c, hipMalloc(&c->rs_phase, ph.size() * sizeof(cx<T>)));
    HIPCHK(c, hipMemcpy(c->rs_phase, ph.data(), ph.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    return DJ_OK;
}

// The extra point sets of one interlacing pass (shift sh) deposited into dens (pm.jl:471-483 with
// _resample_prepare! / _resample_offset! for Fourier, _spawn for linear).
// pts: instead of depositing, write each offset's positions into the point-set lattice pts (3, (nr n)^3)
template <typename T>
dj_status resample_deposit(dj_ctx *c, const void *psi, const double *sh, void *dens, bool fixed, double fscale,
                           void *pts) {
    const int nr = pts ? c->cfg.n_resample : 1, o0[3] = {0, 0, 0};
    const int n = c->n, H = n / 2 + 1;
    const double L = c->cfg.boxsize;
    const size_t n3 = (size_t)c->np, hc = (size_t)n * n * H;
    T *xr = (T *)c->rs_xr;
    if (c->cfg.resampling == 1) {
        T *w = (T *)c->rs_psi;
        for (int comp = 0; comp < 3; comp++)
            c->rs.wrap(psi, c->q1, n, comp, sh[comp], L, false, w + comp * n3, c->st);
        for (int o = 0; o < c->n_offsets; o++) {
            const double *d = c->offsets[o];
            for (int comp = 0; comp < 3; comp++) {
                c->rs.trilinear(w + comp * n3, n, d, w + 3 * n3, c->st);
                c->rs.place(w + 3 * n3, c->q1, n, comp, (double)(T)(d[comp] * L / n), L, false, pts ? pts : xr, nr,
                            pts ? c->ioffsets[o] : o0, c->st);
            }
            if (!pts) c->rs.scatter(xr, n, c->cfg.worder, dens, c->res, L, fixed, fscale, c->st);
        }
        return check_launch(c, "linear sheet resampling");
    }
    cx<T> *hat = (cx<T> *)c->rs_hat;
    for (int comp = 0; comp < 3; comp++) {
        c->rs.wrap(psi, c->q1, n, comp, sh[comp], L, true, c->rs_psi, c->st);
        c->fftn.xf(c->rs_psi, false, c->rs_A, c->rs_tw, 1.0, c->st);
        c->fftn.yf(c->rs_A, c->rs_B, c->rs_tw, c->st);
        c->fftn.zm(ZM_NONE, true, false, c->rs_B, hat + comp * hc, c->rs_tw, c->kmult, c->st);
    }
    const double scale = (double)((T)1 / ((T)n * (T)n * (T)n));
    for (int o = 0; o < c->n_offsets; o++) {
        const double *d = c->offsets[o];
        const cx<T> *px = (const cx<T> *)c->rs_phase + (size_t)o * (2 * n + H), *py = px + n, *pz = px + 2 * n;
        for (int comp = 0; comp < 3; comp++) {
            c->rs.phase(hat + comp * hc, c->rs_B, n, px, py, pz, scale, c->st);
            c->fftn.zm(ZM_NONE, false, true, c->rs_B, c->rs_B, c->rs_tw, c->kmult, c->st);
            c->fftn.yi(c->rs_B, c->rs_A, c->rs_tw, c->st);
            c->fftn.xi(c->rs_A, c->rs_psi, c->rs_tw, c->st);
            c->rs.place(c->rs_psi, c->q1, n, comp, (double)(T)(d[comp] * L / n), L, true, pts ? pts : xr, nr,
                        pts ? c->ioffsets[o] : o0, c->st);
        }
        if (!pts) c->rs.scatter(xr, n, c->cfg.worder, dens, c->res, L, fixed, fscale, c->st);
    }
    return check_launch(c, "Fourier sheet resampling");
}

// Reverse of the resampled points' deposit of one interlacing pass: psibar += (d density / d psi)^T rhobar
// through the points of every offset (the particles' own term is the force adjoint's).  The point lattice
// rl_pts holds the pass's points (the density of the same pass placed them); the gather of grad W rhobar
// at them gives their cotangents, and the transposed interpolation (trilinear, or the conjugate Fourier
// phase) takes those back to Psi.
template <typename T>
dj_status resample_vjp(dj_ctx *c, const void *rhobar, void *psibar) {
    const int nr = c->cfg.n_resample, n = c->n, H = n / 2 + 1, N = c->res;
    const size_t n3 = (size_t)c->np, NR = (size_t)nr * n, NR3 = NR * NR * NR;
    const double L = c->cfg.boxsize;
    if (!c->vb_pl) HIPCHK(c, hipMalloc(&c->vb_pl, 3 * NR3 * sizeof(T)));
    c->part.zero(c->vb_pl, 3 * NR3 * sizeof(T), c->st);
    const double zero3[3] = {0, 0, 0};
    c->adj.gather_grad(c->rl_pts, c->rl_q0, (int)NR, N, L, zero3, rhobar, nullptr, nullptr, 0.0, c->fdc, c->cfg.worder,
                       c->vb_pl, c->st);
    T *pb = (T *)psibar;
    if (c->cfg.resampling == 1) {
        for (int o = 0; o < c->n_offsets; o++)
            for (int comp = 0; comp < 3; comp++)
                c->rs.trilinear_adj(c->vb_pl, n, nr, c->ioffsets[o], comp, c->offsets[o], pb + comp * n3, c->st);
        return check_launch(c, "linear sheet resampling adjoint");
    }
    cx<T> *hb = (cx<T> *)c->rs_hat;  // the accumulated cotangent spectrum of one component
    const double scale = (double)((T)1 / ((T)n * (T)n * (T)n));
    for (int comp = 0; comp < 3; comp++) {
        for (int o = 0; o < c->n_offsets; o++) {
            const cx<T> *px = (const cx<T> *)c->rs_phase + (size_t)o * (2 * n + H), *py = px + n, *pz = px + 2 * n;
            c->rs.extract(c->vb_pl, n, nr, c->ioffsets[o], comp, true, c->rs_psi, c->st);
            c->fftn.xf(c->rs_psi, false, c->rs_A, c->rs_tw, 1.0, c->st);
            c->fftn.yf(c->rs_A, c->rs_B, c->rs_tw, c->st);
            c->fftn.zm(ZM_NONE, true, false, c->rs_B, c->rs_B, c->rs_tw, c->kmult, c->st);
            c->rs.phase_adj(c->rs_B, hb, n, px, py, pz, scale, o > 0, c->st);
        }
        c->fftn.zm(ZM_NONE, false, true, hb, hb, c->rs_tw, c->kmult, c->st);
        c->fftn.yi(hb, c->rs_A, c->rs_tw, c->st);
        c->fftn.xi(c->rs_A, c->rs_psi, c->rs_tw, c->st);
        c->rs.add_transposed(c->rs_psi, n, pb + comp * n3, c->st);
    }
    return check_launch(c, "Fourier sheet resampling adjoint");
}

// tile shape (FP32: 32 x 16 x 16 cells, FP64: 16^3) and the stencil reach over the shifts
dj_status setup_tiles(dj_ctx *c) {
    const int N = c->res;
    auto pick = [N](std::initializer_list<int> cand) {
        for (int t : cand)
            if (t <= N && N % t == 0) return t;
        return N;
    };
    TileGeom &g = c->tg;
    g.N = N;
    // FP32: 32 x 16 x 8 cells, so that the gather's staged region (tile + reach + FD halo, 79 KB for
    // TSC at order 6) lets two workgroups share a CU: one stages while the other computes
    if (c->tsize == 4) { g.TX = pick({32, 16, 8}); g.TY = pick({16, 8}); g.TZ = pick({8}); }
    else { g.TX = pick({16, 8}); g.TY = pick({16, 8}); g.TZ = pick({16, 8}); }
    if ((g.TX & (g.TX - 1)) || (g.TY & (g.TY - 1)) || (g.TZ & (g.TZ - 1)))
        return fail(c, DJ_ERR_UNSUPPORTED, "mesh size needs power-of-two tiles (multiple of 8)");
    g.ntx = N / g.TX; g.nty = N / g.TY; g.ntz = N / g.TZ;
    auto lg = [](int v) { int l = 0; while ((1 << l) < v) l++; return l; };
    g.lT[0] = lg(g.TX); g.lT[1] = lg(g.TY); g.lT[2] = lg(g.TZ);
    g.nt = g.ntx * g.nty * g.ntz;
    // cells of the stencil relative to floor(g): [floor(s_min) + lo, floor(s_max) + (W odd ? 2 : 1) + hi]
    const int W = c->cfg.worder, lo = W == 2 ? 0 : -1, hi = lo + W - 1;
    double smin = 0, smax = 0;
    for (int s = 0; s < c->nshifts; s++)
        for (int a = 0; a < 3; a++) {
            double sig = c->shifts[s][a] * N / c->cfg.boxsize;
            smin = std::min(smin, sig);
            smax = std::max(smax, sig);
        }
    g.RL = -((int)std::floor(smin) + lo);
    g.RR = (int)std::floor(smax) + (W % 2 ? 2 : 1) + hi;
    // The gather stages the reach that exact arithmetic allows: base cells of g in [c0 + s, c0 + 1 + s)
    // are at most c0 + ceil(1.5 + s) - 1 (odd W, round) or c0 + ceil(1 + s) - 1 (even W, floor).
    // A particle that rounding pushes one cell further is gathered from global memory.
    int rs = 0;
    for (int s = 0; s < c->nshifts; s++)
        for (int a = 0; a < 3; a++) {
            const double sig = c->shifts[s][a] * N / c->cfg.boxsize;
            rs = std::max(rs, (int)std::ceil((W % 2 ? 1.5 : 1.0) + sig) - 1);
        }
    g.RS = std::min(g.RR, rs + hi);
    if (g.nt > (1 << 20)) return fail(c, DJ_ERR_UNSUPPORTED, "more than 2^20 mesh tiles");
    HIPCHK(c, hipMalloc(&c->d_shifts, sizeof(double) * 18 * 3));
    double sh18[18][3] = {};
    std::memcpy(sh18, c->shifts, sizeof(double) * 16 * 3);
    HIPCHK(c, hipMemcpy(c->d_shifts, sh18, sizeof(double) * 18 * 3, hipMemcpyHostToDevice));
    // own + ghost entries: about 2.1 per particle for 32 x 16 x 8 tiles with TSC and interlacing, twice
    // that plus 64 per list once the lists are reserved from a previous sort (grown on demand)
    return alloc_lists(c, c->tl, (size_t)(2.5 * (double)c->np) + 1024);
}

dj_status validate(const dj_pm_config &cfg, std::string &msg) {
    if (cfg.worder < 2 || cfg.worder > 4) { msg = "worder must be 2, 3 or 4"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.grad_order != 0 && cfg.grad_order != 2 && cfg.grad_order != 4 && cfg.grad_order != 6) {
        msg = "grad_order must be 0, 2, 4 or 6"; return DJ_ERR_INVALID_CONFIG;
    }
    if (cfg.lap_order != 0 && cfg.lap_order != 2 && cfg.lap_order != 4 && cfg.lap_order != 6) {
        msg = "lap_order must be 0, 2, 4 or 6"; return DJ_ERR_INVALID_CONFIG;
    }
    if (cfg.antialias < -1 || cfg.antialias > 3) { msg = "antialias must be -1, 0, 1, 2 or 3"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.n_part < 1 || cfg.res_pm < 2 || !(cfg.boxsize > 0)) { msg = "n_part, res_pm, boxsize must be positive"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.resampling != 0 && cfg.resampling != 1) { msg = "resampling must be fourier (0) or linear (1)"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.n_resample < 1) { msg = "n_resample must be >= 1"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.lean && !(cfg.antialias == 0 && cfg.n_resample == 1 && cfg.grad_order != 0)) {
        msg = "PMSolverLean supports antialias = 0, n_resample = 1 and finite-difference gradients (2/4/6)";
        return DJ_ERR_INVALID_CONFIG;
    }
    if (cfg.precision != 32 && cfg.precision != 64) { msg = "precision must be 32 or 64"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.deposit != DJ_DEPOSIT_FLOAT && cfg.deposit != DJ_DEPOSIT_FIXED64) { msg = "unknown deposit mode"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.fixed_frac_bits < 0 || cfg.fixed_frac_bits > 40) { msg = "fixed_frac_bits must be in 0..40"; return DJ_ERR_INVALID_CONFIG; }
    if (cfg.n_resample > 4) { msg = "n_resample above 4 (63 offsets) is not built into this library"; return DJ_ERR_UNSUPPORTED; }
    if (cfg.n_part > 1024) { msg = "n_part above 1024 is not supported"; return DJ_ERR_UNSUPPORTED; }
    return DJ_OK;
}

dj_status upload_particles(dj_ctx *c, void *dst, const void *src, hipMemcpyKind kind) {
    size_t bytes = (size_t)c->n3 * c->tsize;
    if (src) HIPCHK(c, hipMemcpyAsync(dst, src, bytes, kind, c->st));
    else c->part.zero(dst, bytes, c->st);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return DJ_OK;
}

}  // namespace

extern "C" {

const char *dj_version(void) { return "discodj_hip 0.1 (gfx950)"; }

dj_status dj_create(const dj_pm_config *cfg, dj_ctx **out) {
    if (!cfg || !out) return fail(nullptr, DJ_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::string msg;
    dj_status v = validate(*cfg, msg);
    if (v) return fail(nullptr, v, msg);
    dj_ctx *c = new (std::nothrow) dj_ctx();
    if (!c) return fail(nullptr, DJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->cfg = *cfg;
    c->tsize = cfg->precision == 64 ? 8 : 4;
    c->dev = cfg->device;
    c->n = cfg->n_part;
    c->res = cfg->res_pm;
    c->np = (int64_t)c->n * c->n * c->n;
    c->n3 = 3 * c->np;
    c->M = (size_t)c->res * c->res * c->res;
    c->HC = (size_t)c->res * c->res * (c->res / 2 + 1);
    c->nshifts = aa_shifts(cfg->antialias, 0.5 * cfg->boxsize / cfg->res_pm, c->shifts);
    double h = cfg->boxsize / cfg->res_pm;
    if (cfg->grad_order == 2) { c->fdc[0] = 1 / (2 * h); }
    else if (cfg->grad_order == 4) { c->fdc[0] = 8 / (12 * h); c->fdc[1] = -1 / (12 * h); }
    else if (cfg->grad_order == 6) { c->fdc[0] = 45 / (60 * h); c->fdc[1] = -9 / (60 * h); c->fdc[2] = 1 / (60 * h); }
    dj_status r = DJ_OK;
    auto bail = [&](dj_status code) {
        g_create_error = c->err;
        free_all(c);
        delete c;
        return code;
    };
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) { hip_fail(c, e, "hipSetDevice"); return bail(DJ_ERR_DEVICE); }
    // the device's copy of the allocation-free spectrum's scratch array (once per device)
    e = pk_scratch_init(c->dev);
    if (e != hipSuccess) { hip_fail(c, e, "hipGetSymbolAddress(scratch)"); return bail(DJ_ERR_DEVICE); }
    e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) { hip_fail(c, e, "hipStreamCreate"); return bail(DJ_ERR_DEVICE); }
    r = c->tsize == 8 ? setup_ops<double>(c) : setup_ops<float>(c);
    if (r) return bail(r);
    r = setup_tiles(c);
    if (r) return bail(r);
    if (cfg->n_resample > 1) {
        r = c->tsize == 8 ? setup_resampling<double>(c) : setup_resampling<float>(c);
        if (r) return bail(r);
    }
    size_t pbytes = (size_t)c->n3 * c->tsize, cbytes = c->HC * 2 * c->tsize;
    void **bufs[] = {&c->psi, &c->mom, &c->acc};
    for (void **b : bufs) {
        e = hipMalloc(b, pbytes);
        if (e != hipSuccess) { hip_fail(c, e, "hipMalloc(particles)"); return bail(DJ_ERR_OUT_OF_MEMORY); }
    }
    e = hipMalloc(&c->mesh, c->M * c->tsize);
    if (e == hipSuccess && cfg->deposit == DJ_DEPOSIT_FIXED64) e = hipMalloc(&c->meshfix, c->M * 8);
    if (e == hipSuccess) e = hipMalloc(&c->A, cbytes);
    if (e == hipSuccess) e = hipMalloc(&c->B, cbytes);
    if (e == hipSuccess && cfg->grad_order == 0) e = hipMalloc(&c->C, cbytes);
    if (e != hipSuccess) { hip_fail(c, e, "hipMalloc(mesh)"); return bail(DJ_ERR_OUT_OF_MEMORY); }
    c->part.zero(c->psi, pbytes, c->st);
    c->part.zero(c->mom, pbytes, c->st);
    e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) { hip_fail(c, e, "initialise"); return bail(DJ_ERR_DEVICE); }
    set_bytes(c);
    *out = c;
    return DJ_OK;
}

dj_status dj_destroy(dj_ctx *c) {
    if (!c) return DJ_OK;
    (void)hipSetDevice(c->dev);
    if (c->st) (void)hipStreamSynchronize(c->st);
    free_all(c);
    delete c;
    return DJ_OK;
}

const char *dj_last_error(const dj_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

dj_status dj_set_state(dj_ctx *c, const void *psi, const void *pi) {
    if (!c || !psi) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = upload_particles(c, c->psi, psi, hipMemcpyHostToDevice);
    return r ? r : upload_particles(c, c->mom, pi, hipMemcpyHostToDevice);
}

dj_status dj_set_state_device(dj_ctx *c, const void *psi, const void *pi) {
    if (!c || !psi) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = upload_particles(c, c->psi, psi, hipMemcpyDeviceToDevice);
    return r ? r : upload_particles(c, c->mom, pi, hipMemcpyDeviceToDevice);
}

static dj_status get_state(dj_ctx *c, void *psi, void *pi, hipMemcpyKind kind) {
    HIPCHK(c, hipSetDevice(c->dev));
    size_t bytes = (size_t)c->n3 * c->tsize;
    if (psi) HIPCHK(c, hipMemcpyAsync(psi, c->psi, bytes, kind, c->st));
    if (pi) HIPCHK(c, hipMemcpyAsync(pi, c->mom, bytes, kind, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return DJ_OK;
}

dj_status dj_get_state(dj_ctx *c, void *psi, void *pi) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    return get_state(c, psi, pi, hipMemcpyDeviceToHost);
}

dj_status dj_get_state_device(dj_ctx *c, void *psi, void *pi) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    return get_state(c, psi, pi, hipMemcpyDeviceToDevice);
}

dj_status dj_acceleration(dj_ctx *c, const void *psi, void *acc_out) {
    if (!c || !acc_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipSetDevice(c->dev));
    const void *src = c->psi;
    if (psi) {
        if (!c->psi_in) HIPCHK(c, hipMalloc(&c->psi_in, (size_t)c->n3 * c->tsize));
        HIPCHK(c, hipMemcpyAsync(c->psi_in, psi, (size_t)c->n3 * c->tsize, hipMemcpyHostToDevice, c->st));
        src = c->psi_in;
    }
    KickArgs k{};
    k.nshifts = c->nshifts;
    dj_status r = force(c, src, FORCE_ACC, k);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(acc_out, c->acc, (size_t)c->n3 * c->tsize, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return DJ_OK;
}

dj_status dj_step(dj_ctx *c, double dd1, double alpha, double beta, double dd2) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(c, hipSetDevice(c->dev));
    { Prof p(c, K_DRIFT); c->part.drift(c->psi, c->mom, c->n3, dd1, c->st); }
    KickArgs k{alpha, beta, dd2, 0.0, 0, c->nshifts};
    dj_status r = force(c, c->psi, FORCE_KICK, k);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return DJ_OK;
}

dj_status dj_run(dj_ctx *c, int32_t n_steps, const double *dd1, const double *alpha, const double *beta,
                 const double *dd2, const uint8_t *want, dj_step_cb cb, void *user) {
    if (!c || n_steps < 0 || (n_steps > 0 && (!dd1 || !alpha || !beta || !dd2)))
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid run arguments");
    HIPCHK(c, hipSetDevice(c->dev));
    auto wants = [&](int k) { return cb && (!want || want[k]); };
    if (wants(0)) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        cb(0, user);
    }
    if (n_steps == 0) return DJ_OK;
    { Prof p(c, K_DRIFT); c->part.drift(c->psi, c->mom, c->n3, dd1[0], c->st); }
    for (int s = 0; s < n_steps; s++) {
        if (c->nn_on && (size_t)s < c->nn_sched.size()) c->nn_a = c->nn_sched[s];
        const bool stop = wants(s + 1);
        const bool has_next = s + 1 < n_steps && !stop;
        KickArgs k{alpha[s], beta[s], dd2[s], has_next ? dd1[s + 1] : 0.0, has_next ? 1 : 0, c->nshifts};
        dj_status r = force(c, c->psi, FORCE_KICK, k);
        if (r) return r;
        if (stop) {
            HIPCHK(c, hipStreamSynchronize(c->st));
            cb(s + 1, user);
            if (s + 1 < n_steps) {
                Prof p(c, K_DRIFT);
                c->part.drift(c->psi, c->mom, c->n3, dd1[s + 1], c->st);
            }
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return check_launch(c, "run");
}

// the unshifted-or-shifted deposit of psi (NULL: the context's Psi) into the context's mesh (or its
// fixed-point mesh); *dens_out names it
static dj_status deposit_mesh(dj_ctx *c, const void *psi, const double *shift, void **dens_out) {
    const bool fixed = c->cfg.deposit == DJ_DEPOSIT_FIXED64;
    const void *src = c->psi;
    if (psi) {
        if (!c->psi_in) HIPCHK(c, hipMalloc(&c->psi_in, (size_t)c->n3 * c->tsize));
        HIPCHK(c, hipMemcpyAsync(c->psi_in, psi, (size_t)c->n3 * c->tsize, hipMemcpyHostToDevice, c->st));
        src = c->psi_in;
    }
    PartArgs pa{};
    pa.psi = src;
    pa.q1 = c->q1;
    pa.n = c->n;
    pa.res = c->res;
    pa.L = c->cfg.boxsize;
    for (int a = 0; a < 3; a++) pa.shift[a] = shift ? shift[a] : 0.0;
    void *dens = fixed ? c->meshfix : c->mesh;
    HIPCHK(c, hipMemcpyAsync(c->d_shifts + 48, pa.shift, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
    dj_status r = tile_sort(c, src, c->d_shifts + 48, 1, c->cfg.worder);
    if (r) return r;
    TileArgs ta = tile_args(c, src, c->d_shifts + 48, 1);
    { Prof p(c, K_SCATTER); c->tiles.deposit(ta, c->cfg.worder, fixed, pa.shift, dens, fscale_dep(c->cfg), c->st); }
    *dens_out = dens;
    return check_launch(c, "deposit");
}

static dj_status deposit_to(dj_ctx *c, const void *psi, const double *shift, void *mesh_out, hipMemcpyKind out_kind) {
    if (!c || !mesh_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipSetDevice(c->dev));
    void *dens = nullptr;
    dj_status r = deposit_mesh(c, psi, shift, &dens);
    if (r) return r;
    const size_t bytes = c->M * (c->cfg.deposit == DJ_DEPOSIT_FIXED64 ? 8 : c->tsize);
    HIPCHK(c, hipMemcpyAsync(mesh_out, dens, bytes, out_kind, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return DJ_OK;
}

dj_status dj_deposit(dj_ctx *c, const void *psi, const double *shift, void *mesh_out) {
    return deposit_to(c, psi, shift, mesh_out, hipMemcpyDeviceToHost);
}

dj_status dj_deposit_device(dj_ctx *c, void *d_mesh_out) {
    return deposit_to(c, nullptr, nullptr, d_mesh_out, hipMemcpyDeviceToDevice);
}

static dj_status ics_lpt(dj_ctx *c, uint32_t seed, const double *white, bool white_on_device, const double *k_tab,
                         const double *pk_tab, int32_t ntab, double D, int32_t order) {
    if (!c || !k_tab || !pk_tab || ntab < 2 || (order != 1 && order != 2))
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid IC arguments");
    const int n = c->n;
    FftOps probe{};
    if (n % 2 || !fft_dispatch<double>(n, probe) || (order == 2 && !fft_dispatch<double>(3 * n / 2, probe)))
        return fail(c, DJ_ERR_UNSUPPORTED, "n_part " + std::to_string(n) + " has no built IC transform size");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    int e = ics_lpt_device(c->icw, n, c->cfg.boxsize, order, white, white_on_device, seed, k_tab, pk_tab, ntab, D,
                           c->cfg.precision, c->psi, c->mom, c->st);
    if (e) return hip_fail(c, (hipError_t)e, "dj_ics_lpt");
    return DJ_OK;
}

dj_status dj_ics_lpt(dj_ctx *c, uint32_t seed, const double *white, const double *k_tab, const double *pk_tab,
                     int32_t ntab, double D, int32_t order) {
    return ics_lpt(c, seed, white, false, k_tab, pk_tab, ntab, D, order);
}

dj_status dj_ics_lpt_device(dj_ctx *c, uint32_t seed, const double *d_white, const double *k_tab,
                            const double *pk_tab, int32_t ntab, double D, int32_t order) {
    return ics_lpt(c, seed, d_white, true, k_tab, pk_tab, ntab, D, order);
}

// ------------------------------------------------------------------------ reverse mode
namespace {
dj_status adj_supported(dj_ctx *c) {
    if (c->nn_on) return fail(c, DJ_ERR_UNSUPPORTED, "the adjoint is built without the learned correction");
    if (c->cfg.n_resample > 1 && !c->rs_tiled)
        return fail(c, DJ_ERR_UNSUPPORTED, "the adjoint of sheet resampling needs (n_resample n_part)^3 < 2^31 points");
    return DJ_OK;
}

dj_status adj_buffers(dj_ctx *c) {
    const size_t pb = (size_t)c->n3 * c->tsize, mb = c->M * c->tsize;
    void **pbufs[] = {&c->vb_psi1, &c->vb_abar, &c->vb_psibar, &c->vb_pibar};
    for (void **b : pbufs)
        if (!*b) HIPCHK(c, hipMalloc(b, pb));
    if (!c->vb_fbar) HIPCHK(c, hipMalloc(&c->vb_fbar, 3 * mb));
    if (!c->vb_pbar) HIPCHK(c, hipMalloc(&c->vb_pbar, mb));
    if (!c->vb_fik) HIPCHK(c, hipMalloc(&c->vb_fik, 4 * mb));  // three force meshes, or four packed values per cell
    if (!c->loss_work) HIPCHK(c, hipMalloc(&c->loss_work, sizeof(double) * LOSS_PARTS));
    return DJ_OK;
}

// psibar += (d acc / d psi)^T abar at psi (device arrays), for the averaged interlacing passes: per pass the
// potential phi is recomputed with the forward kernels, abar / ns is deposited to Fbar, phibar = sum_a D_a
// Fbar_a goes through the same Poisson operator (self-adjoint), and one gather adds the weight-gradient terms.
// tail (nullable): the step's end (AdjTail), applied in the last pass's gather when the path allows it
// (FD gradient, no resampling), which *tail_done reports; abar is then overwritten.
// presorted: the context's tile lists already hold psi's sort (the checkpointed step's own lists)
dj_status vjp_force(dj_ctx *c, const void *psi, void *abar, void *psibar, const AdjTail *tail = nullptr,
                    bool *tail_done = nullptr, bool presorted = false) {
    if (tail_done) *tail_done = false;
    const dj_pm_config &cfg = c->cfg;
    const bool fixed = cfg.deposit == DJ_DEPOSIT_FIXED64;
    const int ns = c->nshifts, N = c->res;
    void *dens = fixed ? c->meshfix : c->mesh;
    const double fscale = std::ldexp(1.0, -(cfg.fixed_frac_bits ? cfg.fixed_frac_bits : 32));
    dj_status r = presorted ? DJ_OK : tile_sort(c, psi, c->d_shifts, ns, cfg.worder);
    if (r) return r;
    for (int s = 0; s < ns; s++) {
        TileArgs ta = tile_args(c, psi, c->d_shifts, ns);
        if (cfg.n_resample > 1) {
            // the forward's density: the particles and every offset's points through the point lattice
            const int nrn = cfg.n_resample * c->n;
            c->rs.shifted(psi, c->q1, c->n, c->shifts[s], cfg.n_resample, c->rl_pts, c->st);
            r = c->tsize == 8 ? resample_deposit<double>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg), c->rl_pts)
                              : resample_deposit<float>(c, psi, c->shifts[s], dens, fixed, fscale_dep(cfg), c->rl_pts);
            if (r) return r;
            r = tile_sort_of(c, c->tlr, c->rl_pts, c->rl_q0, nrn, c->d_shifts + 51, 1, cfg.worder);
            if (r) return r;
            const double zero3[3] = {0, 0, 0};
            TileArgs tr = tile_args_of(c, c->tlr, c->rl_pts, c->rl_q0, nrn, c->d_shifts + 51, 1);
            c->tiles.deposit(tr, cfg.worder, fixed, zero3, dens, fscale_dep(cfg), c->st);
            ta = tile_args(c, psi, c->d_shifts, ns);  // the particles' lists, for the cotangent's vector deposit
        } else {
            c->tiles.deposit(ta, cfg.worder, fixed, c->shifts[s], dens, fscale_dep(cfg), c->st);
        }
        c->fft.xf(dens, fixed, c->A, c->tw, fscale, c->st);
        c->fft.yf(c->A, c->B, c->tw, c->st);
        if (cfg.grad_order == 0) {
            // F_a = B_a rho (the forward's force meshes), Fbar = G^T abar / ns, rhobar = -sum_a B_a Fbar_a
            const size_t M = c->M;
            char *F = (char *)c->vb_fik, *Fb = (char *)c->vb_fbar;
            const size_t mb = M * c->tsize;
            c->fft.zm(ZM_NONE, true, false, c->B, c->B, c->tw, c->kmult, c->st);
            for (int ax = 0; ax < 3; ax++) {
                c->fft.zm(ZM_GRADX + ax, false, true, c->B, c->C, c->tw, c->kmult, c->st);
                c->fft.yi(c->C, c->A, c->tw, c->st);
                c->fft.xi(c->A, F + ax * mb, c->tw, c->st);
            }
            if (!c->tiles.deposit3(ta, cfg.worder, c->shifts[s], abar, 1.0 / ns, c->vb_fbar, c->loss_work, c->st)) {
                c->part.zero(c->vb_fbar, 3 * c->M * c->tsize, c->st);
                c->adj.scatter3(psi, c->q1, c->n, N, cfg.boxsize, c->shifts[s], abar, 1.0 / ns, cfg.worder, c->vb_fbar,
                                c->st);
            }
            c->part.zero(c->vb_pbar, mb, c->st);
            for (int ax = 0; ax < 3; ax++) {
                c->fft.xf(Fb + ax * mb, false, c->A, c->tw, 1.0, c->st);
                c->fft.yf(c->A, c->B, c->tw, c->st);
                c->fft.zm(ZM_NONE, true, false, c->B, c->B, c->tw, c->kmult, c->st);
                c->fft.zm(ZM_GRADX + ax, false, true, c->B, c->C, c->tw, c->kmult, c->st);
                c->fft.yi(c->C, c->A, c->tw, c->st);
                c->fft.xi(c->A, c->mesh, c->tw, c->st);
                c->adj.axpy(c->vb_pbar, c->mesh, M, -1.0, c->st);
            }
            c->adj.gather_grad_f(psi, c->q1, c->n, N, cfg.boxsize, c->shifts[s], c->vb_pbar, c->vb_fik, abar, 1.0 / ns,
                                 cfg.worder, psibar, c->st);
            if (cfg.n_resample > 1) {
                r = c->tsize == 8 ? resample_vjp<double>(c, c->vb_pbar, psibar)
                                  : resample_vjp<float>(c, c->vb_pbar, psibar);
                if (r) return r;
            }
            r = check_launch(c, "force adjoint (spectral gradient)");
            if (r) return r;
            continue;
        }
        c->fft.zm(ZM_POT, true, true, c->B, c->B, c->tw, c->kmult, c->st);
        c->fft.yi(c->B, c->A, c->tw, c->st);
        c->fft.xi(c->A, c->mesh, c->tw, c->st);
        // Fbar = G^T abar / ns: the tiled vector deposit, or device atomics on a single-tile axis
        if (!c->tiles.deposit3(ta, cfg.worder, c->shifts[s], abar, 1.0 / ns, c->vb_fbar, c->loss_work, c->st)) {
            c->part.zero(c->vb_fbar, 3 * c->M * c->tsize, c->st);
            c->adj.scatter3(psi, c->q1, c->n, N, cfg.boxsize, c->shifts[s], abar, 1.0 / ns, cfg.worder, c->vb_fbar,
                            c->st);
        }
        c->adj.divergence(c->vb_fbar, N, c->fdc, c->vb_pbar, c->st);
        c->fft.xf(c->vb_pbar, false, c->A, c->tw, 1.0, c->st);
        c->fft.yf(c->A, c->B, c->tw, c->st);
        c->fft.zm(ZM_POT, true, true, c->B, c->B, c->tw, c->kmult, c->st);
        c->fft.yi(c->B, c->A, c->tw, c->st);
        c->fft.xi(c->A, c->vb_pbar, c->tw, c->st);
        // the forward's force meshes from phi, packed with rhobar, then the weight-gradient gather over them
        c->adj.fd_force3_packed(c->mesh, c->vb_pbar, N, c->fdc, c->vb_fik, c->st);
        const bool tl = tail && s == ns - 1 && cfg.n_resample <= 1;
        c->adj.gather_grad_v(psi, c->q1, c->n, N, cfg.boxsize, c->shifts[s], c->vb_fik, abar, 1.0 / ns, cfg.worder,
                             psibar, tl ? tail : nullptr, c->st);
        if (tl && tail_done) *tail_done = true;
        if (cfg.n_resample > 1) {
            r = c->tsize == 8 ? resample_vjp<double>(c, c->vb_pbar, psibar)
                              : resample_vjp<float>(c, c->vb_pbar, psibar);
            if (r) return r;
        }
        r = check_launch(c, "force adjoint");
        if (r) return r;
    }
    return DJ_OK;
}

}  // namespace

dj_status dj_vjp_acceleration(dj_ctx *c, const void *psi, const void *acc_bar, void *psi_bar_out) {
    if (!c || !acc_bar || !psi_bar_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    dj_status r = adj_supported(c);
    if (r) return r;
    HIPCHK(c, hipSetDevice(c->dev));
    r = adj_buffers(c);
    if (r) return r;
    const size_t pb = (size_t)c->n3 * c->tsize;
    const void *src = c->psi;
    if (psi) {
        HIPCHK(c, hipMemcpyAsync(c->vb_psi1, psi, pb, hipMemcpyHostToDevice, c->st));
        src = c->vb_psi1;
    }
    HIPCHK(c, hipMemcpyAsync(c->vb_abar, acc_bar, pb, hipMemcpyHostToDevice, c->st));
    c->vb_resident = false;
    HIPCHK(c, hipMemsetAsync(c->vb_psibar, 0, pb, c->st));
    r = vjp_force(c, src, c->vb_abar, c->vb_psibar);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(psi_bar_out, c->vb_psibar, pb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return DJ_OK;
}

static dj_status vjp_deposit(dj_ctx *c, const void *psi, const void *mesh_bar, hipMemcpyKind bar_kind,
                             void *psi_bar_out) {
    if (!c || !mesh_bar) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = adj_buffers(c);
    if (r) return r;
    const size_t pb = (size_t)c->n3 * c->tsize;
    const void *src = c->psi;
    if (psi) {
        HIPCHK(c, hipMemcpyAsync(c->vb_psi1, psi, pb, hipMemcpyHostToDevice, c->st));
        src = c->vb_psi1;
    }
    HIPCHK(c, hipMemcpyAsync(c->vb_pbar, mesh_bar, c->M * c->tsize, bar_kind, c->st));
    c->vb_resident = false;
    HIPCHK(c, hipMemsetAsync(c->vb_psibar, 0, pb, c->st));
    const double zero[3] = {0, 0, 0};
    c->adj.gather_grad(src, c->q1, c->n, c->res, c->cfg.boxsize, zero, c->vb_pbar, nullptr, nullptr, 0.0, c->fdc,
                       c->cfg.worder, c->vb_psibar, c->st);
    r = check_launch(c, "deposit adjoint");
    if (r) return r;
    if (psi_bar_out) {
        HIPCHK(c, hipMemcpyAsync(psi_bar_out, c->vb_psibar, pb, hipMemcpyDeviceToHost, c->st));
    } else {
        HIPCHK(c, hipMemsetAsync(c->vb_pibar, 0, pb, c->st));  // resident cotangents (Psibar, Pibar = 0)
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->vb_resident = !psi_bar_out;
    return DJ_OK;
}

dj_status dj_vjp_deposit(dj_ctx *c, const void *psi, const void *mesh_bar, void *psi_bar_out) {
    return vjp_deposit(c, psi, mesh_bar, hipMemcpyHostToDevice, psi_bar_out);
}

dj_status dj_vjp_deposit_device(dj_ctx *c, const void *d_mesh_bar) {
    return vjp_deposit(c, nullptr, d_mesh_bar, hipMemcpyDeviceToDevice, nullptr);
}

dj_status dj_density_loss_device(dj_ctx *c, const double *d_target, double *loss_out) {
    if (!c || !loss_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    if (c->cfg.deposit == DJ_DEPOSIT_FIXED64)
        return fail(c, DJ_ERR_UNSUPPORTED, "the density loss is built on the float deposit");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = adj_buffers(c);
    if (r) return r;
    void *dens = nullptr;
    r = deposit_mesh(c, nullptr, nullptr, &dens);
    if (r) return r;
    const double rho = (double)c->np / (double)c->M;  // particles per cell
    c->adj.mse(dens, d_target, c->M, rho, c->vb_pbar, c->loss_work, c->st);
    // the deposit's reverse at the context's Psi, left resident (Psibar, Pibar = 0)
    const size_t pb = (size_t)c->n3 * c->tsize;
    c->vb_resident = false;
    HIPCHK(c, hipMemsetAsync(c->vb_psibar, 0, pb, c->st));
    HIPCHK(c, hipMemsetAsync(c->vb_pibar, 0, pb, c->st));
    const double zero[3] = {0, 0, 0};
    c->adj.gather_grad(c->psi, c->q1, c->n, c->res, c->cfg.boxsize, zero, c->vb_pbar, nullptr, nullptr, 0.0, c->fdc,
                       c->cfg.worder, c->vb_psibar, c->st);
    r = check_launch(c, "density loss");
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(loss_out, c->loss_work + LOSS_PARTS - 1, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    c->vb_resident = true;
    return DJ_OK;
}

dj_status dj_run_checkpointed(dj_ctx *c, int32_t n_steps, const double *dd1, const double *alpha, const double *beta,
                              const double *dd2) {
    if (!c || n_steps < 1 || !dd1 || !alpha || !beta || !dd2)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid run arguments");
    // options without a reverse pass are refused before the checkpoints are allocated and the run made
    dj_status ra = adj_supported(c);
    if (ra) return ra;
    HIPCHK(c, hipSetDevice(c->dev));
    const size_t pb = (size_t)c->n3 * c->tsize;
    if (c->ckpt && c->ck_steps < n_steps) {
        HIPCHK(c, hipFree(c->ckpt));
        c->ckpt = nullptr;
    }
    if (!c->ckpt) HIPCHK(c, hipMalloc(&c->ckpt, 2 * pb * (size_t)n_steps));
    c->ck_steps = n_steps;
    c->ck_dd1.assign(dd1, dd1 + n_steps);
    c->ck_alpha.assign(alpha, alpha + n_steps);
    c->ck_beta.assign(beta, beta + n_steps);
    c->ck_dd2.assign(dd2, dd2 + n_steps);
    // Each step's tile lists are kept (swapped out of the context after its force) when the device has room
    // for them: about twice the current sort's reserve per step, with 8 GB to spare
    bool keep = c->cfg.n_resample <= 1;
    if (keep) {
        size_t fr = 0, tot = 0;
        const size_t per = (c->tl.ent_cap ? c->tl.ent_cap : 4 * (size_t)c->n3) * c->tiles.entry_bytes;
        size_t have = 0;
        for (auto &t : c->ck_tl) have += t.ent_cap * c->tiles.entry_bytes;
        keep = hipMemGetInfo(&fr, &tot) == hipSuccess && fr + have > 2 * per * (size_t)n_steps + ((size_t)8 << 30);
    }
    if (!keep) {
        for (auto &t : c->ck_tl) free_lists(t);
        c->ck_tl.clear();
    } else {
        if ((int)c->ck_tl.size() < n_steps) c->ck_tl.resize(n_steps);
        for (int s = 0; s < n_steps; s++)
            if (!c->ck_tl[s].cnt) {
                dj_status ra = alloc_lists(c, c->ck_tl[s], c->tl.ent_cap);
                if (ra) return ra;
            }
    }
    c->ck_tl_ok.assign(n_steps, 0);
    for (int s = 0; s < n_steps; s++) {
        // the step's checkpoint (psi, mom before its drift) written by the drift as it reads them
        char *ck = (char *)c->ckpt + 2 * pb * (size_t)s;
        c->part.drift_to(c->psi, c->psi, c->mom, ck, ck + pb, c->n3, dd1[s], c->st);
        KickArgs k{alpha[s], beta[s], dd2[s], 0.0, 0, c->nshifts};
        dj_status r = force(c, c->psi, FORCE_KICK, k);
        if (r) return r;
        if (keep) {
            std::swap(c->tl, c->ck_tl[s]);  // the next sort reserves from the swapped-in set's counts
            c->ck_tl_ok[s] = 1;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return check_launch(c, "checkpointed run");
}

dj_status dj_vjp_run(dj_ctx *c, void *psi_bar, void *pi_bar) {
    if (!c || (!psi_bar) != (!pi_bar)) return fail(c, DJ_ERR_INVALID_ARGUMENT, "psi_bar and pi_bar: both or neither");
    if (!c->ckpt || c->ck_steps < 1) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no checkpointed run to reverse");
    if (!psi_bar && !c->vb_resident)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "no resident cotangents (dj_vjp_deposit with psi_bar_out NULL first)");
    dj_status r = adj_supported(c);
    if (r) return r;
    HIPCHK(c, hipSetDevice(c->dev));
    r = adj_buffers(c);
    if (r) return r;
    const size_t pb = (size_t)c->n3 * c->tsize, np3 = (size_t)c->n3;
    if (psi_bar) {
        HIPCHK(c, hipMemcpyAsync(c->vb_psibar, psi_bar, pb, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(c->vb_pibar, pi_bar, pb, hipMemcpyHostToDevice, c->st));
    }
    bool kicked = false;  // the step's kick backward already made at the end of the later step's reverse
    for (int s = c->ck_steps - 1; s >= 0; s--) {
        const char *ck = (const char *)c->ckpt + 2 * pb * (size_t)s;
        // the drifted positions of the step's force, recomputed from its checkpoint
        c->part.drift_to(c->vb_psi1, ck, ck + pb, nullptr, nullptr, c->n3, c->ck_dd1[s], c->st);
        if (!kicked)
            c->adj.kick(c->vb_psibar, c->vb_pibar, c->vb_abar, np3, c->ck_alpha[s], c->ck_beta[s], c->ck_dd2[s], c->st);
        AdjTail tail;
        tail.pibar = c->vb_pibar;
        tail.dd1 = c->ck_dd1[s];
        tail.kick = s > 0;
        if (s > 0) {
            tail.alpha = c->ck_alpha[s - 1];
            tail.beta = c->ck_beta[s - 1];
            tail.dd2 = c->ck_dd2[s - 1];
        }
        bool done = false;
        // the step's kept lists: the drifted checkpoint is the position its force sorted, bit for bit
        const bool kept = s < (int)c->ck_tl_ok.size() && c->ck_tl_ok[s] && s < (int)c->ck_tl.size();
        if (kept) std::swap(c->tl, c->ck_tl[s]);
        r = vjp_force(c, c->vb_psi1, c->vb_abar, c->vb_psibar, &tail, &done, kept);
        if (kept) std::swap(c->tl, c->ck_tl[s]);
        if (r) return r;
        if (!done) c->adj.drift(c->vb_psibar, c->vb_pibar, np3, c->ck_dd1[s], c->st);
        kicked = done && s > 0;
    }
    r = check_launch(c, "run adjoint");
    if (r) return r;
    if (psi_bar) {
        HIPCHK(c, hipMemcpyAsync(psi_bar, c->vb_psibar, pb, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(pi_bar, c->vb_pibar, pb, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->vb_resident = !psi_bar;
    return DJ_OK;
}

static dj_status vjp_ics_lpt(dj_ctx *c, uint32_t seed, const double *white, bool on_device, const double *k_tab,
                             const double *pk_tab, int32_t ntab, double D, int32_t order, const void *psi_bar,
                             const void *pi_bar, double *white_bar_out) {
    if (!c || !k_tab || !pk_tab || ntab < 2 || (order != 1 && order != 2) || (!psi_bar) != (!pi_bar) ||
        !white_bar_out)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid IC adjoint arguments");
    if (!psi_bar && !c->vb_resident)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "no resident cotangents (dj_vjp_deposit / dj_vjp_run with NULL first)");
    const int n = c->n;
    FftOps probe{};
    if (n % 2 || !fft_dispatch<double>(n, probe) || (order == 2 && !fft_dispatch<double>(3 * n / 2, probe)))
        return fail(c, DJ_ERR_UNSUPPORTED, "n_part " + std::to_string(n) + " has no built IC transform size");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = adj_buffers(c);
    if (r) return r;
    const size_t pb = (size_t)c->n3 * c->tsize;
    if (psi_bar) {
        HIPCHK(c, hipMemcpyAsync(c->vb_psibar, psi_bar, pb, hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(c->vb_pibar, pi_bar, pb, hipMemcpyHostToDevice, c->st));
        c->vb_resident = false;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    int e = ics_lpt_vjp_device(c->icw, n, c->cfg.boxsize, order, white, on_device, seed, k_tab, pk_tab, ntab, D,
                               c->cfg.precision, c->vb_psibar, c->vb_pibar, white_bar_out, on_device, c->st);
    if (e) return hip_fail(c, (hipError_t)e, "dj_vjp_ics_lpt");
    return DJ_OK;
}

dj_status dj_vjp_ics_lpt(dj_ctx *c, uint32_t seed, const double *white, const double *k_tab, const double *pk_tab,
                         int32_t ntab, double D, int32_t order, const void *psi_bar, const void *pi_bar,
                         double *white_bar_out) {
    return vjp_ics_lpt(c, seed, white, false, k_tab, pk_tab, ntab, D, order, psi_bar, pi_bar, white_bar_out);
}

dj_status dj_vjp_ics_lpt_device(dj_ctx *c, uint32_t seed, const double *d_white, const double *k_tab,
                                const double *pk_tab, int32_t ntab, double D, int32_t order, double *d_white_bar_out) {
    return vjp_ics_lpt(c, seed, d_white, true, k_tab, pk_tab, ntab, D, order, nullptr, nullptr, d_white_bar_out);
}

// ------------------------------------------------------------------------ learned correction
namespace {
enum { NN_L = NN_PARAMS_PER_LAYER, NN_W = 27 * 16 * 16 };

static float *nn_layer(void *p, int l) { return (float *)p + (size_t)l * NN_L; }
static void *nn_op(void *w, int l) { return (char *)w + (size_t)l * 27 * 64 * 8; }
static void *nn_pop(void *w, int l) { return (char *)w + (size_t)l * 4 * 64 * 8; }

dj_status nn_supported(dj_ctx *c) {
    if (c->cfg.precision != 32 || c->cfg.deposit != DJ_DEPOSIT_FLOAT)
        return fail(c, DJ_ERR_UNSUPPORTED, "the learned correction is built for FP32 meshes with the float deposit");
    if (c->cfg.grad_order == 0)
        return fail(c, DJ_ERR_UNSUPPORTED, "the learned correction is built for finite-difference gradients");
    if (c->res % 64) return fail(c, DJ_ERR_UNSUPPORTED, "the learned correction needs res_pm to be a multiple of 64");
    return DJ_OK;
}

dj_status nn_alloc(dj_ctx *c) {
    if (c->nn_p) return DJ_OK;
    const size_t pb = sizeof(float) * NN_LAYERS * NN_L, tb = c->M * 16 * 2;
    void **pv[] = {&c->nn_p, &c->nn_g, &c->nn_m, &c->nn_v, &c->nn_mask};
    for (void **b : pv) HIPCHK(c, hipMalloc(b, pb));
    HIPCHK(c, hipMalloc(&c->nn_wf, (size_t)NN_LAYERS * 27 * 64 * 8));
    HIPCHK(c, hipMalloc(&c->nn_wb, (size_t)NN_LAYERS * 27 * 64 * 8));
    HIPCHK(c, hipMalloc(&c->nn_pf, (size_t)NN_LAYERS * 4 * 64 * 8));
    HIPCHK(c, hipMalloc(&c->nn_pb, (size_t)NN_LAYERS * 4 * 64 * 8));
    for (void *&t : c->nn_t) HIPCHK(c, hipMalloc(&t, tb));
    HIPCHK(c, hipMalloc(&c->nn_dphi, c->M * 4));
    HIPCHK(c, hipMalloc(&c->nn_dens, c->M * 4));
    HIPCHK(c, hipMalloc(&c->nn_tgt, (size_t)c->n3 * 4));
    HIPCHK(c, hipMalloc(&c->nn_sum, sizeof(double) * c->nn.residual_sum_doubles()));
    HIPCHK(c, hipMalloc(&c->nn_wpart, c->nn.wgrad_scratch_bytes()));
    // the real parameters: layer 1 has inputs 0, 1; layer 4 has output 0
    std::vector<float> mask((size_t)NN_LAYERS * NN_L, 0.0f);
    for (int l = 0; l < NN_LAYERS; l++)
        for (int t = 0; t < 27; t++)
            for (int co = 0; co < 16; co++)
                for (int ci = 0; ci < 16; ci++)
                    if ((l != 0 || ci < 2) && (l != 3 || co == 0)) mask[(size_t)l * NN_L + (t * 16 + co) * 16 + ci] = 1;
    for (int l = 0; l < NN_LAYERS; l++)
        for (int co = 0; co < 16; co++)
            if (l != 3 || co == 0) mask[(size_t)l * NN_L + NN_W + co] = 1;
    HIPCHK(c, hipMemcpy(c->nn_mask, mask.data(), pb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->nn_m, 0, pb));
    HIPCHK(c, hipMemset(c->nn_v, 0, pb));
    return DJ_OK;
}

void nn_swizzle_all(dj_ctx *c) {
    for (int l = 0; l < NN_LAYERS; l++)
        c->nn.swizzle(nn_layer(c->nn_p, l), nn_op(c->nn_wf, l), nn_op(c->nn_wb, l), nn_pop(c->nn_pf, l),
                      nn_pop(c->nn_pb, l), c->st);
}

// the net on a density mesh (float, unnormalised deposit) into nn_dphi; pre-activations kept when training
void nn_forward_dev(dj_ctx *c, const void *dens, double a) {
    const int N = c->res;
    const double inv_rho = (double)c->M / (double)c->np;
    void **t = c->nn_t;
    const bool keep = c->nn_train;
    c->nn.conv(1, 0, N, nullptr, dens, inv_rho, a, nn_pop(c->nn_pf, 0), nn_layer(c->nn_p, 0) + NN_W, nullptr, nullptr,
               keep ? t[0] : nullptr, t[3], nullptr, 0.0, c->st);
    c->nn.conv(0, 0, N, t[3], nullptr, 0.0, 0.0, nn_op(c->nn_wf, 1), nn_layer(c->nn_p, 1) + NN_W, t[3], nullptr,
               keep ? t[1] : nullptr, t[4], nullptr, 0.0, c->st);
    c->nn.conv(0, 0, N, t[4], nullptr, 0.0, 0.0, nn_op(c->nn_wf, 2), nn_layer(c->nn_p, 2) + NN_W, t[4], nullptr,
               keep ? t[2] : nullptr, t[5], nullptr, 0.0, c->st);
    c->nn.conv(0, 1, N, t[5], nullptr, 0.0, 0.0, nn_op(c->nn_wf, 3), nn_layer(c->nn_p, 3) + NN_W, nullptr, nullptr,
               nullptr, nullptr, c->nn_dphi, c->nn_oscale, c->st);
}

// parameter gradients into nn_g from gphi = dL / d(dphi) (float mesh), after nn_forward_dev with nn_train set
dj_status nn_backward_dev(dj_ctx *c, const void *gphi, const void *dens, double a) {
    const int N = c->res;
    const double inv_rho = (double)c->M / (double)c->np, os = c->nn_oscale;
    void **t = c->nn_t;
    HIPCHK(c, hipMemsetAsync(c->nn_g, 0, sizeof(float) * NN_LAYERS * NN_L, c->st));
    float *g3 = nn_layer(c->nn_g, 3), *g2 = nn_layer(c->nn_g, 2), *g1 = nn_layer(c->nn_g, 1), *g0 = nn_layer(c->nn_g, 0);
    // layer 4: dphi = os (W4 * h3 + b4)
    c->nn.wgrad(0, 1, N, t[5], nullptr, 0.0, 0.0, nullptr, gphi, os, g3, g3 + NN_W, c->nn_wpart, c->st);
    c->nn.conv(2, 2, N, nullptr, gphi, os, 0.0, nn_pop(c->nn_pb, 3), nullptr, nullptr, t[2], t[7], t[6], nullptr, 0.0,
               c->st);  // gh3 (t6), gz3 (t7)
    // layer 3: h3 = h2 + relu(z3)
    c->nn.wgrad(0, 0, N, t[4], nullptr, 0.0, 0.0, t[7], nullptr, 0.0, g2, g2 + NN_W, c->nn_wpart, c->st);
    c->nn.conv(0, 2, N, t[7], nullptr, 0.0, 0.0, nn_op(c->nn_wb, 2), nullptr, t[6], t[1], t[9], t[8], nullptr, 0.0,
               c->st);  // gh2 (t8), gz2 (t9)
    // layer 2: h2 = h1 + relu(z2)
    c->nn.wgrad(0, 0, N, t[3], nullptr, 0.0, 0.0, t[9], nullptr, 0.0, g1, g1 + NN_W, c->nn_wpart, c->st);
    c->nn.conv(0, 2, N, t[9], nullptr, 0.0, 0.0, nn_op(c->nn_wb, 1), nullptr, t[8], t[0], t[7], t[6], nullptr, 0.0,
               c->st);  // gh1 (t6), gz1 (t7)
    // layer 1: h1 = relu(z1), input (log(1 + delta), a)
    c->nn.wgrad(1, 0, N, nullptr, dens, inv_rho, a, t[7], nullptr, 0.0, g0, g0 + NN_W, c->nn_wpart, c->st);
    return check_launch(c, "learned correction backward");
}

}  // namespace

dj_status dj_nn_set_params(dj_ctx *c, const float *params) {
    if (!c || !params) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    dj_status r = nn_supported(c);
    if (r) return r;
    HIPCHK(c, hipSetDevice(c->dev));
    r = nn_alloc(c);
    if (r) return r;
    HIPCHK(c, hipMemcpy(c->nn_p, params, sizeof(float) * NN_LAYERS * NN_L, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->nn_m, 0, sizeof(float) * NN_LAYERS * NN_L));
    HIPCHK(c, hipMemset(c->nn_v, 0, sizeof(float) * NN_LAYERS * NN_L));
    nn_swizzle_all(c);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_launch(c, "learned correction parameters");
}

dj_status dj_nn_get_params(dj_ctx *c, float *params) {
    if (!c || !params) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->nn_p) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no parameters set");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpy(params, c->nn_p, sizeof(float) * NN_LAYERS * NN_L, hipMemcpyDeviceToHost));
    return DJ_OK;
}

dj_status dj_nn_enable(dj_ctx *c, int32_t on, double out_scale) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    if (on && !c->nn_p) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no parameters set");
    c->nn_on = on != 0;
    c->nn_oscale = out_scale;
    return DJ_OK;
}

extern "C++" {
namespace {
template <typename T>
dj_status set_kgain_t(dj_ctx *c, const double *gain, int32_t nbins, double kmax) {
    KMult<T> K;
    std::memcpy(&K, c->kmult, sizeof(K));
    if (!gain || nbins == 0) {
        K.kg = nullptr;
        K.kg_n = 0;
    } else {
        std::vector<T> g(gain, gain + nbins);
        if (c->kgain) HIPCHK(c, hipFree(c->kgain));
        c->kgain = nullptr;
        HIPCHK(c, hipMalloc(&c->kgain, sizeof(T) * nbins));
        HIPCHK(c, hipMemcpy(c->kgain, g.data(), sizeof(T) * nbins, hipMemcpyHostToDevice));
        K.kg = (const T *)c->kgain;
        K.kg_n = nbins;
        K.kg_inv_dk = (T)((nbins - 1) / kmax);
    }
    std::memcpy(c->kmult, &K, sizeof(K));
    return DJ_OK;
}
}  // namespace
}  // extern "C++"

dj_status dj_set_kgain(dj_ctx *c, const double *gain, int32_t nbins, double kmax) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    if (gain && nbins != 0 && (nbins < 2 || nbins > 65536 || !(kmax > 0)))
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "transfer function: nbins in [2, 65536] and kmax > 0");
    if (gain)
        for (int i = 0; i < nbins; i++)
            if (!std::isfinite(gain[i])) return fail(c, DJ_ERR_INVALID_ARGUMENT, "transfer function not finite");
    HIPCHK(c, hipSetDevice(c->dev));
    // the kernels of earlier calls may still read the old table
    HIPCHK(c, hipStreamSynchronize(c->st));
    return c->tsize == 8 ? set_kgain_t<double>(c, gain, nbins, kmax) : set_kgain_t<float>(c, gain, nbins, kmax);
}

dj_status dj_nn_set_schedule(dj_ctx *c, int32_t n, const double *a_force) {
    if (!c || n < 0 || (n > 0 && !a_force)) return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid schedule");
    c->nn_sched.assign(a_force, a_force + n);
    if (n > 0) c->nn_a = a_force[0];
    return DJ_OK;
}

dj_status dj_nn_forward(dj_ctx *c, const float *dens, double a, float *dphi_out) {
    if (!c || !dens || !dphi_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->nn_p) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no parameters set");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpyAsync(c->nn_dens, dens, c->M * 4, hipMemcpyHostToDevice, c->st));
    const bool tr = c->nn_train;
    c->nn_train = true;
    nn_forward_dev(c, c->nn_dens, a);
    c->nn_train = tr;
    dj_status r = check_launch(c, "learned correction");
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(dphi_out, c->nn_dphi, c->M * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return DJ_OK;
}

dj_status dj_nn_backward(dj_ctx *c, const float *dens, double a, const float *gphi, float *grad_out) {
    if (!c || !dens || !gphi || !grad_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->nn_p) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no parameters set");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpyAsync(c->nn_dens, dens, c->M * 4, hipMemcpyHostToDevice, c->st));
    const bool tr = c->nn_train;
    c->nn_train = true;
    nn_forward_dev(c, c->nn_dens, a);
    c->nn_train = tr;
    HIPCHK(c, hipMemcpyAsync(c->nn_dphi, gphi, c->M * 4, hipMemcpyHostToDevice, c->st));
    dj_status r = nn_backward_dev(c, c->nn_dphi, c->nn_dens, a);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(grad_out, c->nn_g, sizeof(float) * NN_LAYERS * NN_L, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return DJ_OK;
}

dj_status dj_nn_train_step(dj_ctx *c, const void *psi, const void *acc_target, double a, double weight, double lr,
                           int32_t step, double *loss_out) {
    if (!c || !psi || !acc_target || step < 1) return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid training arguments");
    if (!c->nn_p) return fail(c, DJ_ERR_INVALID_ARGUMENT, "no parameters set");
    if (c->nshifts != 1) return fail(c, DJ_ERR_UNSUPPORTED, "training is built for antialias = 0 (one pass)");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = adj_buffers(c);
    if (r) return r;
    const size_t pb = (size_t)c->n3 * 4;
    HIPCHK(c, hipMemcpyAsync(c->vb_psi1, psi, pb, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->nn_tgt, acc_target, pb, hipMemcpyHostToDevice, c->st));
    // forward: the corrected acceleration of the snapshot
    const bool on = c->nn_on;
    c->nn_on = true;
    c->nn_train = true;
    c->nn_a = a;
    KickArgs k{};
    k.nshifts = c->nshifts;
    r = force(c, c->vb_psi1, FORCE_ACC, k);
    c->nn_train = false;
    c->nn_on = on;
    if (r) return r;
    // loss = weight x mean over components of (acc - target)^2, abar = dL / d acc
    c->nn.residual(c->acc, c->nn_tgt, c->vb_abar, (size_t)c->n3, 2.0 * weight / (double)c->n3, c->nn_sum, c->st);
    // dL / dphi: the gather's transpose (deposit of abar) and the FD stencil's transpose.  The deposit runs
    // on the snapshot's tile lists in fixed point (order-free sums, so a training run repeats bit for bit),
    // or with device atomics on a single-tile axis
    const double zero[3] = {0, 0, 0};
    r = tile_sort(c, c->vb_psi1, c->d_shifts, c->nshifts, c->cfg.worder);
    if (r) return r;
    TileArgs ta = tile_args(c, c->vb_psi1, c->d_shifts, c->nshifts);
    if (!c->tiles.deposit3(ta, c->cfg.worder, zero, c->vb_abar, 1.0, c->vb_fbar, c->loss_work, c->st)) {
        c->part.zero(c->vb_fbar, 3 * c->M * 4, c->st);
        c->adj.scatter3(c->vb_psi1, c->q1, c->n, c->res, c->cfg.boxsize, zero, c->vb_abar, 1.0, c->cfg.worder,
                        c->vb_fbar, c->st);
    }
    c->adj.divergence(c->vb_fbar, c->res, c->fdc, c->vb_pbar, c->st);
    r = nn_backward_dev(c, c->vb_pbar, c->nn_dens, a);
    if (r) return r;
    const double b1 = 0.9, b2 = 0.999;
    c->nn.adam(c->nn_p, c->nn_g, c->nn_m, c->nn_v, c->nn_mask, NN_LAYERS * NN_L, lr, b1, b2, 1e-8,
               1.0 - std::pow(b1, step), 1.0 - std::pow(b2, step), c->st);
    nn_swizzle_all(c);
    double sum = 0;
    HIPCHK(c, hipMemcpyAsync(&sum, (double *)c->nn_sum + (c->nn.residual_sum_doubles() - 1), sizeof(double),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (loss_out) *loss_out = weight * sum / (double)c->n3;
    return check_launch(c, "learned correction training step");
}

dj_status dj_power_spectrum(dj_ctx *c, const void *psi, int32_t W, int32_t nbins, double *k_out, double *pk_out,
                            int64_t *nmodes_out) {
    if (!c || !k_out || !pk_out || W < 0 || W > 4 || nbins < 1 || nbins > 2048)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid power spectrum arguments");
    HIPCHK(c, hipSetDevice(c->dev));
    const int dw = W == 0 ? 2 : W;
    const void *src = c->psi;
    if (psi) {
        if (!c->psi_in) HIPCHK(c, hipMalloc(&c->psi_in, (size_t)c->n3 * c->tsize));
        HIPCHK(c, hipMemcpyAsync(c->psi_in, psi, (size_t)c->n3 * c->tsize, hipMemcpyHostToDevice, c->st));
        src = c->psi_in;
    }
    PartArgs pa{};
    pa.psi = src;
    pa.q1 = c->q1;
    pa.n = c->n;
    pa.res = c->res;
    pa.L = c->cfg.boxsize;
    const double zero3[3] = {0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(c->d_shifts + 48, zero3, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
    dj_status rs = tile_sort(c, src, c->d_shifts + 48, 1, dw);
    if (rs) return rs;
    TileArgs ta = tile_args(c, src, c->d_shifts + 48, 1);
    c->tiles.deposit(ta, dw, false, zero3, c->mesh, 1.0, c->st);
    c->fft.xf(c->mesh, false, c->A, c->tw, 1.0, c->st);
    c->fft.yf(c->A, c->B, c->tw, c->st);
    c->fft.zm(ZM_NONE, true, false, c->B, c->B, c->tw, c->kmult, c->st);
    void *sums = nullptr;
    HIPCHK(c, hipMalloc(&sums, (size_t)nbins * 24));
    c->part.zero(sums, (size_t)nbins * 24, c->st);
    double *sp = (double *)sums, *sw = sp + nbins;
    unsigned long long *cnt = (unsigned long long *)(sw + nbins);
    const double np6 = (double)c->np * (double)c->np, V = c->cfg.boxsize * c->cfg.boxsize * c->cfg.boxsize;
    if (c->tsize == 8) pk_bin<double>(c->B, c->res, c->cfg.boxsize, W, nbins, V / np6, sp, sw, cnt, c->st);
    else pk_bin<float>(c->B, c->res, c->cfg.boxsize, W, nbins, V / np6, sp, sw, cnt, c->st);
    std::vector<double> h(3 * (size_t)nbins);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), sums, (size_t)nbins * 24, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void)hipFree(sums);
    if (e != hipSuccess) return hip_fail(c, e, "dj_power_spectrum");
    const double kf = 2 * M_PI / c->cfg.boxsize;
    const unsigned long long *hc = (const unsigned long long *)(h.data() + 2 * nbins);
    for (int b = 0; b < nbins; b++) {
        k_out[b] = kf * (b + 1);
        pk_out[b] = h[nbins + b] > 0 ? h[b] / h[nbins + b] : 0.0;
        if (nmodes_out) nmodes_out[b] = (int64_t)hc[b];
    }
    return DJ_OK;
}

// ------------------------------------------------------------------------ binned spectra
static_assert(DJ_PK_MAX_BINS == PK_MAX_BINS, "the header's bin limit is the kernels'");
extern "C++" {
namespace {
// evaluate_power_spectrum's bins (power_spectrum.jl:46-51): nbins + 1 edges between kf = 2 pi / L and
// k_Nyq sqrt(3), k_Nyq = pi N / L, as LinRange ((1 - t) a + t b, t = j / nbins) of k or of log10 k
PkEdges reference_edges(int nbins, int logarithmic, int N, double L) {
    PkEdges ed{};
    ed.nbins = nbins;
    const double dk = 2 * M_PI / L, kny = M_PI * N / L, kmin = dk, kmax = kny * std::sqrt(3.0);
    for (int b = 0; b <= nbins; b++) {
        const double t = (double)b / nbins;
        ed.e[b] = logarithmic ? exp10((1 - t) * std::log10(kmin) + t * std::log10(kmax)) : (1 - t) * kmin + t * kmax;
    }
    return ed;
}

dj_status pk_work(dj_ctx *c) {
    // [0, 4B) bin sums, [4B, 8B) k, P, mode counts, a second P; [8B, 12B) a second set of bin sums
    if (!c->pk_work) HIPCHK(c, hipMalloc(&c->pk_work, sizeof(double) * 12 * DJ_PK_MAX_BINS));
    return DJ_OK;
}

// rfft of a real (res^3) device field of the context's precision into the spectrum S (x-halved B layout)
void field_fft(dj_ctx *c, const void *field, void *S) {
    c->fft.xf(field, false, c->A, c->tw, 1.0, c->st);
    c->fft.yf(c->A, S, c->tw, c->st);
    c->fft.zm(ZM_NONE, true, false, S, S, c->tw, c->kmult, c->st);
}

void bin(dj_ctx *c, const void *F1, const void *F2, int N, int W, int wmode, double scale, const PkEdges &ed,
         double *sums) {
    if (c->tsize == 8) pk_edges<double>(F1, F2, N, c->cfg.boxsize, W, wmode, scale, ed, sums, c->st);
    else pk_edges<float>(F1, F2, N, c->cfg.boxsize, W, wmode, scale, ed, sums, c->st);
}

// outputs of the bins (device pk_work areas) to the host
dj_status pk_download(dj_ctx *c, int nbins, double *k_out, double *pk_out, int64_t *nm_out, double *pk2_out) {
    const size_t B = DJ_PK_MAX_BINS;
    double *w = c->pk_work;
    if (k_out) HIPCHK(c, hipMemcpyAsync(k_out, w + 4 * B, nbins * sizeof(double), hipMemcpyDeviceToHost, c->st));
    if (pk_out) HIPCHK(c, hipMemcpyAsync(pk_out, w + 5 * B, nbins * sizeof(double), hipMemcpyDeviceToHost, c->st));
    if (nm_out) HIPCHK(c, hipMemcpyAsync(nm_out, w + 6 * B, nbins * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    if (pk2_out) HIPCHK(c, hipMemcpyAsync(pk2_out, w + 7 * B, nbins * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return check_launch(c, "binned spectrum");
}

template <typename T>
dj_status setup_pgrid(dj_ctx *c) {
    if (c->pg_A) return DJ_OK;
    const int n = c->n;
    if (!fft_dispatch<T>(n, c->pg_fft))
        return fail(c, DJ_ERR_UNSUPPORTED, "the displacement spectrum needs n_part in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024}");
    std::vector<cx<T>> tw(n);
    for (int k = 0; k < n; k++) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = {(T)std::cos(a), (T)std::sin(a)};
    }
    const size_t hc = (size_t)n * n * (n / 2 + 1);
    HIPCHK(c, hipMalloc(&c->pg_tw, sizeof(cx<T>) * n));
    HIPCHK(c, hipMemcpy(c->pg_tw, tw.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->pg_B, hc * sizeof(cx<T>)));
    HIPCHK(c, hipMalloc(&c->pg_A, hc * sizeof(cx<T>)));
    return DJ_OK;
}
}  // namespace
}  // extern "C++"

dj_status dj_field_power_spectrum(dj_ctx *c, const void *field, int32_t nbins, int32_t logarithmic, double *k_out,
                                  double *pk_out, int64_t *nmodes_out) {
    if (!c || !field || !k_out || !pk_out || nbins < 1 || nbins > DJ_PK_MAX_BINS)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid power spectrum arguments (1 <= nbins <= 256)");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = pk_work(c);
    if (r) return r;
    const int N = c->res;
    const PkEdges ed = reference_edges(nbins, logarithmic, N, c->cfg.boxsize);
    const double V = c->cfg.boxsize * c->cfg.boxsize * c->cfg.boxsize, n3 = (double)N * N * N;
    HIPCHK(c, hipMemcpyAsync(c->mesh, field, c->M * c->tsize, hipMemcpyHostToDevice, c->st));
    field_fft(c, c->mesh, c->B);
    const size_t B = DJ_PK_MAX_BINS;
    zero_doubles(c->pk_work, 4 * B, c->st);
    bin(c, c->B, nullptr, N, 0, 0, V / (n3 * n3), ed, c->pk_work);
    pk_finalize(c->pk_work, ed, 1.0, 1, false, c->pk_work + 4 * B, c->pk_work + 5 * B, (int64_t *)(c->pk_work + 6 * B),
                c->st);
    return pk_download(c, nbins, k_out, pk_out, nmodes_out, nullptr);
}

dj_status dj_field_cross_power_spectrum(dj_ctx *c, const void *field1, const void *field2, int32_t nbins,
                                        int32_t logarithmic, double *k_out, double *pk_out) {
    if (!c || !field1 || !field2 || !k_out || !pk_out || nbins < 1 || nbins > DJ_PK_MAX_BINS)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid cross power spectrum arguments (1 <= nbins <= 256)");
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = pk_work(c);
    if (r) return r;
    if (!c->C) HIPCHK(c, hipMalloc(&c->C, c->HC * 2 * c->tsize));
    const int N = c->res;
    const PkEdges ed = reference_edges(nbins, logarithmic, N, c->cfg.boxsize);
    const double V = c->cfg.boxsize * c->cfg.boxsize * c->cfg.boxsize, n3 = (double)N * N * N;
    HIPCHK(c, hipMemcpyAsync(c->mesh, field1, c->M * c->tsize, hipMemcpyHostToDevice, c->st));
    field_fft(c, c->mesh, c->B);
    HIPCHK(c, hipMemcpyAsync(c->mesh, field2, c->M * c->tsize, hipMemcpyHostToDevice, c->st));
    field_fft(c, c->mesh, c->C);
    const size_t B = DJ_PK_MAX_BINS;
    zero_doubles(c->pk_work, 4 * B, c->st);
    bin(c, c->B, c->C, N, 0, 1, V / (n3 * n3), ed, c->pk_work);
    pk_finalize(c->pk_work, ed, 1.0, 1, false, c->pk_work + 4 * B, c->pk_work + 5 * B, nullptr, c->st);
    return pk_download(c, nbins, k_out, pk_out, nullptr, nullptr);
}

dj_status dj_power_spectra_binned(dj_ctx *c, const void *psi, int32_t nbins, const double *edges, double *k_out,
                                  double *pk_delta_out, double *pk_psi_out) {
    if (!c || !edges || nbins < 1 || nbins > DJ_PK_MAX_BINS || (!pk_delta_out && !pk_psi_out))
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "invalid binned spectrum arguments (1 <= nbins <= 256)");
    PkEdges ed{};
    ed.nbins = nbins;
    for (int b = 0; b <= nbins; b++) {
        ed.e[b] = edges[b];
        if (!std::isfinite(edges[b]) || (b && !(edges[b] > edges[b - 1])))
            return fail(c, DJ_ERR_INVALID_ARGUMENT, "bin edges must be finite and ascending");
    }
    HIPCHK(c, hipSetDevice(c->dev));
    dj_status r = pk_work(c);
    if (r) return r;
    const void *src = c->psi;
    if (psi) {
        if (!c->psi_in) HIPCHK(c, hipMalloc(&c->psi_in, (size_t)c->n3 * c->tsize));
        HIPCHK(c, hipMemcpyAsync(c->psi_in, psi, (size_t)c->n3 * c->tsize, hipMemcpyHostToDevice, c->st));
        src = c->psi_in;
    }
    const size_t B = DJ_PK_MAX_BINS;
    const double L = c->cfg.boxsize, V = L * L * L;
    if (pk_delta_out) {
        // CIC density of q + psi on the mesh, delta = mesh N^3 / n^3 - 1: |delta_k|^2 V / N^6 = |mesh_k|^2 V / n^6
        const double zero3[3] = {0, 0, 0};
        HIPCHK(c, hipMemcpyAsync(c->d_shifts + 48, zero3, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
        r = tile_sort(c, src, c->d_shifts + 48, 1, 2);
        if (r) return r;
        TileArgs ta = tile_args(c, src, c->d_shifts + 48, 1);
        c->tiles.deposit(ta, 2, false, zero3, c->mesh, 1.0, c->st);
        field_fft(c, c->mesh, c->B);
        zero_doubles(c->pk_work, 4 * B, c->st);
        bin(c, c->B, nullptr, c->res, 2, 0, V / ((double)c->np * (double)c->np), ed, c->pk_work);
        pk_finalize(c->pk_work, ed, 1.0, 1, true, c->pk_work + 4 * B, c->pk_work + 5 * B, nullptr, c->st);
    }
    if (pk_psi_out) {
        // the displacement's components on the particle grid, transformed one at a time
        const int n = c->n;
        const FftOps *f = &c->fft;
        const void *tw = c->tw;
        void *A = c->A, *S = c->B;
        if (n != c->res) {
            r = c->tsize == 8 ? setup_pgrid<double>(c) : setup_pgrid<float>(c);
            if (r) return r;
            f = &c->pg_fft; tw = c->pg_tw; A = c->pg_A; S = c->pg_B;
        }
        double *sums = c->pk_work + 8 * B;
        zero_doubles(sums, 4 * B, c->st);
        for (int comp = 0; comp < 3; comp++) {
            f->xf((const char *)src + (size_t)comp * c->np * c->tsize, false, A, tw, 1.0, c->st);
            f->yf(A, S, tw, c->st);
            f->zm(ZM_NONE, true, false, S, S, tw, c->kmult, c->st);
            bin(c, S, nullptr, n, 0, 0, V / ((double)c->np * (double)c->np), ed, sums);
        }
        pk_finalize(sums, ed, 3.0, 3, true, c->pk_work + 4 * B, c->pk_work + 7 * B, nullptr, c->st);
    }
    return pk_download(c, nbins, k_out, pk_delta_out, nullptr, pk_psi_out);
}

dj_status dj_field_power_spectrum_noalloc(dj_ctx *c, const void *d_field, int32_t nbins, int32_t logarithmic,
                                          double *d_k_out, double *d_pk_out, int64_t *d_nmodes_out) {
    if (!c || !d_field || !d_pk_out)
        return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    if (nbins < 1 || nbins > DJ_PK_MAX_BINS)
        return fail(c, DJ_ERR_INVALID_ARGUMENT,
                    "nbins " + std::to_string(nbins) + " outside 1..256, the bins the scratch array holds");
    double *sums = pk_scratch(c->dev);
    if (!sums) return fail(c, DJ_ERR_DEVICE, "the scratch array of this device was not set up");
    HIPCHK(c, hipSetDevice(c->dev));
    const int N = c->res;
    const PkEdges ed = reference_edges(nbins, logarithmic, N, c->cfg.boxsize);
    const double V = c->cfg.boxsize * c->cfg.boxsize * c->cfg.boxsize, n3 = (double)N * N * N;
    field_fft(c, d_field, c->B);
    zero_doubles(sums, 4 * nbins, c->st);
    bin(c, c->B, nullptr, N, 0, 0, V / (n3 * n3), ed, sums);
    pk_finalize(sums, ed, 1.0, 1, false, d_k_out, d_pk_out, d_nmodes_out, c->st);
    HIPCHK(c, hipGetLastError());
    return DJ_OK;
}

dj_status dj_ngenic_field_on_device(uint32_t seed, int32_t res, int32_t device, double *out) {
    if (!out || res < 2 || res % 2) return fail(nullptr, DJ_ERR_INVALID_ARGUMENT, "invalid N-GenIC arguments");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
    hipStream_t st = nullptr;
    double *d = nullptr;
    const size_t bytes = 2 * sizeof(double) * (size_t)res * res * (res / 2 + 1);
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d, bytes);
    if (e == hipSuccess) e = (hipError_t)ngenic_field_device(seed, res, d, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d) (void)hipFree(d);
    if (st) (void)hipStreamDestroy(st);
    return e == hipSuccess ? DJ_OK : hip_fail(nullptr, e, "dj_ngenic_field_on_device");
}

dj_status dj_ngenic_field(uint32_t seed, int32_t res, double *out) {
    if (!out || res < 2 || res % 2) return fail(nullptr, DJ_ERR_INVALID_ARGUMENT, "invalid N-GenIC arguments");
    ngenic_field(seed, res, out);
    return DJ_OK;
}

dj_status dj_ranlxd1_uniforms(uint32_t seed, int32_t n, double *out) {
    if (!out || n < 0) return fail(nullptr, DJ_ERR_INVALID_ARGUMENT, "invalid arguments");
    ranlxd1_uniforms(seed, n, out);
    return DJ_OK;
}

dj_status dj_check_finite(dj_ctx *c) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    size_t bytes = (size_t)c->n3 * c->tsize;
    std::vector<unsigned char> h(bytes);
    void *bufs[2] = {c->psi, c->mom};
    for (void *b : bufs) {
        HIPCHK(c, hipMemcpyAsync(h.data(), b, bytes, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int64_t i = 0; i < c->n3; i++) {
            double v = c->tsize == 8 ? ((double *)h.data())[i] : (double)((float *)h.data())[i];
            if (!std::isfinite(v)) return fail(c, DJ_ERR_NONFINITE, "non-finite particle state");
        }
    }
    return DJ_OK;
}

dj_status dj_synchronize(dj_ctx *c) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    return DJ_OK;
}

dj_status dj_set_profiling(dj_ctx *c, int32_t on) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    c->prof = on != 0;
    return DJ_OK;
}

dj_status dj_reset_stats(dj_ctx *c) {
    if (!c) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null context");
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    for (int i = 0; i < K_COUNT; i++) { c->calls[i] = 0; c->ms[i] = 0; }
    return DJ_OK;
}

dj_status dj_kernel_stats(dj_ctx *c, dj_kernel_stat *out, int32_t cap, int32_t *n_out) {
    if (!c || !n_out) return fail(c, DJ_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_prof(c);
    int n = 0;
    for (int i = 0; i < K_COUNT && n < cap; i++) {
        if (!out) break;
        std::snprintf(out[n].name, sizeof(out[n].name), "%s", kernel_names[i]);
        out[n].calls = c->calls[i];
        out[n].ms_total = c->ms[i];
        out[n].bytes_per_call = c->bytes[i];
        n++;
    }
    *n_out = out ? n : K_COUNT;
    return DJ_OK;
}

dj_status dj_fft3d(int32_t N, int32_t precision, int32_t direction, const void *in, void *out, int32_t device) {
    if (!in || !out || (direction != -1 && direction != 1) || (precision != 32 && precision != 64))
        return fail(nullptr, DJ_ERR_INVALID_ARGUMENT, "invalid fft arguments");
    dj_pm_config cfg{};
    cfg.res_pm = N; cfg.n_part = 1; cfg.boxsize = 1.0; cfg.worder = 2; cfg.grad_order = 2;
    cfg.n_resample = 1; cfg.precision = precision; cfg.device = device;
    dj_ctx *c = nullptr;
    dj_status r = dj_create(&cfg, &c);
    if (r) return r;
    const int T = c->tsize, H = N / 2 + 1;
    const size_t M = c->M, HC = c->HC;
    std::vector<unsigned char> hb(HC * 2 * T);
    auto fin = [&](dj_status code) {
        if (code) g_create_error = c->err;
        dj_destroy(c);
        return code;
    };
    if (direction < 0) {
        HIPCHK(c, hipMemcpy(c->mesh, in, M * T, hipMemcpyHostToDevice));
        c->fft.xf(c->mesh, false, c->A, c->tw, 1.0, c->st);
        c->fft.yf(c->A, c->B, c->tw, c->st);
        c->fft.zm(ZM_NONE, true, false, c->B, c->B, c->tw, c->kmult, c->st);
        if ((r = check_launch(c, "fft3d"))) return fin(r);
        HIPCHK(c, hipMemcpyAsync(hb.data(), c->B, HC * 2 * T, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        // B[kz + N (kx + H ky)] -> out[kx + H (ky + N kz)]
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++)
                for (int kz = 0; kz < N; kz++)
                    std::memcpy((unsigned char *)out + 2 * T * (kx + (size_t)H * (ky + (size_t)N * kz)),
                                hb.data() + 2 * T * (kz + (size_t)N * (kx + (size_t)H * ky)), 2 * T);
    } else {
        for (int ky = 0; ky < N; ky++)
            for (int kx = 0; kx < H; kx++)
                for (int kz = 0; kz < N; kz++)
                    std::memcpy(hb.data() + 2 * T * (kz + (size_t)N * (kx + (size_t)H * ky)),
                                (const unsigned char *)in + 2 * T * (kx + (size_t)H * (ky + (size_t)N * kz)), 2 * T);
        HIPCHK(c, hipMemcpy(c->B, hb.data(), HC * 2 * T, hipMemcpyHostToDevice));
        c->fft.zm(ZM_NONE, false, true, c->B, c->B, c->tw, c->kmult, c->st);
        c->fft.yi(c->B, c->A, c->tw, c->st);
        c->fft.xi(c->A, c->mesh, c->tw, c->st);
        if ((r = check_launch(c, "fft3d"))) return fin(r);
        HIPCHK(c, hipMemcpyAsync(out, c->mesh, M * T, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    return fin(DJ_OK);
}

}  // extern "C"
