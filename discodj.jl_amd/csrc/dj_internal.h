// Internal interfaces between the C-ABI layer (api.cpp) and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include <cstdint>

namespace dj {

// FFT passes of one mesh size and precision (fft.hip)
struct FftOps {
    void (*xf)(const void *mesh, bool fixed, void *A, const void *tw, double fscale, hipStream_t st);
    // FP64: xf of the real field -(mub a) (neg) or mub (a + b) (b nullable), formed as the pass loads it
    void (*xf_partner)(const double *mub, const double *a, const double *b, int neg, void *A, const void *tw,
                       hipStream_t st);
    // FP64, on the ext grid of the 2LPT source: the inverse z pass of A_ik formed from phi (F, n grid) as it
    // is loaded, and the forward z pass of an A_ik cotangent whose output is added into Fb (n grid)
    void (*zm_deriv_pad)(const void *F, void *Be, int n, double L, int i, int k, double s, const void *tw,
                         hipStream_t st);
    void (*zm_deriv_adj)(const void *Be, void *Fb, int n, double L, int i, int k, double sme, double s,
                         const void *tw, hipStream_t st);
    // FP64 IC spectra (mode: IC_PSI1 / IC_PSI2, and IC_DERIV / IC_PSI2_ADJ of ic_modes.h): the inverse z pass
    // of the multiplied spectrum, and the forward z pass whose output is weighted, multiplied, accumulated
    void (*zm_mult_inv)(int mode, const void *in, void *out, int c, double L, const void *tw, hipStream_t st);
    void (*zm_fwd_mult_acc)(int mode, const void *in, void *acc, int c, double L, double s, const void *tw,
                            hipStream_t st);
    void (*xi)(const void *A, void *mesh, const void *tw, hipStream_t st);
    // the same with the output multiplied by s as it is stored
    void (*xi_scaled)(const void *A, void *mesh, const void *tw, double s, hipStream_t st);
    void (*yf)(const void *A, void *B, const void *tw, hipStream_t st);
    void (*yi)(const void *B, void *A, const void *tw, hipStream_t st);
    void (*zm)(int mode, bool fwd, bool inv, const void *in, void *out, const void *tw, const void *K,
               hipStream_t st);
    int lines_per_block;
};
template <typename T> bool fft_dispatch(int N, FftOps &ops);

// particle-side kernels (particles.hip)
struct PartArgs {
    const void *psi;      // (3, Np) displacement, component-major
    void *mom;            // (3, Np) momentum
    void *acc;            // (3, Np) acceleration accumulator
    const void *q1;       // (n) Lagrangian lattice coordinates (i * L/n in FP64, rounded to T)
    int n;                // particles per side
    int res;              // mesh cells per side
    double L;             // box size
    double shift[3];      // interlacing shift of this pass
};

enum { GATHER_ACC = 0, GATHER_ASSIGN = 1, GATHER_KICK = 2 };

struct KickArgs {
    double alpha, beta, dd2, dd1_next;  // FP64 step coefficients, rounded to T at use
    int has_next;                       // apply the next step's first drift in the same pass
    int nshifts;                        // interlacing passes averaged
};

struct PartOps {
    // unit-mass deposit of X = q + psi + shift (float atomics, or fixed-point int64 when fixed)
    void (*scatter)(const PartArgs &a, int W, bool fixed, int frac_bits, void *mesh, hipStream_t st);
    // gather -D phi (FD order 2/4/6) at X; mode GATHER_ACC (acc += s), GATHER_ASSIGN (acc = s, lean
    // per-axis semantics are equal), GATHER_KICK (final pass fused with kick and drift)
    void (*gather_fd)(const PartArgs &a, int W, int order, const void *phi, const double *c, int mode,
                      const KickArgs &k, void *psi_out, hipStream_t st);
    // gather one component of a mesh field at X into acc[comp] (+=)
    void (*gather_comp)(const PartArgs &a, int W, const void *field, int comp, hipStream_t st);
    // psi += dd * mom
    void (*drift)(void *psi, const void *mom, int64_t n3, double dd, hipStream_t st);
    // out = psi + dd * mom (out may be psi); psi and mom also copied to ck_psi, ck_mom when both are non-null
    void (*drift_to)(void *out, const void *psi, const void *mom, void *ck_psi, void *ck_mom, int64_t n3, double dd,
                     hipStream_t st);
    // mom = alpha mom + beta acc/ns ; psi += dd2 mom [; psi += dd1_next mom]
    void (*kick_drift)(void *psi, void *mom, const void *acc, int64_t n3, const KickArgs &k, hipStream_t st);
    // x *= s
    void (*scale)(void *x, int64_t n, double s, hipStream_t st);
    // x = 0 (a fill kernel, graph-capture safe)
    void (*zero)(void *x, size_t bytes, hipStream_t st);
};
template <typename T> void part_ops(PartOps &ops);

// tile binning (tiles.hip): mesh tiles of TX x TY x TZ cells; RL / RR: stencil reach in cells left
// and right of a particle's cell over all interlacing shifts (sizes the staged potential of the gather)
struct TileGeom {
    int N, TX, TY, TZ, ntx, nty, ntz, nt, RL, RR;
    int lT[3];  // log2 of the (power-of-two) tile dimensions
    int RS;     // right reach of the gather's staged region (RR without the rounding-tie margin)
};
struct TileArgs {
    const void *psi;       // (3, Np) displacement
    void *mom, *acc;       // (3, Np)
    const void *q1;        // lattice coordinates
    int n;
    double L;
    const double *shifts;  // device (ns, 3) interlacing shifts used by the binning
    int ns;
    TileGeom g;
    unsigned *cnt;         // (2 nt) own then ghost counts of the counting pass
    unsigned *start;       // (nt + 1) first entry of each tile's own list; start[nt] = entries reserved
    unsigned *gstart;      // (nt) first entry of each tile's ghost list (own capacity after start)
    unsigned *cur;         // (2 nt) fill cursors: after the fill, the own and ghost counts
    unsigned *flag;        // (1) set by the fill when a list outgrows its capacity or the buffer
    unsigned long long *scan_part;  // (nt / 1024 + 1) block sums of the list scan
    void *ent;             // entries (X, particle index)
    size_t ent_cap;        // entries the buffer holds
    int acc_by_entry;      // the gather's acc is (ent_cap, 3) in list order (between interlacing passes)
    unsigned *own_at;      // (Np, nullable): the fill records each particle's own-list entry
};
struct TileOps {
    void (*count)(const TileArgs &a, int W, hipStream_t st);  // exact counts into cnt
    // list capacities from src (2 nt counts; may be cur), plus slack when slack != 0; zeroes cur and flag
    void (*scan)(const TileArgs &a, const unsigned *src, int slack, hipStream_t st);
    void (*fill)(const TileArgs &a, int W, hipStream_t st);
    // reverse mode: out_c = sum_p W(X_p + sh - cell) vals[p + c np] scale (3 meshes, every cell written);
    // false when the mesh has a single tile along an axis or the accumulators exceed the LDS.  work: 16 bytes
    // of device memory (the FP32 fixed-point scales)
    bool (*deposit3)(const TileArgs &a, int W, const double *sh, const void *vals, double scale, void *out,
                     void *work, hipStream_t st);
    void (*deposit)(const TileArgs &a, int W, bool fixed, const double *shift, void *mesh, double fscale,
                    hipStream_t st);
    void (*gather)(const TileArgs &a, int W, int order, const double *shift, const void *phi, const double *c,
                   int mode, const KickArgs &k, void *psi_out, hipStream_t st);
    // the kick and drifts of a step in particle order, the force summed over the ns passes in list order
    // (acc_list[3 e + c], e = own_at[p]): mom = alpha mom + beta acc / ns, psi += (dd2 + dd1n) steps of mom
    void (*kick_listed)(const unsigned *own_at, const void *acc_list, void *psi, void *mom, size_t np,
                        const KickArgs &k, hipStream_t st);
    size_t entry_bytes;
};
template <typename T> void tile_ops(TileOps &ops);

// sheet resampling (resample.hip): the steps of one offset's extra point set
struct ResampleOps {
    // psi_c = mod((psi + q)_c + s_c - q_c + L/2, L) - L/2 of every particle, natural or transposed layout
    void (*wrap)(const void *psi, const void *q1, int n, int c, double sc, double L, bool transposed, void *out,
                 hipStream_t st);
    // out = in * px py pz * scale on the transposed field's half spectrum
    void (*phase)(const void *in, void *out, int n, const void *px, const void *py, const void *pz, double scale,
                  hipStream_t st);
    // periodic trilinear shift by d (grid cells)
    void (*trilinear)(const void *f, int n, const double *d, void *out, hipStream_t st);
    // Xr_c = mod(q_c + npsi_c + shift + L, L), at point (nr i + o[0], nr j + o[1], nr k + o[2]) of the
    // (nr n)^3 point-set lattice xr (3 components); nr = 1, o = 0: the particle layout
    void (*place)(const void *npsi, const void *q1, int n, int c, double shift, double L, bool transposed, void *xr,
                  int nr, const int *o, hipStream_t st);
    // the particles' shifted positions (psi + q) + sh at points (nr i, nr j, nr k) of the point-set lattice
    void (*shifted)(const void *psi, const void *q1, int n, const double *sh, int nr, void *out, hipStream_t st);
    // unit-mass deposit of positions (3, n^3) with device atomics (float, or int64 fixed point)
    void (*scatter)(const void *X, int n, int W, void *mesh, int res, double L, bool fixed, double fscale,
                    hipStream_t st);
    // reverse mode: point set o's component c of a point-lattice cotangent into an n^3 field (natural or
    // transposed); hb (+)= in conj(px py pz scale); psibar_c += transposed v; psibar_c += the transposed
    // trilinear shift by d of point set o's cotangent
    void (*extract)(const void *pl, int n, int nr, const int *o, int c, bool transposed, void *out, hipStream_t st);
    void (*phase_adj)(const void *in, void *hb, int n, const void *px, const void *py, const void *pz, double scale,
                      bool acc, hipStream_t st);
    void (*add_transposed)(const void *v, int n, void *psibar_c, hipStream_t st);
    void (*trilinear_adj)(const void *pl, int n, int nr, const int *o, int c, const double *d, void *psibar_c,
                          hipStream_t st);
};
template <typename T> void resample_ops(ResampleOps &ops);

// reverse mode of the force and the DKD step (adjoint.hip); particle arrays (3, n^3), meshes N^3
// The end of a DKD step's reverse pass after its force: Pibar += dd1 Psibar (first drift backward), then,
// when kick, the previous step's kick backward with its alpha, beta, dd2 (abar of that step, Pibar)
struct AdjTail {
    void *pibar = nullptr;
    double dd1 = 0;
    int kick = 0;
    double alpha = 0, beta = 0, dd2 = 0;
};
struct AdjointOps {
    // Fbar_a += sum_p W(X_p - c) abar_p,a s   (3 meshes)
    void (*scatter3)(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *abar,
                     double s, int W, void *fbar, hipStream_t st);
    // phibar = sum_a D_a Fbar_a (FD coefficients c[3])
    void (*divergence)(const void *fbar, int N, const double *c, void *phibar, hipStream_t st);
    // psibar += N/L sum_c grad W(X - c) (rhobar_c + sum_a abar_a s F_a(c)), F = -D phi; rhobar, phi nullable
    void (*gather_grad)(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *rhobar,
                        const void *phi, const void *abar, double s, const double *c, int W, void *psibar,
                        hipStream_t st);
    // psibar += N/L sum_c grad W(X - c) (rhobar_c + sum_a abar_a s F_a(c)) with F the three force meshes (3 N^3)
    void (*gather_grad_f)(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *rhobar,
                          const void *F, const void *abar, double s, int W, void *psibar, hipStream_t st);
    // y += s x
    void (*axpy)(void *y, const void *x, size_t n, double s, hipStream_t st);
    // kick and second drift backward: Pibar_t = Pibar + dd2 Psibar; abar = beta Pibar_t; Pibar = alpha Pibar_t
    void (*kick)(const void *psibar, void *pibar, void *abar, size_t n, double alpha, double beta, double dd2,
                 hipStream_t st);
    // first drift backward: Pibar += dd1 Psibar
    void (*drift)(const void *psibar, void *pibar, size_t n, double dd1, hipStream_t st);
    // F_a = -D_a phi (FD coefficients c[3]), summed as gather_grad sums them, packed with rhobar:
    // V[c] = {rhobar, F_0, F_1, F_2} (4 N^3 values)
    void (*fd_force3_packed)(const void *phi, const void *rhobar, int N, const double *c, void *V, hipStream_t st);
    // gather_grad_f over the packed cells V of fd_force3_packed; with tail (nullable), the step's drift and
    // the previous step's kick backward per particle (AdjTail), abar then overwritten with that step's
    void (*gather_grad_v)(const void *psi, const void *q1, int n, int N, double L, const double *sh, const void *V,
                          void *abar, double s, int W, void *psibar, const AdjTail *tail, hipStream_t st);
    // L = sum (mesh / rho - 1 - target)^2 into part[LOSS_PARTS - 1] (part: LOSS_PARTS doubles), mesh_bar = dL/dmesh;
    // target nullable (zero)
    void (*mse)(const void *mesh, const double *target, size_t n, double rho, void *mesh_bar, double *part,
                hipStream_t st);
};
constexpr int LOSS_PARTS = 1025;
template <typename T> void adjoint_ops(AdjointOps &ops);

// learned correction (nn.hip): FP32 fields, BF16 channel-last tensors t[v * 16 + c]
enum { NN_PARAMS_PER_LAYER = 27 * 16 * 16 + 16, NN_LAYERS = 4 };
struct NnOps {
    // one 3x3x3 convolution layer (in: 0 tensor, 1 density field, 2 gradient field; epi: 0 hidden, 1 out, 2 bwd)
    void (*conv)(int in, int epi, int N, const void *hin, const void *fin, double inv_rho, double cval, const void *wsw,
                 const void *bias, const void *res, const void *zmask, void *out_z, void *out_h, void *out_f,
                 double oscale, hipStream_t st);
    // dW[27][16][16] += sum_v gz(v) x(v + off)^T, db += sum_v gz(v)  (ga: gz is a 1-channel field gf * gscale),
    // summed in a fixed order through part (wgrad_scratch_bytes())
    void (*wgrad)(int in, int ga, int N, const void *hin, const void *fin, double inv_rho, double cval, const void *gz,
                  const void *gf, double gscale, void *dW, void *db, void *part, hipStream_t st);
    size_t (*wgrad_scratch_bytes)();
    // lane-ordered BF16 operands of a layer: per tap (fwd, bwd) and packed for two-channel inputs
    void (*swizzle)(const void *W, void *fwd, void *bwd, void *pfwd, void *pbwd, hipStream_t st);
    void (*adam)(void *p, const void *g, void *m, void *v, const void *mask, int n, double lr, double b1, double b2,
                 double eps, double c1, double c2, hipStream_t st);
    void (*axpy)(void *out, const void *in, size_t n, double s, hipStream_t st);
    // sum: residual_sum_doubles() doubles, the total of r^2 in the last one (summed in a fixed order)
    void (*residual)(const void *acc, const void *tgt, void *abar, size_t n, double gs, void *sum, hipStream_t st);
    size_t (*residual_sum_doubles)();
};
void nn_ops(NnOps &ops);

// power spectrum binning of an x-halved B-layout spectrum (analysis.hip)
template <typename T>
void pk_bin(const void *B, int N, double L, int W, int nbins, double scale, double *sum_p, double *sum_w,
            unsigned long long *cnt, hipStream_t st);
// bin edges, passed by value to the kernels (no device copy): bin b is [e[b], e[b + 1])
enum { PK_MAX_BINS = 256 };
struct PkEdges {
    double e[PK_MAX_BINS + 1];
    int nbins;
};
// sums (4 nbins: sum |k|, sum w v, sum w, count) += the binned modes of the spectrum F1 (|F1|^2) or of
// F1, F2 (Re conj(F1) F2), v scaled by scale and divided by the window of order W (0: none); wmode 0:
// weights 1/2 on the reference's half-axis end planes, 1: unweighted
template <typename T>
void pk_edges(const void *F1, const void *F2, int N, double L, int W, int wmode, double scale, const PkEdges &ed,
              double *sums, hipStream_t st);
// k_out, pk_out = mult sum(w v) / sum(w), nm_out = count / ncomp (device pointers, each may be null)
void pk_finalize(const double *sums, const PkEdges &ed, double mult, int ncomp, bool centre_only, double *k_out,
                 double *pk_out, int64_t *nm_out, hipStream_t st);
void zero_doubles(double *p, int n, hipStream_t st);
// the allocation-free entry point's bin sums (4 PK_MAX_BINS doubles) on device dev: set up once per
// device with that device current (pk_scratch_init), then its address (pk_scratch; null before set-up)
hipError_t pk_scratch_init(int dev);
double *pk_scratch(int dev);

// initial conditions (ics.hip, ngenic.cpp)
void ranlxd1_uniforms(uint32_t seed, int n, double *out);
void ngenic_field(uint32_t seed, int res, double *out);
void ngenic_seed_table(uint32_t seed, int res, uint32_t *tab);
// the N-GenIC field of seed on the device (the columns' generators run one per thread), (res, res, res/2+1)
// complex, element (i, j, k) at i + res (j + res k), into device memory d_field; returns a HIP error code
int ngenic_field_device(uint32_t seed, int res, double *d_field, hipStream_t st);
// Device buffers of the IC pipeline and its reverse for one (n, order), kept between calls (the 2LPT
// transforms at 3n/2 need about 20 n^3 doubles: 5.3 GB at n = 256, 42 GB at n = 512); freed by
// ic_work_free or when a call needs another size
struct IcWork {
    int n = 0, order = 0, ntab = 0;
    std::vector<void *> bufs;
    void *twn = nullptr, *twe = nullptr, *kt = nullptr, *pt = nullptr;
    void *A = nullptr, *Bm = nullptr, *F = nullptr, *Fb = nullptr, *fLb = nullptr;
    void *w = nullptr, *p1 = nullptr, *p2 = nullptr;
    void *Ae = nullptr, *Be = nullptr, *mu = nullptr, *fld[3] = {nullptr, nullptr, nullptr};
    // The nine second-order fields A_ik of the last forward (ext grid), kept for the reverse pass of the
    // same ICs, and what they were computed from: n, L, order, the P(k) table, and the seed or (wkey) a
    // copy of the white noise.  fk_room: 0 not tried, 1 allocated, -1 no room (the forward then uses fld).
    void *fk[9] = {}, *wkey = nullptr, *flag = nullptr;
    int fk_room = 0;
    bool fk_valid = false, fk_seeded = false;
    int fk_n = 0, fk_order = 0;
    uint32_t fk_seed = 0;
    double fk_L = 0;
    std::vector<double> fk_kt, fk_pt;
};
void ic_work_free(IcWork &ws);
// White noise: `white` (n^3 real values, on the host, or on the device when white_on_device), or else
// (NULL) the N-GenIC field of seed generated on the device.  The reverse writes the white-noise cotangent
// to wbar (host, or device when wbar_on_device).  Both return a HIP error code.
int ics_lpt_device(IcWork &ws, int n, double L, int order, const double *white, bool white_on_device, uint32_t seed,
                   const double *kt, const double *pt, int ntab, double D, int prec, void *psi_out,
                   void *pi_out, hipStream_t st);
int ics_lpt_vjp_device(IcWork &ws, int n, double L, int order, const double *white, bool white_on_device,
                       uint32_t seed, const double *kt, const double *pt, int ntab, double D, int prec,
                       const void *psibar, const void *pibar, double *wbar, bool wbar_on_device, hipStream_t st);

}  // namespace dj
