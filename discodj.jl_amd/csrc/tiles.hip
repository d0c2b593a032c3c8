// Tile-binned mass assignment and force gather.
//
// The mesh is cut into tiles of TX x TY x TZ cells.  Once per force evaluation the particles are
// binned by tile (a counting sort): each particle is listed in the tile of its cell c0 = floor(g)
// ("own") and, as a ghost, in every other tile that a cell of its mass-assignment stencil falls in
// for any of the interlacing shifts.  The stencil cells are computed with exactly the arithmetic
// of the deposit (same X, same shift, same base rule), so a tile's list holds every particle that
// deposits into it.  Per-workgroup LDS hash tables aggregate the counters, so the sort issues one
// global atomic per (workgroup, tile) instead of one per particle.
//
// Deposit: one workgroup per tile accumulates its list in LDS (float / double / int64 atomics) and
// writes the tile with plain coalesced stores: no global atomics, no zero fill of the mesh.
// Gather: one workgroup per tile stages the potential of the tile plus the stencil reach and the
// finite-difference halo in LDS, evaluates -D phi at the 27 (64) stencil cells from LDS for its own
// particles, and applies the kick and drifts in place.  A particle whose stencil leaves the staged
// region (not expected; kept for safety) reads the potential from global memory instead.
#include <hip/hip_runtime.h>

#include "dj_internal.h"

namespace dj {

template <typename T> struct Entry;
template <> struct alignas(16) Entry<float> { float x, y, z; int idx; };
template <> struct alignas(16) Entry<double> { double x, y, z; int idx; int pad; };

// Julia's float mod(x, L) for L > 0.  On [-L, 2L) it is computed without fmod: there fmod returns x,
// x - L (exact, Sterbenz) or x (for -L < x < 0, then shifted by L), and every branch below gives
// the same bits as the general rule (rem, then + L for a negative remainder, +0 for a zero one).
// The three in-range cases are computed with the same operations and selected without branches;
// only positions outside [-L, 2L) take the general path.
template <typename T> __device__ __noinline__ T tjmod_general(T x, T L) {
    T r = fmod(x, L);
    if (r == (T)0) return copysign(r, L);
    if ((r > (T)0) != (L > (T)0)) return r + L;
    return r;
}
template <typename T> __device__ __forceinline__ T tjmod(T x, T L) {
    T r = x + (T)0;            // [0, L): + 0 turns -0 into +0, as copysign(0, L) does
    if (x >= L) r = x - L;     // [L, 2L)
    if (x < (T)0) r = x + L;   // [-L, 0)
    if (__builtin_expect(x < -L || x >= (T)2 * L, 0)) r = tjmod_general(x, L);
    return r;
}
// a mod n for a in [-n, 2n)
__device__ __forceinline__ int wrapn(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
template <int W, typename T> __device__ __forceinline__ int twbase(T g) {
    return (W % 2 == 0) ? (int)floor(g) : (int)rint(g);
}
template <int W> __device__ __forceinline__ constexpr int twlo() { return W == 2 ? 0 : -1; }
__device__ __forceinline__ int timod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}
// 1-D assignment weights (pm.jl:49-51).  Every piece is evaluated with the reference's operations
// and the applicable one selected, so the code has no branches and the values are the same bits.
template <int W, typename T> __device__ __forceinline__ T twk(T d) {
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

// The weight of stencil cell o (0 <= o < W) at d = |g - (b + lo + o)|, b the base cell of g.  Each
// cell's d stays in one piece of the window, so only that piece is evaluated: TSC (b = round(g),
// |g - b| <= 1/2) has d <= 1/2 at the centre (inner piece) and 1/2 <= d <= 3/2 at the sides (outer
// piece; at d = 1/2 both pieces give 1/2 and at d = 3/2 the outer one gives the reference's 0); CIC
// (b = floor(g)) has d <= 1, where the weight is 1 - d.  The values are the same bits as twk's.
template <int W, typename T> __device__ __forceinline__ T twk_at(int o, T d) {
    if constexpr (W == 3) {
        if (o == 1) return (T)0.75 - d * d;
        const T t = (T)1.5 - d;
        return (T)0.5 * (t * t);
    } else if constexpr (W == 2) {
        return (T)1 - d;
    } else {
        return twk<W>(d);
    }
}

// ------------------------------------------------------------------------------------ binning
constexpr unsigned EMPTY = 0xffffffffu;

// Tiles of one particle.  The stencil cells of every shift along an axis lie within c0 + [rlo, rhi],
// a span shorter than the tile, so the axis touches at most the two tiles {tile(c0 + rlo),
// tile(c0 + rhi)}.  The particle's ghost tiles are the combinations of those per-axis tiles other
// than its own tile (that of c0, whose own list it always joins, even when a shift moves every
// stencil cell out of it): combination c (0 <= c < 8) takes th along the axes whose bit of c is set
// (x = 1, y = 2, z = 4) and tl along the others, and exists when th != tl on each set axis.  The
// cells are computed with the arithmetic of the deposit (same X, same shift, same base rule).
struct PartTiles {
    int tl[3], th[3], to[3];
};

template <typename T, int W>
__device__ __forceinline__ PartTiles particle_tiles(const T X[3], const TileGeom &g, const double *shifts, int ns,
                                                   T L) {
    PartTiles P;
    for (int a = 0; a < 3; a++) {
        const T g0 = tjmod(X[a], L) / L * (T)g.N;
        const int c0 = wrapn((int)floor(g0), g.N);
        int rlo = 1 << 20, rhi = -(1 << 20);
        for (int s = 0; s < ns; s++) {
            const T sh = (T)shifts[3 * s + a];
            // X + 0 == X: the unshifted pass reuses g0
            const T gs = sh == (T)0 ? g0 : tjmod(X[a] + sh, L) / L * (T)g.N;
            const int b = twbase<W>(gs);
            const int d = wrapn(b - c0 + g.N / 2, g.N) - g.N / 2;  // b in [0, N], c0 in [0, N)
            rlo = min(rlo, d + twlo<W>());
            rhi = max(rhi, d + twlo<W>() + W - 1);
        }
        P.to[a] = c0 >> g.lT[a];
        P.tl[a] = wrapn(c0 + rlo, g.N) >> g.lT[a];
        P.th[a] = wrapn(c0 + rhi, g.N) >> g.lT[a];
    }
    return P;
}

// Per workgroup, the lists its particles join are counted in a window of tiles around the tile of its
// first particle: BWX x BWY x BWZ tiles (own and ghost list each), directly indexed, so a lane finds its
// counter without a search.  The window covers the spread of a 16 x 4 x 4 particle block (32 x 8 x 8
// cells at the start) plus displacements of some tens of cells; a list outside it counts with one global
// atomic per particle.  Window cells that alias one tile on a small mesh never both occur (a tile's offset
// from the window's centre is taken in (-n/2, n/2]).
constexpr int BWX = 5, BWY = 7, BWZ = 13, BSLOTS = 2 * BWX * BWY * BWZ;

__device__ __forceinline__ int tdiff(int t, int b, int nt) {
    int d = t - b;
    d = d > nt / 2 ? d - nt : d;
    return d <= -((nt + 1) / 2) ? d + nt : d;
}

// list keys (tile * 2, + 1 for a ghost list) and window slots (-1: outside): index 8 is the own tile's,
// c that of ghost combination c, or EMPTY when the combination does not exist or is the own tile.  The
// per-axis parts (list index, window offset, in-window and own-tile tests of tl, th and to) are formed
// once; each combination then selects and adds them (its selections are compile-time constants).
struct AxisParts {
    int k[3], sl[3];       // [tl, th, to]: tile index term, window slot term
    bool in[3], eqo[2];    // [tl, th, to] inside the window; tl, th equal to the own tile's coordinate
    bool two;              // th != tl
};

__device__ __forceinline__ void axis_parts(int tl, int th, int to, int tb, int nt, int wd, int mulk, int muls,
                                           AxisParts &A) {
    const int t3[3] = {tl, th, to};
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const int d = tdiff(t3[h], tb, nt) + wd / 2;
        A.in[h] = (unsigned)d < (unsigned)wd;
        A.k[h] = t3[h] * mulk;
        A.sl[h] = d * muls;
    }
    A.eqo[0] = tl == to;
    A.eqo[1] = th == to;
    A.two = th != tl;
}

template <int C>
__device__ __forceinline__ void combo_key(const AxisParts (&A)[3], unsigned &key, int &slot) {
    constexpr int sx = C & 1, sy = (C >> 1) & 1, sz = (C >> 2) & 1;
    const bool exists = (!sx || A[0].two) && (!sy || A[1].two) && (!sz || A[2].two) &&
                        !(A[0].eqo[sx] && A[1].eqo[sy] && A[2].eqo[sz]);
    key = exists ? (unsigned)(A[0].k[sx] + A[1].k[sy] + A[2].k[sz]) * 2u + 1u : EMPTY;
    slot = exists && A[0].in[sx] && A[1].in[sy] && A[2].in[sz] ? 2 * (A[0].sl[sx] + A[1].sl[sy] + A[2].sl[sz]) + 1
                                                                : -1;
}

template <int C = 0>
__device__ __forceinline__ void combo_keys(const AxisParts (&A)[3], unsigned (&key)[9], int (&slot)[9]) {
    if constexpr (C < 8) {
        combo_key<C>(A, key[C], slot[C]);
        combo_keys<C + 1>(A, key, slot);
    } else {
        key[8] = (unsigned)(A[0].k[2] + A[1].k[2] + A[2].k[2]) * 2u;
        slot[8] = A[0].in[2] && A[1].in[2] && A[2].in[2] ? 2 * (A[0].sl[2] + A[1].sl[2] + A[2].sl[2]) : -1;
    }
}

template <typename T, int W, bool FILL>
__global__ __launch_bounds__(256) void k_tile_bin(const T *__restrict__ psi, const T *__restrict__ q1, int n,
                                                  const double *__restrict__ shifts, int ns, T L, TileGeom g,
                                                  unsigned *cnt, const unsigned *__restrict__ start,
                                                  const unsigned *__restrict__ gstart, unsigned *cur, unsigned *flag,
                                                  Entry<T> *__restrict__ ent, size_t ent_cap, unsigned *own_at) {
    __shared__ unsigned vals[BSLOTS], base[BSLOTS], lend[BSLOTS];
    __shared__ int tbs[3];
    for (int s = threadIdx.x; s < BSLOTS; s += 256) vals[s] = 0;
    // 16 x 4 x 4 Lagrangian block of particles per workgroup (thread 0's particle always exists)
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 4 + ((threadIdx.x >> 4) & 3),
              k = blockIdx.z * 4 + (threadIdx.x >> 6);
    const bool ok = i < n && j < n && k < n;
    T X[3] = {0, 0, 0};
    size_t p = 0;
    PartTiles P{};
    if (ok) {
        const size_t np = (size_t)n * n * n;
        p = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
        X[0] = psi[p] + q1[i];
        X[1] = psi[p + np] + q1[j];
        X[2] = psi[p + 2 * np] + q1[k];
        P = particle_tiles<T, W>(X, g, shifts, ns, L);
    }
    if (threadIdx.x == 0) {
        tbs[0] = P.to[0];
        tbs[1] = P.to[1];
        tbs[2] = P.to[2];
    }
    __syncthreads();
    const int tb[3] = {tbs[0], tbs[1], tbs[2]};
    // the particle's list keys and window slots (every index below is a compile-time constant once the
    // loops are unrolled, so the arrays stay in registers without indexed selects)
    unsigned key[9];
    int slot[9], rank[9];
    if (ok) {
        AxisParts A[3];
        axis_parts(P.tl[0], P.th[0], P.to[0], tb[0], g.ntx, BWX, 1, 1, A[0]);
        axis_parts(P.tl[1], P.th[1], P.to[1], tb[1], g.nty, BWY, g.ntx, BWX, A[1]);
        axis_parts(P.tl[2], P.th[2], P.to[2], tb[2], g.ntz, BWZ, g.ntx * g.nty, BWX * BWY, A[2]);
        combo_keys(A, key, slot);
    } else {
#pragma unroll
        for (int c = 0; c < 9; c++) {
            key[c] = EMPTY;
            slot[c] = -1;
        }
    }
    // Ranks within the workgroup's counters.  Where every active lane of the wave joins the same list
    // (the common case), one lane adds the group's size and each lane takes the group base plus its
    // position, so a wave's particles stay contiguous (in Lagrangian order) in their list; otherwise each
    // lane makes its own LDS atomic.
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        rank[c] = -1;
        const bool act = key[c] != EMPTY && slot[c] >= 0;
        const unsigned long long am = __ballot(act);
        if (am) {
            const int leader = __ffsll((long long)am) - 1;
            const int ls = __builtin_amdgcn_readlane(slot[c], leader);
            const unsigned long long sm = __ballot(act && slot[c] == ls);
            if (sm == am) {
                unsigned b = 0;
                if (lane == leader) b = atomicAdd(&vals[ls], (unsigned)__popcll(am));
                b = (unsigned)__builtin_amdgcn_readlane((int)b, leader);
                if (act) rank[c] = (int)(b + (unsigned)__popcll(am & below));
            } else if (act) {
                rank[c] = (int)atomicAdd(&vals[slot[c]], 1u);
            }
        }
    }
    __syncthreads();
    // one global atomic per window list in use; the fill also records the group's first entry and its
    // list's end (own list [start[t], gstart[t]), ghost list [gstart[t], start[t + 1]))
    for (int s = threadIdx.x; s < BSLOTS; s += 256) {
        const unsigned v = vals[s];
        if (!v) continue;
        const unsigned gh = (unsigned)s & 1u;
        const int r = s >> 1;
        const int tx = timod(tb[0] + r % BWX - BWX / 2, g.ntx), ty = timod(tb[1] + (r / BWX) % BWY - BWY / 2, g.nty),
                  tz = timod(tb[2] + r / (BWX * BWY) - BWZ / 2, g.ntz);
        const unsigned t = (unsigned)(tx + g.ntx * (ty + g.nty * tz));
        if (!FILL) {
            atomicAdd(&cnt[gh * g.nt + t], v);
        } else {
            const unsigned b = atomicAdd(&cur[gh * g.nt + t], v);
            base[s] = (gh ? gstart[t] : start[t]) + b;
            lend[s] = gh ? start[t + 1] : gstart[t];
        }
    }
    if (!FILL) {
        // lists outside the window count directly
#pragma unroll
        for (int c = 0; c < 9; c++)
            if (key[c] != EMPTY && slot[c] < 0) atomicAdd(&cnt[(key[c] & 1u) * g.nt + (key[c] >> 1)], 1u);
        return;
    }
    __syncthreads();
    bool over = false;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        if (key[c] == EMPTY) continue;
        const unsigned t = key[c] >> 1;
        const bool gh = key[c] & 1u;
        unsigned at, end;
        if (slot[c] >= 0) {
            at = base[slot[c]] + (unsigned)rank[c];
            end = lend[slot[c]];
        } else {  // a list outside the workgroup's window
            const unsigned r = atomicAdd(&cur[(gh ? g.nt : 0) + t], 1u);
            at = (gh ? gstart[t] : start[t]) + r;
            end = gh ? start[t + 1] : gstart[t];
        }
        if (at < end && at < ent_cap) {
            Entry<T> e;
            e.x = X[0];
            e.y = X[1];
            e.z = X[2];
            e.idx = (int)p;
            ent[at] = e;
            if (c == 8 && own_at) own_at[p] = at;
        } else {
            over = true;
        }
    }
    // capacity exceeded: the host repeats the sort with exact counts (one flag atomic per wave; a
    // buffer that is much too small would otherwise have every lane hit the one flag word)
    if (__any(over) && __lane_id() == 0) atomicOr(flag, 1u);
}

// List capacities of every tile from counts src (own, then ghost; src may be the fill cursors of
// the previous sort), reserving 2 c + 64 per list when slack != 0 (with some 2^18 tiles at 512^3,
// some list almost surely grows by more than a quarter between steps; memory is plentiful):
// start[t] = sum_{u<t} capacities, gstart[t] = start[t] + own capacity, start[nt] = total reserved.
// Three parallel kernels: per-block sums, a scan of the block sums, per-block scan and write-out
// (which also clears the cursors, after reading them when src is cur, and the overflow flag).
__device__ __forceinline__ unsigned list_cap(unsigned c, int slack) { return slack ? 2 * c + 64 : c; }

constexpr int SCAN_TILES = 1024;  // tiles per block of the scan (4 per thread)

__global__ __launch_bounds__(256) void k_scan_sums(const unsigned *src, int nt, int slack, unsigned long long *part) {
    __shared__ unsigned long long red[256];
    const int t0 = blockIdx.x * SCAN_TILES;
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_TILES / 256; k++) {
        const int t = t0 + k * 256 + threadIdx.x;
        if (t < nt) s += (unsigned long long)list_cap(src[t], slack) + list_cap(src[nt + t], slack);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// exclusive scan of nb <= 1024 block sums in place; part[nb] = total
__global__ __launch_bounds__(1024) void k_scan_top(unsigned long long *part, int nb, unsigned *start, int nt,
                                                   unsigned *flag) {
    __shared__ unsigned long long v[1024];
    const int i = threadIdx.x;
    const unsigned long long x = i < nb ? part[i] : 0ull;
    v[i] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned long long y = i >= off ? v[i - off] : 0ull;
        __syncthreads();
        v[i] += y;
        __syncthreads();
    }
    if (i < nb) part[i] = v[i] - x;
    if (i == 1023) {
        start[nt] = (unsigned)min(v[1023], 0xffffffffull);
        flag[0] = v[1023] > 0xffffffffull ? 1u : 0u;
    }
}

__global__ __launch_bounds__(256) void k_scan_write(const unsigned *src, int nt, int slack,
                                                    const unsigned long long *__restrict__ part, unsigned *__restrict__ start,
                                                    unsigned *__restrict__ gstart, unsigned *cur) {
    constexpr int PER = SCAN_TILES / 256;  // consecutive tiles per thread
    __shared__ unsigned long long v[256];
    const int t0 = blockIdx.x * SCAN_TILES + threadIdx.x * PER;
    unsigned co[PER], cg[PER];
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int t = t0 + k;
        co[k] = t < nt ? list_cap(src[t], slack) : 0u;
        cg[k] = t < nt ? list_cap(src[nt + t], slack) : 0u;
        s += (unsigned long long)co[k] + cg[k];
    }
    v[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned long long y = threadIdx.x >= off ? v[threadIdx.x - off] : 0ull;
        __syncthreads();
        v[threadIdx.x] += y;
        __syncthreads();
    }
    unsigned long long run = part[blockIdx.x] + v[threadIdx.x] - s;
    // src may be cur: this thread read its tiles' counts above, before clearing them here
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int t = t0 + k;
        if (t >= nt) break;
        start[t] = (unsigned)run;
        gstart[t] = (unsigned)(run + co[k]);
        run += (unsigned long long)co[k] + cg[k];
        cur[t] = 0;
        cur[nt + t] = 0;
    }
}

// ------------------------------------------------------------------------------------ deposit
// Accumulators of a tile in LDS:
//   ACC_T      T atomics (FP64 meshes)
//   ACC_Q31    FP32 meshes: each weight w <= 1 enters as the integer nearest to w 2^31 (the float
//              product wx (wy wz) with the 2^31 in wx, plus one half in the same fused multiply-add,
//              then truncated), summed in 64-bit integers and rounded once to T at the tile store.
//              Integer LDS atomics run at about five times the rate of float ones on gfx950, and the
//              sum is independent of the order of the additions (the deposit is reproducible bit for
//              bit).  Each term carries the rounding of the float product (relative 2^-23 at most, no
//              bias) and, below 2^-8, one of at most 2^-32; the sum adds no error until the final
//              rounding, so the result is closer to the exact mass than an FP32 running sum.
//              A tile whose lists hold E < 2^12 entries sums in 32-bit integers instead (ACC_Q32):
//              the weights enter as w 2^b with b = 32 - bits(E), so that no cell's sum (at most E
//              terms of at most 2^b) can exceed 2^32; a 32-bit LDS atomic moves half the bytes of a
//              64-bit one.  Each term then carries an absolute rounding of at most 2^-(b+1) <= 2^-21.
//   ACC_FIXED  the fixed-point mode: llrint(w 2^frac) in 64-bit integers, stored as integers
enum { ACC_T = 0, ACC_Q31 = 1, ACC_FIXED = 2, ACC_Q32 = 3 };
constexpr int Q32_MIN_BITS = 20;

template <typename T, int ACC, typename Acc>
__device__ __forceinline__ void dep_add(Acc *a, T wx, T wy, T wz, T wyz, double fscale) {
    if constexpr (ACC == ACC_FIXED) {
        // the oracle's product order, (wx wy) wz, then the fixed-point scale in FP64
        const T wv = wx * wy * wz;
        atomicAdd(a, (unsigned long long)__double2ll_rn((double)wv * fscale));
    } else if constexpr (ACC == ACC_Q31) {
        atomicAdd(a, (unsigned long long)(unsigned)fmaf(wx, wyz, 0.5f));  // wx carries the 2^31
    } else if constexpr (ACC == ACC_Q32) {
        atomicAdd(a, (unsigned)fmaf(wx, wyz, 0.5f));  // wx carries the tile's 2^b
    } else {
        atomicAdd(a, wx * wy * wz);
    }
}

// Padded accumulator (PAD, every axis of the mesh holds at least two tiles): rows of SXD = TX + 2 (W-1)
// cells, the tile's x cells at [W-1, W-1+TX), for the TY TZ rows of the tile.  An entry's stencil starts at local x cell l0 in
// [-(W-1), TX) whenever it overlaps the tile, so its W x cells are always in the row (cells outside
// the tile land in the pads) and each row makes W unmasked atomics at immediate offsets from one row
// address.  Rows outside the tile are skipped: a ghost entry's stencil is mostly outside, and those
// lanes sit out the row's atomics.  Pads are not stored; their cells belong to neighbour tiles, which
// receive the same particles through their ghost lists.  With a single tile along an axis the stencil
// wraps onto the tile itself, and the masked form (!PAD) wraps each cell.
template <typename T, int W, int ACC, bool PAD>
__device__ __forceinline__ void tile_deposit(const Entry<T> *__restrict__ ent, const unsigned *__restrict__ start,
                                             const unsigned *__restrict__ gstart, const unsigned *__restrict__ cnt,
                                             const TileGeom &g, T sx, T sy, T sz, T L, void *mesh, double fscale,
                                             int qbits) {
    using Acc = typename std::conditional<ACC == ACC_T, T,
                                          typename std::conditional<ACC == ACC_Q32, unsigned, unsigned long long>::type>::type;
    extern __shared__ unsigned char smem[];
    Acc *acc = (Acc *)smem;
    constexpr int P = W - 1;
    const int t = blockIdx.x, TV = g.TX * g.TY * g.TZ;
    const int SXD = PAD ? g.TX + 2 * P : g.TX, ROWS = g.TY * g.TZ, CELLS = PAD ? SXD * ROWS : TV;
    for (int c = threadIdx.x; c < CELLS; c += blockDim.x) acc[c] = 0;
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    const int ox = tx << g.lT[0], oy = ty << g.lT[1], oz = tz << g.lT[2];
    __syncthreads();
    // the tile's own list, then its ghost list
    const unsigned n_own = cnt[t], n_all = n_own + cnt[g.nt + t], s_own = start[t], s_gh = gstart[t] - n_own;
    const int o3[3] = {ox, oy, oz}, T3[3] = {g.TX, g.TY, g.TZ};
    for (unsigned i = threadIdx.x; i < n_all; i += blockDim.x) {
        const Entry<T> en = ent[i < n_own ? s_own + i : s_gh + i];
        const T xs[3] = {en.x + sx, en.y + sy, en.z + sz};
        T w[3][W];
        int b0[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T gg = tjmod(xs[a], L) / L * (T)g.N;
            const int b = twbase<W>(gg);
#pragma unroll
            for (int o = 0; o < W; o++) w[a][o] = twk_at<W>(o, fabs(gg - (T)(b + twlo<W>() + o)));
            b0[a] = b + twlo<W>();
        }
        if constexpr (ACC == ACC_Q31 || ACC == ACC_Q32) {
            // x weights scaled by 2^31 or the tile's 2^b (exact); one w_y w_z product per row
            const T q = ACC == ACC_Q31 ? (T)2147483648.0f : (T)ldexpf(1.0f, qbits);
#pragma unroll
            for (int ix = 0; ix < W; ix++) w[0][ix] *= q;
        }
        if constexpr (PAD) {
            // first stencil cell relative to the tile, nearest periodic image (N >= 2 T per axis)
            int l0[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int l = b0[a] - o3[a];
                l = l >= (g.N >> 1) ? l - g.N : (l < -(g.N >> 1) ? l + g.N : l);
                l0[a] = l;
            }
            // x: start inside the padded row or discard the whole entry (it overlaps this tile only
            // for another interlacing shift)
            const bool okx = l0[0] >= -P && l0[0] < g.TX;
            const int xo = okx ? l0[0] + P : 0;
            // row offsets (cells) per y and z stencil cell, out-of-tile ones pushed past the last row
            constexpr int BIG = 1 << 28;
            const int SZ = SXD * g.TY, END = SXD * ROWS;
            int ly[W], lz[W];
#pragma unroll
            for (int o = 0; o < W; o++) {
                const int y = l0[1] + o, z = l0[2] + o;
                ly[o] = (unsigned)y < (unsigned)g.TY ? __mul24(y, SXD) : BIG;
                lz[o] = (okx && (unsigned)z < (unsigned)g.TZ) ? __mul24(z, SZ) : BIG;
            }
            // rows outside the tile are skipped (their lanes sit out the row's atomics); sending them to a
            // shared discard row instead made every lane's atomics meet on one row
#pragma unroll
            for (int iz = 0; iz < W; iz++)
#pragma unroll
                for (int iy = 0; iy < W; iy++) {
                    const int ro = ly[iy] + lz[iz];
                    if (ro >= END) continue;
                    Acc *r = acc + (xo + ro);
                    const T wyz = w[1][iy] * w[2][iz];
#pragma unroll
                    for (int ix = 0; ix < W; ix++) dep_add<T, ACC>(r + ix, w[0][ix], w[1][iy], w[2][iz], wyz, fscale);
                }
        } else {
            int loc[3][W];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int o = 0; o < W; o++) {
                    const int l = wrapn(b0[a] + o, g.N) - o3[a];
                    loc[a][o] = (l >= 0 && l < T3[a]) ? l : -1;
                }
#pragma unroll
            for (int iz = 0; iz < W; iz++) {
                if (loc[2][iz] < 0) continue;
#pragma unroll
                for (int iy = 0; iy < W; iy++) {
                    if (loc[1][iy] < 0) continue;
                    const int row = g.TX * (loc[1][iy] + g.TY * loc[2][iz]);
                    const T wyz = w[1][iy] * w[2][iz];
#pragma unroll
                    for (int ix = 0; ix < W; ix++) {
                        if (loc[0][ix] < 0) continue;
                        dep_add<T, ACC>(&acc[row + loc[0][ix]], w[0][ix], w[1][iy], w[2][iz], wyz, fscale);
                    }
                }
            }
        }
    }
    __syncthreads();
    // one x row of the tile per TX threads (the block size is a multiple of TX); cell indices < N^3 <= 2^30
    const unsigned N = (unsigned)g.N, lx = threadIdx.x & (g.TX - 1);
    for (int r = threadIdx.x >> g.lT[0]; r < ROWS; r += blockDim.x >> g.lT[0]) {
        const unsigned ly = r & (g.TY - 1), lz = r >> g.lT[1];
        const unsigned gi = (ox + lx) + N * ((oy + ly) + N * (oz + lz));
        const Acc v = acc[(PAD ? lx + P : lx) + SXD * r];
        if constexpr (ACC == ACC_Q31) ((T *)mesh)[gi] = (T)((double)v * 0x1p-31);
        else if constexpr (ACC == ACC_Q32) ((T *)mesh)[gi] = (T)ldexp((double)v, -qbits);
        else ((Acc *)mesh)[gi] = v;
    }
}

// FP32 tiles (ACC_Q31) with fewer than 2^12 entries sum in 32 bits (ACC_Q32); the LDS holds 64-bit
// accumulators either way
template <typename T, int W, int ACC, bool PAD>
__global__ __launch_bounds__(512) void k_tile_deposit(const Entry<T> *__restrict__ ent, const unsigned *__restrict__ start,
                                                      const unsigned *__restrict__ gstart, const unsigned *__restrict__ cnt,
                                                      TileGeom g, T sx, T sy, T sz, T L, void *mesh, double fscale) {
    if constexpr (ACC == ACC_Q31) {
        const unsigned t = blockIdx.x, e = cnt[t] + cnt[g.nt + t];
        const int qbits = 32 - (32 - __clz(e));  // 32 - bits(e): e 2^qbits < 2^32
        if (qbits >= Q32_MIN_BITS) {
            tile_deposit<T, W, ACC_Q32, PAD>(ent, start, gstart, cnt, g, sx, sy, sz, L, mesh, fscale, min(qbits, 31));
            return;
        }
    }
    tile_deposit<T, W, ACC, PAD>(ent, start, gstart, cnt, g, sx, sy, sz, L, mesh, fscale, 31);
}

// Vector deposit (reverse mode): out_c(cell) = sum_p W(X_p + s - cell) vals[p + c np] scale for c = 0, 1, 2,
// per tile in three padded LDS accumulators of T (the layout of the padded deposit above), stored as the
// tile's cells: no device atomics, no zero fill.  Every axis of the mesh holds at least two tiles.
template <typename T, int W>
__global__ __launch_bounds__(512) void k_tile_deposit3(const Entry<T> *__restrict__ ent, const unsigned *__restrict__ start,
                                                       const unsigned *__restrict__ gstart, const unsigned *__restrict__ cnt,
                                                       TileGeom g, T sx, T sy, T sz, T L, const T *__restrict__ vals,
                                                       size_t np, T vscale, T *__restrict__ out) {
    extern __shared__ unsigned char smem[];
    T *acc = (T *)smem;
    constexpr int P = W - 1;
    const int t = blockIdx.x;
    const int SXD = g.TX + 2 * P, ROWS = g.TY * g.TZ, CELLS = SXD * ROWS;
    const int SZ = SXD * g.TY, END = SXD * ROWS;
    for (int c = threadIdx.x; c < 3 * CELLS; c += blockDim.x) acc[c] = 0;
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    const int o3[3] = {tx << g.lT[0], ty << g.lT[1], tz << g.lT[2]};
    __syncthreads();
    const unsigned n_own = cnt[t], n_all = n_own + cnt[g.nt + t], s_own = start[t], s_gh = gstart[t] - n_own;
    for (unsigned i = threadIdx.x; i < n_all; i += blockDim.x) {
        const Entry<T> en = ent[i < n_own ? s_own + i : s_gh + i];
        const size_t p = (size_t)en.idx;
        const T v0 = vals[p] * vscale, v1 = vals[p + np] * vscale, v2 = vals[p + 2 * np] * vscale;
        const T xs[3] = {en.x + sx, en.y + sy, en.z + sz};
        T w[3][W];
        int l0[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T gg = tjmod(xs[a], L) / L * (T)g.N;
            const int b = twbase<W>(gg);
#pragma unroll
            for (int o = 0; o < W; o++) w[a][o] = twk_at<W>(o, fabs(gg - (T)(b + twlo<W>() + o)));
            int l = b + twlo<W>() - o3[a];
            l0[a] = l >= (g.N >> 1) ? l - g.N : (l < -(g.N >> 1) ? l + g.N : l);
        }
        const bool okx = l0[0] >= -P && l0[0] < g.TX;
        const int xo = okx ? l0[0] + P : 0;
        constexpr int BIG = 1 << 28;
        int ly[W], lz[W];
#pragma unroll
        for (int o = 0; o < W; o++) {
            const int y = l0[1] + o, z = l0[2] + o;
            ly[o] = (unsigned)y < (unsigned)g.TY ? __mul24(y, SXD) : BIG;
            lz[o] = (okx && (unsigned)z < (unsigned)g.TZ) ? __mul24(z, SZ) : BIG;
        }
#pragma unroll
        for (int iz = 0; iz < W; iz++)
#pragma unroll
            for (int iy = 0; iy < W; iy++) {
                const int ro = ly[iy] + lz[iz];
                if (ro >= END) continue;  // a row outside the tile
                T *r = acc + (xo + ro);
                const T wyz = w[1][iy] * w[2][iz];
#pragma unroll
                for (int ix = 0; ix < W; ix++) {
                    const T wv = w[0][ix] * wyz;
                    unsafeAtomicAdd(r + ix, wv * v0);
                    unsafeAtomicAdd(r + CELLS + ix, wv * v1);
                    unsafeAtomicAdd(r + 2 * CELLS + ix, wv * v2);
                }
            }
    }
    __syncthreads();
    const unsigned N = (unsigned)g.N, lx = threadIdx.x & (g.TX - 1);
    const size_t M = (size_t)N * N * N;
    for (int r = threadIdx.x >> g.lT[0]; r < ROWS; r += blockDim.x >> g.lT[0]) {
        const unsigned ly = r & (g.TY - 1), lz = r >> g.lT[1];
        const unsigned gi = (o3[0] + lx) + N * ((o3[1] + ly) + N * (o3[2] + lz));
        const int li = (int)lx + P + SXD * r;
        out[gi] = acc[li];
        out[gi + M] = acc[li + CELLS];
        out[gi + 2 * M] = acc[li + 2 * CELLS];
    }
}

// FP32 vector deposit in fixed point: the three components' values are scaled so that the largest
// |value| is 2^30 (s_c = 2^30 / max_c |vals_c vscale|, from k_absmax3), every term w v s_c is rounded to an
// integer (|term| <= 2^30, so the sum of up to 2^33 terms fits in 64 bits) and summed by 64-bit LDS integer
// atomics, which run several times faster than float ones; the store divides by s_c.  The sum is
// independent of the order of the additions; each term carries its float product's rounding plus at most
// 2^-31 max_c.  Block (tile, component): one 64-bit accumulator per padded cell (the layout above).
__global__ __launch_bounds__(256) void k_absmax3(const float *__restrict__ vals, size_t np, float vscale,
                                                 unsigned *__restrict__ mx) {
    __shared__ float red[3][256];
    float m[3] = {0.f, 0.f, 0.f};
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < np; p += (size_t)gridDim.x * 256)
#pragma unroll
        for (int c = 0; c < 3; c++) m[c] = fmaxf(m[c], fabsf(vals[p + c * np] * vscale));
#pragma unroll
    for (int c = 0; c < 3; c++) red[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] = fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + h]);
        __syncthreads();
    }
    // non-negative floats order as their bit patterns
    if (threadIdx.x < 3) atomicMax(&mx[threadIdx.x], __float_as_uint(red[threadIdx.x][0]));
}

__global__ void k_zero3(unsigned *mx) {
    if (threadIdx.x < 3) mx[threadIdx.x] = 0u;
}

__device__ __forceinline__ float fx_scale(unsigned mbits) {
    const float m = __uint_as_float(mbits);
    return m > 0.f ? 1073741824.0f / m : 0.f;
}

template <int W>
__global__ __launch_bounds__(512) void k_tile_deposit3q(const Entry<float> *__restrict__ ent,
                                                        const unsigned *__restrict__ start,
                                                        const unsigned *__restrict__ gstart,
                                                        const unsigned *__restrict__ cnt, TileGeom g, float sx, float sy,
                                                        float sz, float L, const float *__restrict__ vals, size_t np,
                                                        float vscale, const unsigned *__restrict__ mx,
                                                        float *__restrict__ out) {
    extern __shared__ unsigned char smem[];
    unsigned long long *acc = (unsigned long long *)smem;
    constexpr int P = W - 1;
    const int t = blockIdx.x, comp = blockIdx.y;
    const int SXD = g.TX + 2 * P, ROWS = g.TY * g.TZ, CELLS = SXD * ROWS;
    const int SZ = SXD * g.TY, END = SXD * ROWS;
    for (int c = threadIdx.x; c < CELLS; c += blockDim.x) acc[c] = 0ull;
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    const int o3[3] = {tx << g.lT[0], ty << g.lT[1], tz << g.lT[2]};
    const float s = fx_scale(mx[comp]);
    const float *vc = vals + (size_t)comp * np;
    __syncthreads();
    const unsigned n_own = cnt[t], n_all = n_own + cnt[g.nt + t], s_own = start[t], s_gh = gstart[t] - n_own;
    for (unsigned i = threadIdx.x; i < n_all; i += blockDim.x) {
        const Entry<float> en = ent[i < n_own ? s_own + i : s_gh + i];
        const float vs = vc[en.idx] * vscale * s;
        const float xs[3] = {en.x + sx, en.y + sy, en.z + sz};
        float w[3][W];
        int l0[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float gg = tjmod(xs[a], L) / L * (float)g.N;
            const int b = twbase<W>(gg);
#pragma unroll
            for (int o = 0; o < W; o++) w[a][o] = twk_at<W>(o, fabsf(gg - (float)(b + twlo<W>() + o)));
            int l = b + twlo<W>() - o3[a];
            l0[a] = l >= (g.N >> 1) ? l - g.N : (l < -(g.N >> 1) ? l + g.N : l);
        }
#pragma unroll
        for (int o = 0; o < W; o++) w[0][o] *= vs;
        const bool okx = l0[0] >= -P && l0[0] < g.TX;
        const int xo = okx ? l0[0] + P : 0;
        constexpr int BIG = 1 << 28;
        int ly[W], lz[W];
#pragma unroll
        for (int o = 0; o < W; o++) {
            const int y = l0[1] + o, z = l0[2] + o;
            ly[o] = (unsigned)y < (unsigned)g.TY ? __mul24(y, SXD) : BIG;
            lz[o] = (okx && (unsigned)z < (unsigned)g.TZ) ? __mul24(z, SZ) : BIG;
        }
#pragma unroll
        for (int iz = 0; iz < W; iz++)
#pragma unroll
            for (int iy = 0; iy < W; iy++) {
                const int ro = ly[iy] + lz[iz];
                if (ro >= END) continue;  // a row outside the tile
                unsigned long long *r = acc + (xo + ro);
                const float wyz = w[1][iy] * w[2][iz];
#pragma unroll
                for (int ix = 0; ix < W; ix++) {
                    const long long q = (long long)(int)rintf(w[0][ix] * wyz);
                    atomicAdd(r + ix, (unsigned long long)q);
                }
            }
    }
    __syncthreads();
    const unsigned N = (unsigned)g.N, lx = threadIdx.x & (g.TX - 1);
    const size_t M = (size_t)N * N * N;
    const double inv = s > 0.f ? 1.0 / (double)s : 0.0;
    float *oc = out + (size_t)comp * M;
    for (int r = threadIdx.x >> g.lT[0]; r < ROWS; r += blockDim.x >> g.lT[0]) {
        const unsigned ly = r & (g.TY - 1), lz = r >> g.lT[1];
        const unsigned gi = (o3[0] + lx) + N * ((o3[1] + ly) + N * (o3[2] + lz));
        oc[gi] = (float)((double)(long long)acc[(int)lx + P + SXD * r] * inv);
    }
}

// ------------------------------------------------------------------------------------ gather
// -(c1 d1 + c2 d2 + c3 d3) of a line of potential values v[0 .. W+2H) at stencil point o (index o + H):
// the reference's finite-difference gradient at one cell (global-memory path)
template <typename T, int ORDER>
__device__ __forceinline__ T fd_line(const T *v, int i, T c1, T c2, T c3) {
    const T d1 = v[i + 1] - v[i - 1];
    if constexpr (ORDER == 2) return -(c1 * d1);
    const T d2 = v[i + 2] - v[i - 2];
    if constexpr (ORDER == 4) return -(c1 * d1 + c2 * d2);
    const T d3 = v[i + 3] - v[i - 3];
    return -(c1 * d1 + c2 * d2 + c3 * d3);
}

// Sums over the W^3 stencil cells of (-D_a phi) w for the three components, one line of W + 2H values
// live at a time (x rows for F_x, y columns for F_y, z columns for F_z); ld(x, y, z) returns phi at
// stencil-relative cell (x, y, z), the stencil's first cell at H.
template <typename T, int W, int ORDER, typename LD>
__device__ __forceinline__ void gather_sums(const LD &ld, const T (&w)[3][W], T c1, T c2, T c3, T &s1, T &s2, T &s3) {
    constexpr int H = ORDER / 2, Q = W + 2 * H;
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++) {
            T v[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) v[q] = ld(q, oy + H, oz + H);
#pragma unroll
            for (int ox = 0; ox < W; ox++) s1 += fd_line<T, ORDER>(v, ox + H, c1, c2, c3) * (w[0][ox] * w[1][oy] * w[2][oz]);
        }
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int ox = 0; ox < W; ox++) {
            T v[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) v[q] = ld(ox + H, q, oz + H);
#pragma unroll
            for (int oy = 0; oy < W; oy++) s2 += fd_line<T, ORDER>(v, oy + H, c1, c2, c3) * (w[0][ox] * w[1][oy] * w[2][oz]);
        }
#pragma unroll
    for (int oy = 0; oy < W; oy++)
#pragma unroll
        for (int ox = 0; ox < W; ox++) {
            T v[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) v[q] = ld(ox + H, oy + H, q);
#pragma unroll
            for (int oz = 0; oz < W; oz++) s3 += fd_line<T, ORDER>(v, oz + H, c1, c2, c3) * (w[0][ox] * w[1][oy] * w[2][oz]);
        }
}

// The finite difference and the interpolation along a line folded into one filter:
// sum_o w[o] fd(v, o + H) = sum_q u[q] v[q] with u[o+H+t] -= w[o] c_t, u[o+H-t] += w[o] c_t.
// A line then costs Q fused multiply-adds instead of W finite differences; the value equals the
// per-cell form up to rounding (the LDS path).
template <typename T, int W, int ORDER>
__device__ __forceinline__ void fd_filter(const T (&w)[W], T c1, T c2, T c3, T (&u)[W + ORDER]) {
    constexpr int H = ORDER / 2, Q = W + 2 * H;
    const T ct[3] = {c1, c2, c3};
#pragma unroll
    for (int q = 0; q < Q; q++) u[q] = 0;
#pragma unroll
    for (int o = 0; o < W; o++)
#pragma unroll
        for (int t = 1; t <= H; t++) {
            u[o + H + t] = fma(-w[o], ct[t - 1], u[o + H + t]);
            u[o + H - t] = fma(w[o], ct[t - 1], u[o + H - t]);
        }
}

// The filtered sums with every staged cell read once.  F_x, F_y and F_z read the union of an x
// box (Q x W x W cells), a y box (W x Q x W) and a z box (W x W x Q), which share the W^3 core.
// The loop runs over the Q x slices: each slice's W x W core block feeds F_x (weighted by the
// x filter's tap), and the W core slices also read their y and z columns for F_y and F_z.  That is
// Q W^2 + 4 H W^2 reads per particle (189 for TSC at order 6) instead of 3 Q W^2 (243).
template <typename T, int W, int ORDER, typename LD>
__device__ __forceinline__ void gather_sums_union(const LD &ld, const T (&w)[3][W], T c1, T c2, T c3, T &s1, T &s2,
                                                  T &s3) {
    constexpr int H = ORDER / 2, Q = W + 2 * H;
    T u1[Q], u2[Q], u3[Q], wyz[W][W];
    fd_filter<T, W, ORDER>(w[0], c1, c2, c3, u1);
    fd_filter<T, W, ORDER>(w[1], c1, c2, c3, u2);
    fd_filter<T, W, ORDER>(w[2], c1, c2, c3, u3);
#pragma unroll
    for (int oz = 0; oz < W; oz++)
#pragma unroll
        for (int oy = 0; oy < W; oy++) wyz[oz][oy] = w[1][oy] * w[2][oz];
#pragma unroll
    for (int x = 0; x < Q; x++) {
        T core[W][W];  // [oz][oy]
#pragma unroll
        for (int oz = 0; oz < W; oz++)
#pragma unroll
            for (int oy = 0; oy < W; oy++) core[oz][oy] = ld(x, oy + H, oz + H);
        T b = 0;
#pragma unroll
        for (int oz = 0; oz < W; oz++)
#pragma unroll
            for (int oy = 0; oy < W; oy++) b = fma(wyz[oz][oy], core[oz][oy], b);
        s1 = fma(u1[x], b, s1);
        if (x >= H && x < H + W) {
            const int ox = x - H;
            // F_y: y columns at (x, ., oz + H); F_z: z columns at (x, oy + H, .)
            T ty = 0, tz = 0;
#pragma unroll
            for (int oz = 0; oz < W; oz++) {
                T l = 0;
#pragma unroll
                for (int q = 0; q < Q; q++)
                    l = fma(u2[q], (q >= H && q < H + W) ? core[oz][q - H] : ld(x, q, oz + H), l);
                ty = fma(w[2][oz], l, ty);
            }
#pragma unroll
            for (int oy = 0; oy < W; oy++) {
                T l = 0;
#pragma unroll
                for (int q = 0; q < Q; q++)
                    l = fma(u3[q], (q >= H && q < H + W) ? core[q - H][oy] : ld(x, oy + H, q), l);
                tz = fma(w[1][oy], l, tz);
            }
            s2 = fma(w[0][ox], ty, s2);
            s3 = fma(w[0][ox], tz, s3);
        }
    }
}

// workgroup size: FP32 tiles hold ~512 particles, one per thread; FP64 PCS needs 256 threads to keep
// its stencil in registers
template <typename T, int W> constexpr int gather_threads() { return sizeof(T) == 4 ? 512 : (W <= 3 ? 512 : 256); }
// waves per SIMD the register allocation must allow: two FP32 workgroups per CU (128 VGPRs)
template <typename T, int W> constexpr int gather_min_waves() { return sizeof(T) == 4 && W <= 3 ? 4 : 1; }

// a particle's tile entry and the state its gather mode reads
template <typename T, int MODE>
__device__ __forceinline__ void gather_fetch(const Entry<T> *__restrict__ ent, unsigned e, const T *psi_in,
                                             const T *__restrict__ mom, const T *__restrict__ acc, size_t np, int ns,
                                             int lacc, Entry<T> &en, T (&m3)[3], T (&p3)[3], T (&a3)[3]) {
    en = ent[e];
    const size_t p = (size_t)en.idx;
    if constexpr (MODE == GATHER_KICK) {
        // acc (the earlier passes' sum) is read after the stencil, to keep registers for it
#pragma unroll
        for (int c = 0; c < 3; c++) {
            m3[c] = mom[p + c * np];
            p3[c] = psi_in[p + c * np];
        }
    } else if constexpr (MODE == GATHER_ACC) {
        const size_t ai = lacc ? 3 * (size_t)e : p, as = lacc ? 1 : np;
#pragma unroll
        for (int c = 0; c < 3; c++) a3[c] = acc[ai + c * as];
    }
}

// The hardware hands workgroup b to XCD b % 8.  Tile (b / 8) + (b % 8) (nt / 8) keeps runs of
// consecutive tiles (x neighbours, then the next row of tiles) on one XCD, so the halo that
// neighbouring tiles share is fetched into that XCD's L2 once.
__device__ __forceinline__ int xcd_tile(int b, int nt) { return (nt & 7) ? b : (b >> 3) + (b & 7) * (nt >> 3); }

// Gather order: tiles go brick by brick, 2 x 4 x 8 tiles (64^3 cells for 32 x 16 x 8 tiles, about
// as many workgroups as an XCD runs at once), x fastest inside a brick, and each brick on one XCD.
// The staged regions of a tile reach its neighbours along every axis; in a brick the neighbours
// along z are in flight together too, and their shared halo planes are fetched into the XCD's L2
// once instead of once per tile.  When the brick count is a multiple of 8, XCD x takes bricks x,
// x + 8, x + 16, ... spread over the whole mesh, so that clustered regions (long tile lists) load
// every XCD alike; otherwise each XCD takes a contiguous eighth of the tiles.  Tile counts that the
// bricks do not divide keep the row order.
constexpr int BRICK_X = 2, BRICK_Y = 4, BRICK_Z = 8;
__device__ __forceinline__ void gather_tile(int b, const TileGeom &g, int &tx, int &ty, int &tz) {
    constexpr int BV = BRICK_X * BRICK_Y * BRICK_Z;
    const int loc = xcd_tile(b, g.nt);
    if (g.ntx % BRICK_X || g.nty % BRICK_Y || g.ntz % BRICK_Z) {
        tx = loc % g.ntx;
        ty = (loc / g.ntx) % g.nty;
        tz = loc / (g.ntx * g.nty);
        return;
    }
    int w = loc % BV, br = loc / BV;
    if ((g.nt / BV) % 8 == 0) {
        // XCD b % 8 takes bricks b % 8, b % 8 + 8, ...: its k-th workgroup (k = b / 8) takes tile
        // k % BV of its brick k / BV, so that every XCD's bricks spread over the whole mesh
        const int k = b >> 3;
        w = k % BV;
        br = (k / BV) * 8 + (b & 7);
    }
    const int nbx = g.ntx / BRICK_X, nby = g.nty / BRICK_Y;
    tx = (br % nbx) * BRICK_X + w % BRICK_X;
    ty = ((br / nbx) % nby) * BRICK_Y + (w / BRICK_X) % BRICK_Y;
    tz = (br / (nbx * nby)) * BRICK_Z + w / (BRICK_X * BRICK_Y);
}

// Staged region of a tile's gather, shared by the kernel and its launcher: cells
// [tile + RL + H left, tile + RS + H right] along each axis, the x extent widened to whole 16-byte
// groups (4 floats, 2 doubles) that start at a multiple of the group width, so that the region is
// staged with 16-byte LDS DMA and no group straddles the periodic x edge.  A region is staged only
// when it is smaller than the mesh (no periodic self-overlap).
struct StageGeom {
    int xl;        // cells left of the tile's first x cell
    int s[3];      // extents (x padded to whole groups)
    int lo[3];     // cells left of the tile's first cell per axis (xl for x)
    bool ok;
};
template <typename T>
__host__ __device__ __forceinline__ StageGeom stage_geom_of(int TX, int TY, int TZ, int RL, int RS, int N, int H) {
    constexpr int CPG = 16 / (int)sizeof(T);
    StageGeom q;
    q.xl = ((RL + H + CPG - 1) / CPG) * CPG;
    q.s[0] = ((q.xl + TX + RS + H + CPG - 1) / CPG) * CPG;
    q.s[1] = TY + RL + RS + 2 * H;
    q.s[2] = TZ + RL + RS + 2 * H;
    q.lo[0] = q.xl;
    q.lo[1] = q.lo[2] = RL + H;
    q.ok = q.s[0] < N && q.s[1] < N && q.s[2] < N && N % CPG == 0;
    return q;
}
template <typename T>
__host__ __device__ __forceinline__ StageGeom stage_geom(const TileGeom &g, int H) {
    return stage_geom_of<T>(g.TX, g.TY, g.TZ, g.RL, g.RS, g.N, H);
}

// The FP32 tiles of 32 x 16 x 8 cells with the reach of the interlacing shifts 0 and half a cell
// (antialias -1, 0, 1, 2): the staged region's extents and strides are then compile-time constants,
// and every LDS read of the gather is one base register plus an immediate offset.  A smaller reach
// (antialias 0) stages a region that contains the one it needs.
constexpr int FIXED_TX = 32, FIXED_TY = 16, FIXED_TZ = 8;
template <int W> constexpr int fixed_rl() { return W == 2 ? 0 : 1; }
template <int W> constexpr int fixed_rs() { return W == 4 ? 3 : 2; }
template <int W> constexpr bool fixed_geom_fits(const TileGeom &g) {
    return g.TX == FIXED_TX && g.TY == FIXED_TY && g.TZ == FIXED_TZ && g.RL <= fixed_rl<W>() && g.RS <= fixed_rs<W>();
}
template <typename T, int W, bool FG>
__host__ __device__ __forceinline__ StageGeom gather_geom(const TileGeom &g, int H) {
    if constexpr (FG) return stage_geom_of<T>(FIXED_TX, FIXED_TY, FIXED_TZ, fixed_rl<W>(), fixed_rs<W>(), g.N, H);
    else return stage_geom<T>(g, H);
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Staging of the region (S0 x S1 x S2 cells, x rows contiguous in LDS at stride S0) straight from
// global memory into LDS: global_load_lds of 16 bytes writes each active lane's group at the wave's
// LDS base + 16 lane, so the region is read as a flat array of 16-byte groups, 64 consecutive ones
// per wave instruction (rows may change within an instruction).  Each lane tracks its (group, row,
// plane) by increments and wraps its own source cell periodically; the arithmetic is per lane (vector
// ALU), which keeps the CU's single scalar unit free.
template <typename T, int NT>
__device__ __forceinline__ void stage_groups(const T *phi, T *ph, const int (&r0)[3], const int (&S)[3], int N, int lane,
                                             int wave) {
    constexpr int NW = NT / 64, CPG = 16 / (int)sizeof(T);
    const int GR = S[0] / CPG;                 // groups per row
    const int total = GR * S[1] * S[2];
    int f = wave * 64 + lane;                  // this lane's group of the current instruction
    int row = f / GR, cg = f - row * GR;
    int ly = row % S[1], lz = row / S[1];
    const int step = NW * 64, dcg = step % GR, drow = step / GR;
    unsigned *dst0 = (unsigned *)ph;
    if (r0[0] >= 0 && r0[0] + S[0] <= N && r0[1] >= 0 && r0[1] + S[1] <= N && r0[2] >= 0 && r0[2] + S[2] <= N) {
        // the region does not wrap: the lane's source offset advances by the same increments as its
        // (group, row, plane), with one carry from group to row and one from row to plane (the row
        // step drow splits into dly < S1 rows and dlz planes)
        const int dly = drow % S[1], dlz = drow / S[1];
        const unsigned NN = (unsigned)N * (unsigned)N;
        const unsigned inc = (unsigned)(dcg * CPG) + (unsigned)dly * (unsigned)N + (unsigned)dlz * NN;
        const unsigned carry_x = (unsigned)N - (unsigned)(GR * CPG), carry_y = NN - (unsigned)S[1] * (unsigned)N;
        unsigned off = (unsigned)(r0[0] + cg * CPG) + (unsigned)N * (unsigned)(r0[1] + ly) +
                       NN * (unsigned)(r0[2] + lz);  // < N^3 <= 2^30 cells
        for (int c = wave; c * 64 < total; c += NW) {
            if (f < total)
                __builtin_amdgcn_global_load_lds((glb_void_t *)(phi + off), (lds_void_t *)(dst0 + (size_t)c * 256), 16,
                                                 0, 0);
            f += step;
            off += inc;
            cg += dcg;
            int dr = dly;
            if (cg >= GR) { cg -= GR; dr++; off += carry_x; }
            ly += dr;
            if (ly >= S[1]) { ly -= S[1]; off += carry_y; }
        }
        return;
    }
    for (int c = wave; c * 64 < total; c += NW) {
        if (f < total) {
            const unsigned gx = (unsigned)wrapn(r0[0] + cg * CPG, N), gy = (unsigned)wrapn(r0[1] + ly, N),
                           gz = (unsigned)wrapn(r0[2] + lz, N);
            const T *src = phi + (gx + (unsigned)N * (gy + (unsigned)N * gz));  // < N^3 <= 2^30 cells
            __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)(dst0 + (size_t)c * 256), 16, 0, 0);
        }
        f += step;
        cg += dcg;
        int dr = drow;
        if (cg >= GR) { cg -= GR; dr++; }
        ly += dr;
        while (ly >= S[1]) { ly -= S[1]; lz++; }
    }
}

template <typename T, int W, int ORDER, int MODE, bool FG>
__global__ __launch_bounds__((gather_threads<T, W>()), (gather_min_waves<T, W>())) void k_tile_gather(const Entry<T> *__restrict__ ent, const unsigned *__restrict__ start,
                                                      const unsigned *__restrict__ own, TileGeom g, T sx, T sy, T sz,
                                                      T L, const T *__restrict__ phi, T c1, T c2, T c3, const T *psi_in,
                                                      T *psi_out, T *__restrict__ mom, T *__restrict__ acc, size_t np,
                                                      T alpha, T beta, T dd2, T dd1n, int has_next, int ns, int lacc) {
    constexpr int H = ORDER / 2, Q = W + 2 * H, NT = gather_threads<T, W>();
    extern __shared__ unsigned char smem[];
    T *ph = (T *)smem;
    int tx, ty, tz;
    gather_tile(blockIdx.x, g, tx, ty, tz);
    const int t = tx + g.ntx * (ty + g.nty * tz);
    // staged region: tile + stencil reach + FD halo H (gather_geom)
    const StageGeom q = gather_geom<T, W, FG>(g, H);
    const int lT[3] = {FG ? 5 : g.lT[0], FG ? 4 : g.lT[1], FG ? 3 : g.lT[2]};
    const int r0[3] = {(tx << lT[0]) - q.lo[0], (ty << lT[1]) - q.lo[1], (tz << lT[2]) - q.lo[2]};
    const int S[3] = {q.s[0], q.s[1], q.s[2]};
    const int N = g.N, SX = S[0], SXY = S[0] * S[1];
    const bool region_ok = q.ok;

    // this thread's first particle and its state, loaded before the staging so both are in flight together
    const unsigned e0 = start[t], e1 = e0 + own[t];
    unsigned e = e0 + threadIdx.x;
    Entry<T> en{};
    T m3[3] = {0, 0, 0}, p3[3] = {0, 0, 0}, a3[3] = {0, 0, 0};
    if (e < e1) gather_fetch<T, MODE>(ent, e, psi_in, mom, acc, np, ns, lacc, en, m3, p3, a3);

    if (region_ok) {
        const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        stage_groups<T, NT>(phi, ph, r0, S, N, lane, wave);
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();

    const T xsh[3] = {sx, sy, sz};
    while (e < e1) {
        const T xe[3] = {en.x, en.y, en.z};
        T w[3][W];
        int c0[3], lb[3];
        bool staged = region_ok;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T gg = tjmod(xe[a] + xsh[a], L) / L * (T)N;
            const int b = twbase<W>(gg);
#pragma unroll
            for (int o = 0; o < W; o++) w[a][o] = twk_at<W>(o, fabs(gg - (T)(b + twlo<W>() + o)));
            c0[a] = b + twlo<W>() - H;  // first cell of the stencil's FD neighbourhood (unwrapped)
            lb[a] = wrapn(c0[a] - r0[a], N);
            staged &= lb[a] + Q <= S[a];
        }
        T s1 = 0, s2 = 0, s3 = 0;
        if (staged) {
            const int base = lb[0] + SX * lb[1] + SXY * lb[2];
            auto ld = [&](int x, int y, int z) { return ph[base + x + SX * y + SXY * z]; };
            gather_sums_union<T, W, ORDER>(ld, w, c1, c2, c3, s1, s2, s3);
        } else {
            auto ld = [&](int x, int y, int z) {
                return phi[(size_t)wrapn(c0[0] + x, N) +
                           (size_t)N * ((size_t)wrapn(c0[1] + y, N) + (size_t)N * wrapn(c0[2] + z, N))];
            };
            gather_sums<T, W, ORDER>(ld, w, c1, c2, c3, s1, s2, s3);
        }
        const size_t p = (size_t)en.idx;
        const T f3[3] = {s1, s2, s3};
        // acc of this particle: (3, Np) in particle order, or (entries, 3) in list order between the passes
        const size_t ai = lacc ? 3 * (size_t)e : p, as = lacc ? 1 : np;
        if constexpr (MODE == GATHER_ACC) {
#pragma unroll
            for (int c = 0; c < 3; c++) acc[ai + c * as] = a3[c] + f3[c];
        } else if constexpr (MODE == GATHER_ASSIGN) {
#pragma unroll
            for (int c = 0; c < 3; c++) acc[ai + c * as] = f3[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                T ac = f3[c];
                if (ns > 1) ac = (acc[ai + c * as] + ac) / (T)ns;
                T m = alpha * m3[c] + beta * ac;
                T s = p3[c] + dd2 * m;
                if (has_next) s = s + dd1n * m;
                mom[p + c * np] = m;
                psi_out[p + c * np] = s;
            }
        }
        e += NT;
        if (e < e1) gather_fetch<T, MODE>(ent, e, psi_in, mom, acc, np, ns, lacc, en, m3, p3, a3);
    }
}

// Gather that marches along z.  A workgroup takes a segment of G consecutive tiles of one column (fixed
// tx, ty) and keeps the staged potential in an LDS ring of S2 planes: tile t + 1's region is tile t's
// shifted by TZ planes, so after tile t only its TZ new planes are staged, into the ring slots of the TZ
// planes tile t + 1 no longer needs (TZ of S2 = TZ + RL + RS + 2H planes instead of S2: 8 of 17 for TSC
// at order 6).  A particle's Q planes sit at ring slots (sb + lb_z + z) mod S2, so each of its z planes
// gets one base register and every LDS read is that base plus an immediate offset, as in the staged
// gather above.  The segments are dealt so that the workgroups an XCD (workgroup b % 8) runs together
// are the same z segment of a 4 x 4 patch of neighbouring columns, whose x and y halos then meet in that
// XCD's L2; the XCDs' patches interleave over the mesh, which spreads clustered regions over all of them.
// Used for the FP32 32 x 16 x 8 tiles with the compile-time staging geometry; the arithmetic per particle
// is the staged gather's (same sums, same order).
struct MarchMap {
    int G;          // tiles per segment (divides ntz)
    bool patches;   // 4 x 4 column patches dealt round-robin to the XCDs
};
__host__ __device__ __forceinline__ MarchMap march_map(const TileGeom &g) {
    MarchMap m;
    m.G = (g.ntz % 16 == 0) ? 16 : (g.ntz % 8 == 0) ? 8 : (g.ntz % 4 == 0) ? 4 : (g.ntz % 2 == 0) ? 2 : 1;
    m.patches = g.ntx % 4 == 0 && g.nty % 4 == 0 && ((g.ntx / 4) * (g.nty / 4)) % 8 == 0;
    return m;
}
__device__ __forceinline__ void march_segment(int b, const TileGeom &g, const MarchMap &m, int &tx, int &ty, int &tz0) {
    const int ncol = g.ntx * g.nty;
    int col, seg;
    if (m.patches) {
        const int xcd = b & 7, k = b >> 3;
        const int per = ncol / 8;  // columns per XCD: (ncol / 16) / 8 patches of 16
        seg = k / per;
        const int r = k % per, pi = r >> 4, ci = r & 15;
        const int gp = pi * 8 + xcd, npx = g.ntx / 4;
        tx = (gp % npx) * 4 + (ci & 3);
        ty = (gp / npx) * 4 + (ci >> 2);
        tz0 = seg * m.G;
        return;
    }
    col = b % ncol;
    seg = b / ncol;
    tx = col % g.ntx;
    ty = col / g.ntx;
    tz0 = seg * m.G;
}

template <typename T, int W, int ORDER, int MODE>
__global__ __launch_bounds__((gather_threads<T, W>()), (gather_min_waves<T, W>())) void k_tile_gather_march(
    const Entry<T> *__restrict__ ent, const unsigned *__restrict__ start, const unsigned *__restrict__ own, TileGeom g,
    T sx, T sy, T sz, T L, const T *__restrict__ phi, T c1, T c2, T c3, const T *psi_in, T *psi_out,
    T *__restrict__ mom, T *__restrict__ acc, size_t np, T alpha, T beta, T dd2, T dd1n, int has_next, int ns,
    int lacc) {
    constexpr int H = ORDER / 2, Q = W + 2 * H, NT = gather_threads<T, W>();
    extern __shared__ unsigned char smem[];
    T *ph = (T *)smem;
    const StageGeom q = gather_geom<T, W, true>(g, H);
    constexpr int TXL = 5, TYL = 4, TZ = FIXED_TZ;
    const int S[3] = {q.s[0], q.s[1], q.s[2]};
    const int N = g.N, SX = S[0], SXY = S[0] * S[1], R = S[2];  // R: ring planes
    const MarchMap m = march_map(g);
    int tx, ty, tz0;
    march_segment(blockIdx.x, g, m, tx, ty, tz0);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rx = (tx << TXL) - q.lo[0], ry = (ty << TYL) - q.lo[1];
    const int zs0 = tz0 * TZ - q.lo[2];  // first staged plane of the segment (ring slot 0)
    const T xsh[3] = {sx, sy, sz};

    for (int it = 0; it < m.G; it++) {
        const int tz = tz0 + it;
        const int t = tx + g.ntx * (ty + g.nty * tz);
        const int r0[3] = {rx, ry, tz * TZ - q.lo[2]};
        const unsigned e0 = start[t], e1 = e0 + own[t];
        unsigned e = e0 + threadIdx.x;
        Entry<T> en{};
        T m3[3] = {0, 0, 0}, p3[3] = {0, 0, 0}, a3[3] = {0, 0, 0};
        if (e < e1) gather_fetch<T, MODE>(ent, e, psi_in, mom, acc, np, ns, lacc, en, m3, p3, a3);
        // stage: the whole region for the first tile, else the TZ planes after the previous region,
        // into the ring slots (zg - zs0) mod R, split where the slots wrap
        {
            const int zfirst = it == 0 ? r0[2] : r0[2] + S[2] - TZ;
            const int nplanes = it == 0 ? S[2] : TZ;
            const int slot = (zfirst - zs0) % R;
            const int n1 = min(nplanes, R - slot);
            const int ra[3] = {rx, ry, zfirst}, sa[3] = {S[0], S[1], n1};
            stage_groups<T, NT>(phi, ph + (size_t)slot * SXY, ra, sa, N, lane, wave);
            if (n1 < nplanes) {
                const int rb[3] = {rx, ry, zfirst + n1}, sb2[3] = {S[0], S[1], nplanes - n1};
                stage_groups<T, NT>(phi, ph, rb, sb2, N, lane, wave);
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        const int sbase = (r0[2] - zs0) % R;  // ring slot of this tile's first staged plane
        while (e < e1) {
            const T xe[3] = {en.x, en.y, en.z};
            T w[3][W];
            int c0[3], lb[3];
            bool staged = true;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const T gg = tjmod(xe[a] + xsh[a], L) / L * (T)N;
                const int b = twbase<W>(gg);
#pragma unroll
                for (int o = 0; o < W; o++) w[a][o] = twk_at<W>(o, fabs(gg - (T)(b + twlo<W>() + o)));
                c0[a] = b + twlo<W>() - H;
                lb[a] = wrapn(c0[a] - r0[a], N);
                staged &= lb[a] + Q <= S[a];
            }
            T s1 = 0, s2 = 0, s3 = 0;
            if (staged) {
                // plane z of the stencil at ring slot sl0 + z, wrapping once, at plane zw: base A for the
                // planes before it, B = A - R SXY after it, the rest an immediate offset
                int sl0 = sbase + lb[2];
                sl0 = sl0 >= R ? sl0 - R : sl0;
                const int A = lb[0] + SX * lb[1] + sl0 * SXY, B = A - R * SXY, zw = R - sl0;
                auto ld = [&](int x, int y, int z) { return ph[(z >= zw ? B : A) + x + SX * y + SXY * z]; };
                gather_sums_union<T, W, ORDER>(ld, w, c1, c2, c3, s1, s2, s3);
            } else {
                auto ld = [&](int x, int y, int z) {
                    return phi[(size_t)wrapn(c0[0] + x, N) +
                               (size_t)N * ((size_t)wrapn(c0[1] + y, N) + (size_t)N * wrapn(c0[2] + z, N))];
                };
                gather_sums<T, W, ORDER>(ld, w, c1, c2, c3, s1, s2, s3);
            }
            const size_t p = (size_t)en.idx;
            const T f3[3] = {s1, s2, s3};
            const size_t ai = lacc ? 3 * (size_t)e : p, as = lacc ? 1 : np;
            if constexpr (MODE == GATHER_ACC) {
#pragma unroll
                for (int c = 0; c < 3; c++) acc[ai + c * as] = a3[c] + f3[c];
            } else if constexpr (MODE == GATHER_ASSIGN) {
#pragma unroll
                for (int c = 0; c < 3; c++) acc[ai + c * as] = f3[c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    T ac = f3[c];
                    if (ns > 1) ac = (acc[ai + c * as] + ac) / (T)ns;
                    T mm = alpha * m3[c] + beta * ac;
                    T sp = p3[c] + dd2 * mm;
                    if (has_next) sp = sp + dd1n * mm;
                    mom[p + c * np] = mm;
                    psi_out[p + c * np] = sp;
                }
            }
            e += NT;
            if (e < e1) gather_fetch<T, MODE>(ent, e, psi_in, mom, acc, np, ns, lacc, en, m3, p3, a3);
        }
        // every wave is done with this tile's first TZ planes before the next tile's staging overwrites them
        __syncthreads();
    }
}

// The kick and drifts of a step in particle order (run_nbody.jl:156-158), the last interlacing pass having
// added its force into the list-ordered hand-over: the state is read and written in order, and each
// particle's summed force (12 contiguous bytes) is fetched through its own-list entry.  The arithmetic is
// the kick gather's, in its order.
template <typename T>
__global__ __launch_bounds__(256) void k_kick_listed(const unsigned *__restrict__ own_at, const T *__restrict__ accl,
                                                     T *__restrict__ psi, T *__restrict__ mom, size_t np, T alpha,
                                                     T beta, T dd2, T dd1n, int has_next, int ns) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const size_t e = 3 * (size_t)own_at[p];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const T ac = accl[e + c] / (T)ns;
        const T mm = alpha * mom[p + c * np] + beta * ac;
        T sp = psi[p + c * np] + dd2 * mm;
        if (has_next) sp = sp + dd1n * mm;
        mom[p + c * np] = mm;
        psi[p + c * np] = sp;
    }
}

// ------------------------------------------------------------------------------------ launchers
template <typename T>
struct TileLaunch {
    template <int W>
    static void bin_w(const TileArgs &a, hipStream_t st) {
        dim3 grid((a.n + 15) / 16, (a.n + 3) / 4, (a.n + 3) / 4);
        k_tile_bin<T, W, false><<<grid, 256, 0, st>>>((const T *)a.psi, (const T *)a.q1, a.n, a.shifts, a.ns, (T)a.L,
                                                      a.g, a.cnt, a.start, a.gstart, a.cur, a.flag, (Entry<T> *)a.ent,
                                                      a.ent_cap, a.own_at);
    }
    template <int W>
    static void fill_w(const TileArgs &a, hipStream_t st) {
        dim3 grid((a.n + 15) / 16, (a.n + 3) / 4, (a.n + 3) / 4);
        k_tile_bin<T, W, true><<<grid, 256, 0, st>>>((const T *)a.psi, (const T *)a.q1, a.n, a.shifts, a.ns, (T)a.L,
                                                     a.g, a.cnt, a.start, a.gstart, a.cur, a.flag, (Entry<T> *)a.ent,
                                                     a.ent_cap, a.own_at);
    }
    static void count(const TileArgs &a, int W, hipStream_t st) {
        if (W == 2) bin_w<2>(a, st); else if (W == 3) bin_w<3>(a, st); else bin_w<4>(a, st);
    }
    static void scan(const TileArgs &a, const unsigned *src, int slack, hipStream_t st) {
        const int nb = (a.g.nt + SCAN_TILES - 1) / SCAN_TILES;  // <= 1024 (nt <= 2^20)
        k_scan_sums<<<nb, 256, 0, st>>>(src, a.g.nt, slack, a.scan_part);
        k_scan_top<<<1, 1024, 0, st>>>(a.scan_part, nb, a.start, a.g.nt, a.flag);
        k_scan_write<<<nb, 256, 0, st>>>(src, a.g.nt, slack, a.scan_part, a.start, a.gstart, a.cur);
    }
    static void fill(const TileArgs &a, int W, hipStream_t st) {
        if (W == 2) fill_w<2>(a, st); else if (W == 3) fill_w<3>(a, st); else fill_w<4>(a, st);
    }
    template <int W, int ACC>
    static void dep_w(const TileArgs &a, const double *sh, void *mesh, double fscale, hipStream_t st) {
        const TileGeom &g = a.g;
        const bool pad = g.ntx >= 2 && g.nty >= 2 && g.ntz >= 2;
        const size_t cells = pad ? (size_t)(g.TX + 2 * (W - 1)) * (g.TY * g.TZ) : (size_t)g.TX * g.TY * g.TZ;
        const size_t shm = cells * (ACC == ACC_T ? sizeof(T) : 8);
        auto kern = pad ? k_tile_deposit<T, W, ACC, true> : k_tile_deposit<T, W, ACC, false>;
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        // 512 threads: the four workgroups that the accumulators' LDS allows per CU then hold the
        // CU's full 32 waves, twice as many LDS atomics in flight as with 256 (2.97 -> 2.56 ms at
        // 512^3 on 1024^3, TSC)
        kern<<<g.nt, 512, shm, st>>>((const Entry<T> *)a.ent, a.start, a.gstart, a.cur, g, (T)sh[0], (T)sh[1],
                                     (T)sh[2], (T)a.L, mesh, fscale);
    }
    template <int W>
    static void dep_sel(const TileArgs &a, bool fixed, const double *sh, void *mesh, double fscale, hipStream_t st) {
        constexpr int FLOAT_ACC = sizeof(T) == 4 ? ACC_Q31 : ACC_T;
        if (fixed) dep_w<W, ACC_FIXED>(a, sh, mesh, fscale, st);
        else dep_w<W, FLOAT_ACC>(a, sh, mesh, fscale, st);
    }
    template <int W>
    static void dep3_w(const TileArgs &a, const double *sh, const void *vals, double scale, void *out, hipStream_t st) {
        const TileGeom &g = a.g;
        const size_t shm = 3 * (size_t)(g.TX + 2 * (W - 1)) * (g.TY * g.TZ) * sizeof(T);
        (void)hipFuncSetAttribute((const void *)k_tile_deposit3<T, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
        k_tile_deposit3<T, W><<<g.nt, 512, shm, st>>>((const Entry<T> *)a.ent, a.start, a.gstart, a.cur, g, (T)sh[0],
                                                      (T)sh[1], (T)sh[2], (T)a.L, (const T *)vals,
                                                      (size_t)a.n * a.n * a.n, (T)scale, (T *)out);
    }
    template <int W>
    static void dep3q_w(const TileArgs &a, const double *sh, const void *vals, double scale, void *out, void *work,
                        hipStream_t st) {
        const TileGeom &g = a.g;
        const size_t np = (size_t)a.n * a.n * a.n;
        unsigned *mx = (unsigned *)work;
        k_zero3<<<1, 64, 0, st>>>(mx);
        const unsigned nb = (unsigned)std::min<size_t>(2048, (np + 255) / 256);
        k_absmax3<<<nb, 256, 0, st>>>((const float *)vals, np, (float)scale, mx);
        const size_t shm = (size_t)(g.TX + 2 * (W - 1)) * (g.TY * g.TZ) * 8;
        (void)hipFuncSetAttribute((const void *)k_tile_deposit3q<W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
        k_tile_deposit3q<W><<<dim3(g.nt, 3), 512, shm, st>>>((const Entry<float> *)a.ent, a.start, a.gstart, a.cur, g,
                                                             (float)sh[0], (float)sh[1], (float)sh[2], (float)a.L,
                                                             (const float *)vals, np, (float)scale, mx, (float *)out);
    }
    static bool deposit3(const TileArgs &a, int W, const double *sh, const void *vals, double scale, void *out,
                         void *work, hipStream_t st) {
        const TileGeom &g = a.g;
        if (g.ntx < 2 || g.nty < 2 || g.ntz < 2) return false;
        if constexpr (sizeof(T) == 4) {
            if (W == 2) dep3q_w<2>(a, sh, vals, scale, out, work, st);
            else if (W == 3) dep3q_w<3>(a, sh, vals, scale, out, work, st);
            else dep3q_w<4>(a, sh, vals, scale, out, work, st);
            return true;
        }
        if ((size_t)3 * (g.TX + 2 * (W - 1)) * (g.TY * g.TZ) * sizeof(T) > 160 * 1024) return false;
        if (W == 2) dep3_w<2>(a, sh, vals, scale, out, st);
        else if (W == 3) dep3_w<3>(a, sh, vals, scale, out, st);
        else dep3_w<4>(a, sh, vals, scale, out, st);
        return true;
    }
    static void deposit(const TileArgs &a, int W, bool fixed, const double *sh, void *mesh, double fscale,
                        hipStream_t st) {
        if (W == 2) dep_sel<2>(a, fixed, sh, mesh, fscale, st);
        else if (W == 3) dep_sel<3>(a, fixed, sh, mesh, fscale, st);
        else dep_sel<4>(a, fixed, sh, mesh, fscale, st);
    }
    template <int W, int O, int M, bool FG>
    static void gat_womf(const TileArgs &a, const double *sh, const void *phi, const double *c, const KickArgs &k,
                         void *psi_out, hipStream_t st) {
        constexpr int H = O / 2;
        const StageGeom q = gather_geom<T, W, FG>(a.g, H);
        // a region that is not staged (as large as the mesh) is read from global memory by the kernel
        size_t shm = q.ok ? sizeof(T) * (size_t)q.s[0] * q.s[1] * q.s[2] : sizeof(T);
        auto kern = k_tile_gather<T, W, O, M, FG>;
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        kern<<<a.g.nt, gather_threads<T, W>(), shm, st>>>((const Entry<T> *)a.ent, a.start, a.cur, a.g, (T)sh[0], (T)sh[1], (T)sh[2],
                                        (T)a.L, (const T *)phi, (T)c[0], (T)c[1], (T)c[2], (const T *)a.psi,
                                        (T *)psi_out, (T *)a.mom, (T *)a.acc, (size_t)a.n * a.n * a.n, (T)k.alpha,
                                        (T)k.beta, (T)k.dd2, (T)k.dd1_next, k.has_next, k.nshifts, a.acc_by_entry);
    }
    template <int W, int O, int M>
    static void gat_march(const TileArgs &a, const double *sh, const void *phi, const double *c, const KickArgs &k,
                          void *psi_out, hipStream_t st) {
        constexpr int H = O / 2;
        const StageGeom q = gather_geom<T, W, true>(a.g, H);
        const size_t shm = sizeof(T) * (size_t)q.s[0] * q.s[1] * q.s[2];
        const MarchMap mm = march_map(a.g);
        auto kern = k_tile_gather_march<T, W, O, M>;
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        kern<<<a.g.nt / mm.G, gather_threads<T, W>(), shm, st>>>(
            (const Entry<T> *)a.ent, a.start, a.cur, a.g, (T)sh[0], (T)sh[1], (T)sh[2], (T)a.L, (const T *)phi,
            (T)c[0], (T)c[1], (T)c[2], (const T *)a.psi, (T *)psi_out, (T *)a.mom, (T *)a.acc,
            (size_t)a.n * a.n * a.n, (T)k.alpha, (T)k.beta, (T)k.dd2, (T)k.dd1_next, k.has_next, k.nshifts,
            a.acc_by_entry);
    }
    template <int W, int O, int M>
    static void gat_wom(const TileArgs &a, const double *sh, const void *phi, const double *c, const KickArgs &k,
                        void *psi_out, hipStream_t st) {
        if constexpr (sizeof(T) == 4) {
            if (fixed_geom_fits<W>(a.g) && gather_geom<T, W, true>(a.g, O / 2).ok)
                return gat_march<W, O, M>(a, sh, phi, c, k, psi_out, st);
            if (fixed_geom_fits<W>(a.g)) return gat_womf<W, O, M, true>(a, sh, phi, c, k, psi_out, st);
        }
        gat_womf<W, O, M, false>(a, sh, phi, c, k, psi_out, st);
    }
    template <int W, int O>
    static void gat_wo(const TileArgs &a, const double *sh, const void *phi, const double *c, int mode,
                       const KickArgs &k, void *psi_out, hipStream_t st) {
        if (mode == GATHER_ACC) gat_wom<W, O, GATHER_ACC>(a, sh, phi, c, k, psi_out, st);
        else if (mode == GATHER_ASSIGN) gat_wom<W, O, GATHER_ASSIGN>(a, sh, phi, c, k, psi_out, st);
        else gat_wom<W, O, GATHER_KICK>(a, sh, phi, c, k, psi_out, st);
    }
    template <int W>
    static void gat_w(const TileArgs &a, int order, const double *sh, const void *phi, const double *c, int mode,
                      const KickArgs &k, void *psi_out, hipStream_t st) {
        if (order == 2) gat_wo<W, 2>(a, sh, phi, c, mode, k, psi_out, st);
        else if (order == 4) gat_wo<W, 4>(a, sh, phi, c, mode, k, psi_out, st);
        else gat_wo<W, 6>(a, sh, phi, c, mode, k, psi_out, st);
    }
    static void gather(const TileArgs &a, int W, int order, const double *sh, const void *phi, const double *c,
                       int mode, const KickArgs &k, void *psi_out, hipStream_t st) {
        if (W == 2) gat_w<2>(a, order, sh, phi, c, mode, k, psi_out, st);
        else if (W == 3) gat_w<3>(a, order, sh, phi, c, mode, k, psi_out, st);
        else gat_w<4>(a, order, sh, phi, c, mode, k, psi_out, st);
    }
};

template <typename T>
static void kick_listed_l(const unsigned *own_at, const void *acc_list, void *psi, void *mom, size_t np,
                          const KickArgs &k, hipStream_t st) {
    k_kick_listed<T><<<(unsigned)((np + 255) / 256), 256, 0, st>>>(own_at, (const T *)acc_list, (T *)psi, (T *)mom, np,
                                                                   (T)k.alpha, (T)k.beta, (T)k.dd2, (T)k.dd1_next,
                                                                   k.has_next, k.nshifts);
}

template <typename T> void tile_ops(TileOps &ops) {
    ops.kick_listed = kick_listed_l<T>;
    ops.count = TileLaunch<T>::count;
    ops.scan = TileLaunch<T>::scan;
    ops.fill = TileLaunch<T>::fill;
    ops.deposit = TileLaunch<T>::deposit;
    ops.deposit3 = TileLaunch<T>::deposit3;
    ops.gather = TileLaunch<T>::gather;
    ops.entry_bytes = sizeof(Entry<T>);
}

template void tile_ops<float>(TileOps &);
template void tile_ops<double>(TileOps &);

}  // namespace dj
