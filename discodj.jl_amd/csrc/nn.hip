// Learned force correction: a residual 3-D convolution net on the PM mesh, BF16 weights and
// activations, FP32 accumulation on MFMA (v_mfma_f32_16x16x16bf16_1k).  There is no counterpart in the
// reference; the net adds a potential correction to the PM potential before the gradient and the gather.
//
//   x0 = (log(1 + max(delta, -0.999)), a)            delta = mesh / rho_bar - 1, a = scale factor
//   z1 = W1 * x0 + b1,  h1 = relu(z1)
//   z2 = W2 * h1 + b2,  h2 = h1 + relu(z2)
//   z3 = W3 * h2 + b3,  h3 = h2 + relu(z3)
//   dphi = s_out (W4 * h3 + b4)                       (*: periodic 3x3x3 convolution, 16 channels)
//
// Tensors are channel-last: t[v * 16 + c], v = x + N (y + N z), BF16 (round to nearest even).  Every
// layer is stored as W[27][16][16] (tap, out, in) padded with zeros (layer 1 has 2 inputs, layer 4 one
// output).  One wave computes 16 output channels of 16 voxels of an x row: per tap one MFMA with
// A = the tap's weights (16 out x 16 in) and B = the 16 voxels' input channels; the transposed
// convolution of the backward pass is the same kernel with the flipped, transposed taps.
#include <hip/hip_runtime.h>

#include "dj_internal.h"

namespace dj {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// FP32 -> BF16 round to nearest even (finite values); BF16 -> FP32 exact
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ float nn_input(float dens, float inv_rho) {
    const float d = dens * inv_rho - 1.0f;
    return log1pf(fmaxf(d, -0.999f));
}

__device__ __forceinline__ int nn_wrap(int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); }

enum { NN_IN_H = 0, NN_IN_FIELD = 1, NN_IN_GRAD = 2 };  // B: 16-channel tensor / density field / 1-channel float
enum { NN_EPI_HIDDEN = 0, NN_EPI_OUT = 1, NN_EPI_BWD = 2 };

// B operand of one tap: channels 4 (l/16) .. +3 of voxel v
template <int IN>
__device__ __forceinline__ s4v nn_bop(const unsigned short *hin, const float *fin, float inv_rho, float cval, size_t v,
                                      int cg) {
    s4v b = {0, 0, 0, 0};
    if constexpr (IN == NN_IN_H) {
        b = *(const s4v *)(hin + v * 16 + cg);
    } else if (cg == 0) {
        // density field: the net's input transform; gradient field: scaled by inv_rho
        const float f = IN == NN_IN_FIELD ? nn_input(fin[v], inv_rho) : fin[v] * inv_rho;
        b[0] = (short)f2bf(f);
        b[1] = (short)f2bf(cval);
    }
    return b;
}

// One block marches along z over a 64 (x) x 8 (y) footprint, 16 planes: the input planes z-1, z, z+1
// (with a one-voxel x-y halo, 66 x 10 voxels) sit in a three-slot LDS ring, so each input voxel is
// read from memory about 1.3 times instead of once per tap.  Two blocks share a CU (77 KB of LDS
// each), so one stages while the other computes.  Wave w computes y rows 2w, 2w+1 of the plane (8
// segments of 16 voxels), per tap one MFMA per segment with B read from LDS (8 bytes per lane,
// conflict-free) and A (the tap's weights) from LDS.
constexpr int NN_TX = 64, NN_TY = 8, NN_TZ = 16, NN_PX = NN_TX + 2, NN_PY = NN_TY + 2, NN_PV = NN_PX * NN_PY;
constexpr int NN_CHUNKS = NN_PV * 2;                       // 16-byte chunks of a 16-channel plane
constexpr int NN_PER = (NN_CHUNKS + 255) / 256;            // chunks per thread

// stage input plane zz (wrapped) of the block's footprint into an LDS slot as 16-channel BF16; a field
// input becomes (bf16(input), bf16(cval), 0, ...)
template <int IN>
__device__ __forceinline__ void nn_plane_stage(unsigned short *slot, const unsigned short *hin, const float *fin, int N,
                                               int x0, int y0, int zz, float inv_rho, float cval) {
    const size_t zb = (size_t)N * N * nn_wrap(zz, N);
    if constexpr (IN == NN_IN_H) {
#pragma unroll
        for (int k = 0; k < NN_PER; k++) {
            const int ch = threadIdx.x + 256 * k;
            if (ch < NN_CHUNKS) {
                const int vox = ch >> 1, xi = vox % NN_PX, yi = vox / NN_PX;
                const size_t v = zb + (size_t)nn_wrap(x0 - 1 + xi, N) + (size_t)N * nn_wrap(y0 - 1 + yi, N);
                *(uint4 *)(slot + ch * 8) = *(const uint4 *)(hin + v * 16 + (ch & 1) * 8);
            }
        }
    } else {
        // a field input has two channels, (input, cval): one 4-byte pair per voxel
        const unsigned c1 = (unsigned)f2bf(cval) << 16;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int vox = threadIdx.x + 256 * k;
            if (vox < NN_PV) {
                const int xi = vox % NN_PX, yi = vox / NN_PX;
                const float fv = fin[zb + (size_t)nn_wrap(x0 - 1 + xi, N) + (size_t)N * nn_wrap(y0 - 1 + yi, N)];
                const float x = IN == NN_IN_FIELD ? nn_input(fv, inv_rho) : fv * inv_rho;
                ((unsigned *)slot)[vox] = (unsigned)f2bf(x) | c1;
            }
        }
    }
}

template <int IN, int EPI>
__global__ __launch_bounds__(256, 2) void k_nn_conv(int N, const unsigned short *__restrict__ hin,
                                                    const float *__restrict__ fin, float inv_rho, float cval,
                                                    const s4v *__restrict__ wsw, const float *__restrict__ bias,
                                                    const unsigned short *__restrict__ res,
                                                    const unsigned short *__restrict__ zmask,
                                                    unsigned short *__restrict__ out_z, unsigned short *__restrict__ out_h,
                                                    float *__restrict__ out_f, float oscale) {
    extern __shared__ unsigned char nn_smem[];
    s4v *wl = (s4v *)nn_smem;                                           // 27 x 64 lane-ordered A operands
    unsigned short *ring = (unsigned short *)(nn_smem + 27 * 64 * 8);   // 3 x NN_PV x 16 BF16
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, cg = 4 * (lane >> 4), lx = lane & 15;
    const int nbx = N / NN_TX, nby = N / NN_TY;
    const int x0 = (blockIdx.x % nbx) * NN_TX, y0 = ((blockIdx.x / nbx) % nby) * NN_TY,
              z0 = (blockIdx.x / (nbx * nby)) * NN_TZ;
    for (int i = threadIdx.x; i < (IN == NN_IN_H ? 27 : 4) * 64; i += 256) wl[i] = wsw[i];
    auto slot = [&](int z) { return ring + (size_t)(((z % 3) + 3) % 3) * NN_PV * 16; };
    nn_plane_stage<IN>(slot(z0 - 1), hin, fin, N, x0, y0, z0 - 1, inv_rho, cval);
    nn_plane_stage<IN>(slot(z0), hin, fin, N, x0, y0, z0, inv_rho, cval);
    for (int z = z0; z < z0 + NN_TZ; z++) {
        nn_plane_stage<IN>(slot(z + 1), hin, fin, N, x0, y0, z + 1, inv_rho, cval);
        __syncthreads();  // planes z-1, z, z+1 (and the weights) in LDS
        f4v acc[8];
#pragma unroll
        for (int sg = 0; sg < 8; sg++) acc[sg] = {0, 0, 0, 0};
        if constexpr (IN == NN_IN_H) {
#pragma unroll 1
        for (int dz = -1; dz <= 1; dz++) {
            const unsigned short *sh = slot(z + dz);
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const s4v a = wl[((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)) * 64 + lane];
#pragma unroll
                    for (int sg = 0; sg < 8; sg++) {
                        const int yi = 2 * wave + (sg >> 2) + dy + 1, xi = (sg & 3) * 16 + lx + dx + 1;
                        const s4v b = *(const s4v *)(sh + (yi * NN_PX + xi) * 16 + cg);
                        acc[sg] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[sg], 0, 0, 0);
                    }
                }
        }
        } else {
            // two input channels: one MFMA takes 8 taps x 2 channels (K = 16), 4 MFMAs for the 27 taps
            // (the 5 extra taps have zero weights); lane k-group 4 (lane / 16) + j reads taps
            // 8 q + 2 (lane / 16) + j / 2, channel j % 2
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const s4v a = wl[q * 64 + lane];
                const int t1 = min(8 * q + 2 * (lane >> 4), 26), t2 = min(t1 + 1, 26);
                const unsigned *s1 = (const unsigned *)slot(z + t1 / 9 - 1), *s2 = (const unsigned *)slot(z + t2 / 9 - 1);
                const int d1 = ((t1 / 3) % 3) * NN_PX + t1 % 3, d2 = ((t2 / 3) % 3) * NN_PX + t2 % 3;
#pragma unroll
                for (int sg = 0; sg < 8; sg++) {
                    const int v0 = (2 * wave + (sg >> 2)) * NN_PX + (sg & 3) * 16 + lx;  // the (-1,-1) corner cell
                    const unsigned u1 = s1[v0 + d1], u2 = s2[v0 + d2];
                    const s4v b = {(short)(u1 & 0xffffu), (short)(u1 >> 16), (short)(u2 & 0xffffu), (short)(u2 >> 16)};
                    acc[sg] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[sg], 0, 0, 0);
                }
            }
        }
        // epilogue: lane holds D[cg + j][lx] of each segment: channels cg..cg+3 of one voxel
#pragma unroll
        for (int sg = 0; sg < 8; sg++) {
            const int y = y0 + 2 * wave + (sg >> 2), x = x0 + (sg & 3) * 16 + lx;
            const size_t v = (size_t)x + (size_t)N * ((size_t)y + (size_t)N * z);
            if constexpr (EPI == NN_EPI_OUT) {
                if (cg == 0) out_f[v] = oscale * (acc[sg][0] + bias[0]);
            } else {
                s4v zo, ho;
                s4v rr = {0, 0, 0, 0}, m = {0, 0, 0, 0};
                if constexpr (IN == NN_IN_H) {
                    if (res == hin)  // the residual is the layer's input: its centre cell is in the LDS ring
                        rr = *(const s4v *)(slot(z) + ((2 * wave + (sg >> 2) + 1) * NN_PX + (sg & 3) * 16 + lx + 1) * 16 + cg);
                    else if (res)
                        rr = *(const s4v *)(res + v * 16 + cg);
                } else if (res) {
                    rr = *(const s4v *)(res + v * 16 + cg);
                }
                if constexpr (EPI == NN_EPI_BWD) m = *(const s4v *)(zmask + v * 16 + cg);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float rv = bf2f((unsigned short)rr[j]);
                    if constexpr (EPI == NN_EPI_HIDDEN) {
                        const float zz = acc[sg][j] + bias[cg + j];
                        zo[j] = (short)f2bf(zz);
                        ho[j] = (short)f2bf(rv + fmaxf(zz, 0.0f));
                    } else {
                        const float g = rv + acc[sg][j];
                        ho[j] = (short)f2bf(g);
                        zo[j] = (short)f2bf(bf2f((unsigned short)m[j]) > 0.0f ? g : 0.0f);
                    }
                }
                if (out_z) *(s4v *)(out_z + v * 16 + cg) = zo;  // pre-activations: kept for training only
                *(s4v *)(out_h + v * 16 + cg) = ho;
            }
        }
        __syncthreads();  // slot (z - 1) is overwritten by plane z + 2 next
    }
}

// One block's partial weight and bias gradients (27 x 16 x 16 + 16 floats), and the slices of blocks the
// first reduction pass sums
constexpr int NN_WP = 27 * 16 * 16 + 16, NN_WSLICES = 16, NN_WBLOCKS = 2048;

// Weight gradient dW[t][co][ci] += sum_v gz(v, co) x(v + off(t), ci) and db[co] += sum_v gz(v, co).
// A block takes rows of 64 voxels (x chunks), stages gz (64 x 16) and the 9 neighbouring input rows
// (66 x 16 each, periodic) in LDS, and each of its 4 waves accumulates the MFMAs of taps w, w+4, ...
// over the voxels (K = 16 voxels per MFMA) in registers.  Each block writes its sums to its row of part,
// and k_nn_wreduce1/2 add the rows in a fixed order: no atomics, so a training run repeats bit for bit.
template <int IN, int GA>
__global__ __launch_bounds__(256) void k_nn_wgrad(int N, const unsigned short *__restrict__ hin, const float *__restrict__ fin,
                                                  float inv_rho, float cval, const unsigned short *__restrict__ gz,
                                                  const float *__restrict__ gf, float gscale, float *__restrict__ part) {
    __shared__ unsigned short sA[64 * 16];
    __shared__ unsigned short sB[9 * 66 * 16];
    __shared__ float sdbp[4][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int chunks = N / 64, nrows = chunks * N * N;
    float dbacc = 0;  // threads < 16: channel threadIdx.x's bias gradient over the block's rows, in row order
    f4v acc[7];
#pragma unroll
    for (int i = 0; i < 7; i++) acc[i] = {0, 0, 0, 0};
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int x0 = (r % chunks) * 64, y = (r / chunks) % N, z = r / (chunks * N);
        __syncthreads();  // previous row's LDS reads are done
        // the previous row's four partial bias sums, in order (written before that row's MFMAs)
        if (r != (int)blockIdx.x && threadIdx.x < 16)
            dbacc += ((sdbp[0][threadIdx.x] + sdbp[1][threadIdx.x]) + sdbp[2][threadIdx.x]) + sdbp[3][threadIdx.x];
        // gz of the row: 64 voxels x 16 channels (GA: a 1-channel float field in channel 0, scaled)
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            const int vx = e >> 4, c = e & 15;
            const size_t v = (size_t)(x0 + vx) + (size_t)N * ((size_t)y + (size_t)N * z);
            unsigned short g;
            if constexpr (GA) g = c == 0 ? f2bf(gscale * gf[v]) : (unsigned short)0;
            else g = gz[v * 16 + c];
            sA[e] = g;
        }
        // the input rows (y + dy, z + dz), x from x0 - 1 to x0 + 64
        for (int e = threadIdx.x; e < 9 * 66 * 16; e += 256) {
            const int c = e & 15, xi = (e >> 4) % 66, rr = (e >> 4) / 66;
            const int yy = nn_wrap(y + rr % 3 - 1, N), zz = nn_wrap(z + rr / 3 - 1, N), xx = nn_wrap(x0 + xi - 1, N);
            const size_t v = (size_t)xx + (size_t)N * ((size_t)yy + (size_t)N * zz);
            unsigned short h;
            if constexpr (IN == NN_IN_H) h = hin[v * 16 + c];
            else h = c == 0 ? f2bf(nn_input(fin[v], inv_rho)) : (c == 1 ? f2bf(cval) : (unsigned short)0);
            sB[e] = h;
        }
        __syncthreads();
        if (threadIdx.x < 64) {  // bias: one channel per 4 threads, 16 voxels each
            const int c = threadIdx.x & 15, q = threadIdx.x >> 4;
            float s = 0;
            for (int vx = q * 16; vx < q * 16 + 16; vx++) s += bf2f(sA[vx * 16 + c]);
            sdbp[q][c] = s;
        }
        const int cg = 4 * (lane >> 4), cl = lane & 15;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            s4v a;
#pragma unroll
            for (int j = 0; j < 4; j++) a[j] = (short)sA[(kc * 16 + cg + j) * 16 + cl];
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const int t = wave + 4 * i;
                if (t < 27) {
                    const int dx = t % 3 - 1, rr = t / 3;  // rr = (dz + 1) 3 + (dy + 1)
                    s4v b;
#pragma unroll
                    for (int j = 0; j < 4; j++) b[j] = (short)sB[(rr * 66 + kc * 16 + cg + j + dx + 1) * 16 + cl];
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i], 0, 0, 0);
                }
            }
        }
    }
    // D[co = cg + j][ci = lane & 15] of tap t: each element is one lane's
    float *mine = part + (size_t)blockIdx.x * NN_WP;
    const int cg = 4 * (lane >> 4), cl = lane & 15;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const int t = wave + 4 * i;
        if (t < 27) {
#pragma unroll
            for (int j = 0; j < 4; j++) mine[(t * 16 + cg + j) * 16 + cl] = acc[i][j];
        }
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        if ((int)blockIdx.x < nrows)  // the last row's partial sums
            dbacc += ((sdbp[0][threadIdx.x] + sdbp[1][threadIdx.x]) + sdbp[2][threadIdx.x]) + sdbp[3][threadIdx.x];
        mine[27 * 16 * 16 + threadIdx.x] = dbacc;
    }
}

// Fixed-order sums of the blocks' rows of part: slice y of the nb rows into part2[y] (each thread sums
// rows g, g + 4, ... of the slice, then the four threads' sums in order), then the slices in order into
// dW (the first 27 x 16 x 16 values) and db
__global__ __launch_bounds__(256) void k_nn_wreduce1(const float *__restrict__ part, int nb, float *__restrict__ part2) {
    __shared__ float s4[4][64];
    const int o = threadIdx.x & 63, g = threadIdx.x >> 6, i = blockIdx.x * 64 + o;
    const int per = (nb + NN_WSLICES - 1) / NN_WSLICES, b0 = blockIdx.y * per, b1 = min(nb, b0 + per);
    float a = 0;
    if (i < NN_WP)
        for (int b = b0 + g; b < b1; b += 4) a += part[(size_t)b * NN_WP + i];
    s4[g][o] = a;
    __syncthreads();
    if (g == 0 && i < NN_WP) part2[(size_t)blockIdx.y * NN_WP + i] = ((s4[0][o] + s4[1][o]) + s4[2][o]) + s4[3][o];
}
__global__ __launch_bounds__(256) void k_nn_wreduce2(const float *__restrict__ part2, float *__restrict__ dW,
                                                     float *__restrict__ db) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NN_WP) return;
    float a = 0;
    for (int sl = 0; sl < NN_WSLICES; sl++) a += part2[(size_t)sl * NN_WP + i];
    if (i < 27 * 16 * 16) dW[i] += a;
    else db[i - 27 * 16 * 16] += a;
}

// lane-ordered BF16 A operands of the forward (W[t][co][ci] at lane (co, ci / 4)) and of the
// transposed convolution (W[26 - t][co][ci] at lane (ci, co / 4)) of one layer
// packed A operands of a two-channel input (8 taps x 2 channels per K = 16): forward A[co][k] =
// W[t][co][ch], transposed A[ci][k] = W[26 - t][ch][ci], t = 8 q + 2 (k / 4) + (k % 4) / 2, ch = k % 2,
// zero for t >= 27
__global__ void k_nn_swizzle_packed(const float *__restrict__ W, s4v *__restrict__ fwd, s4v *__restrict__ bwd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * 64) return;
    const int q = i / 64, l = i % 64, row = l & 15;
    s4v f, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int t = 8 * q + 2 * (l >> 4) + j / 2, ch = j % 2;
        f[j] = (short)(t < 27 ? f2bf(W[(t * 16 + row) * 16 + ch]) : 0);
        b[j] = (short)(t < 27 ? f2bf(W[((26 - t) * 16 + ch) * 16 + row]) : 0);
    }
    fwd[i] = f;
    bwd[i] = b;
}

__global__ void k_nn_swizzle(const float *__restrict__ W, s4v *__restrict__ fwd, s4v *__restrict__ bwd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27 * 64) return;
    const int t = i / 64, l = i % 64;
    s4v f, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        f[j] = (short)f2bf(W[(t * 16 + (l & 15)) * 16 + 4 * (l >> 4) + j]);
        b[j] = (short)f2bf(W[((26 - t) * 16 + 4 * (l >> 4) + j) * 16 + (l & 15)]);
    }
    fwd[i] = f;
    bwd[i] = b;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; mask zeroes the padding
__global__ void k_nn_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                          const float *__restrict__ mask, int n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * mask[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// out += s in  (the correction into the potential); r = acc - target and its squared sum
__global__ void k_nn_axpy(float *__restrict__ out, const float *__restrict__ in, size_t n, float s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += s * in[i];
}
// Each of the NN_LOSS_BLOCKS blocks sums its grid-stride share of r^2 in a fixed order into sum[block], and
// k_nn_loss_total adds the block sums in a fixed order into sum[NN_LOSS_BLOCKS]: the loss repeats bit for bit
constexpr int NN_LOSS_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_nn_residual(const float *__restrict__ acc, const float *__restrict__ tgt,
                                                     float *__restrict__ abar, size_t n, float gs,
                                                     double *__restrict__ sum) {
    __shared__ double red[256];
    double s = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)NN_LOSS_BLOCKS * 256) {
        const float r = acc[i] - tgt[i];
        abar[i] = gs * r;
        s += (double)r * r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sum[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_nn_loss_total(double *sum) {
    __shared__ double red[256];
    double s = 0;
    for (int i = threadIdx.x; i < NN_LOSS_BLOCKS; i += 256) s += sum[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sum[NN_LOSS_BLOCKS] = red[0];
}

static unsigned ngrid(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct NnLaunch {
    static void conv(int in, int epi, int N, const void *hin, const void *fin, double inv_rho, double cval,
                     const void *wsw, const void *bias, const void *res, const void *zmask, void *out_z, void *out_h,
                     void *out_f, double oscale, hipStream_t st) {
        const dim3 g((N / NN_TX) * (N / NN_TY) * (N / NN_TZ));
        const size_t shm = (size_t)27 * 64 * 8 + (size_t)3 * NN_PV * 16 * 2;
#define NN_C(I, E)                                                                                                   \
    do {                                                                                                             \
    (void)hipFuncSetAttribute((const void *)k_nn_conv<I, E>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);   \
    k_nn_conv<I, E><<<g, 256, shm, st>>>(N, (const unsigned short *)hin, (const float *)fin, (float)inv_rho, (float)cval, \
                                       (const s4v *)wsw, (const float *)bias, (const unsigned short *)res,              \
                                       (const unsigned short *)zmask, (unsigned short *)out_z, (unsigned short *)out_h, \
                                       (float *)out_f, (float)oscale); } while (0)
        if (in == NN_IN_FIELD && epi == NN_EPI_HIDDEN) NN_C(NN_IN_FIELD, NN_EPI_HIDDEN);
        else if (in == NN_IN_H && epi == NN_EPI_HIDDEN) NN_C(NN_IN_H, NN_EPI_HIDDEN);
        else if (in == NN_IN_H && epi == NN_EPI_OUT) NN_C(NN_IN_H, NN_EPI_OUT);
        else if (in == NN_IN_GRAD && epi == NN_EPI_BWD) NN_C(NN_IN_GRAD, NN_EPI_BWD);
        else NN_C(NN_IN_H, NN_EPI_BWD);
#undef NN_C
    }
    static void wgrad(int in, int ga, int N, const void *hin, const void *fin, double inv_rho, double cval, const void *gz,
                      const void *gf, double gscale, void *dW, void *db, void *part, hipStream_t st) {
        const int nrows = (N / 64) * N * N, blocks = nrows < NN_WBLOCKS ? nrows : NN_WBLOCKS;
        float *p1 = (float *)part, *p2 = p1 + (size_t)NN_WBLOCKS * NN_WP;
#define NN_W(I, G)                                                                                                     \
    k_nn_wgrad<I, G><<<blocks, 256, 0, st>>>(N, (const unsigned short *)hin, (const float *)fin, (float)inv_rho,        \
                                             (float)cval, (const unsigned short *)gz, (const float *)gf, (float)gscale, p1)
        if (in == NN_IN_FIELD) NN_W(NN_IN_FIELD, 0);
        else if (ga) NN_W(NN_IN_H, 1);
        else NN_W(NN_IN_H, 0);
#undef NN_W
        k_nn_wreduce1<<<dim3(ngrid(NN_WP, 64), NN_WSLICES), 256, 0, st>>>(p1, blocks, p2);
        k_nn_wreduce2<<<ngrid(NN_WP, 256), 256, 0, st>>>(p2, (float *)dW, (float *)db);
    }
    static size_t wgrad_scratch_bytes() { return sizeof(float) * (size_t)NN_WP * (NN_WBLOCKS + NN_WSLICES); }
    static void swizzle(const void *W, void *fwd, void *bwd, void *pfwd, void *pbwd, hipStream_t st) {
        k_nn_swizzle<<<ngrid(27 * 64, 256), 256, 0, st>>>((const float *)W, (s4v *)fwd, (s4v *)bwd);
        k_nn_swizzle_packed<<<1, 256, 0, st>>>((const float *)W, (s4v *)pfwd, (s4v *)pbwd);
    }
    static void adam(void *p, const void *g, void *m, void *v, const void *mask, int n, double lr, double b1, double b2,
                     double eps, double c1, double c2, hipStream_t st) {
        k_nn_adam<<<ngrid(n, 256), 256, 0, st>>>((float *)p, (const float *)g, (float *)m, (float *)v,
                                                 (const float *)mask, n, (float)lr, (float)b1, (float)b2, (float)eps,
                                                 (float)c1, (float)c2);
    }
    static void axpy(void *out, const void *in, size_t n, double s, hipStream_t st) {
        k_nn_axpy<<<ngrid(n, 256), 256, 0, st>>>((float *)out, (const float *)in, n, (float)s);
    }
    static void residual(const void *acc, const void *tgt, void *abar, size_t n, double gs, void *sum, hipStream_t st) {
        k_nn_residual<<<NN_LOSS_BLOCKS, 256, 0, st>>>((const float *)acc, (const float *)tgt, (float *)abar, n, (float)gs,
                                                     (double *)sum);
        k_nn_loss_total<<<1, 256, 0, st>>>((double *)sum);
    }
    static size_t residual_sum_doubles() { return NN_LOSS_BLOCKS + 1; }
};

void nn_ops(NnOps &ops) {
    ops.conv = NnLaunch::conv;
    ops.wgrad = NnLaunch::wgrad;
    ops.wgrad_scratch_bytes = NnLaunch::wgrad_scratch_bytes;
    ops.swizzle = NnLaunch::swizzle;
    ops.adam = NnLaunch::adam;
    ops.axpy = NnLaunch::axpy;
    ops.residual = NnLaunch::residual;
    ops.residual_sum_doubles = NnLaunch::residual_sum_doubles;
}

}  // namespace dj
