// Learned warm start: physics features, BF16 3x3-conv U-Net kernels (MFMA
// implicit GEMM forward, MFMA weight gradient), pooling/upsampling, loss and
// Adam -- everything the network needs to train and run on the GPU.
//
// Activations are NHWC bf16 [B][H][W][C] with C a multiple of 8.  A 3x3
// "same" convolution with weights w[Cout][Kp] (k = tap*Cin + ci, tap = 3ky+kx,
// bias in column 9*Cin, zero columns up to Kp = roundup(9*Cin+1, 32)) is the
// GEMM  y[p][co] = sum_k w[co][k] * col[p][k]  with col[p][k] the input pixel
// p + (ky-1, kx-1), channel ci (0 outside the image) and col[p][9*Cin] = 1.
// It runs on v_mfma_f32_16x16x32_bf16: A = 16 output channels x 32 k,
// B = 32 k x 16 pixels, D = 16 channels x 16 pixels, FP32 accumulate.
// Lane layout: A[i][k] in lane i + 16*(k/8) element k%8, B[k][j] in lane
// j + 16*(k/8) element k%8, D[i][j] in lane j + 16*(i/4) register i%4.
// The data gradient of a layer is the same kernel with the flipped,
// transposed weights; the weight gradient (bias included through the ones
// column) is a GEMM over pixels with both operands staged transposed in LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "common.h"
#include "fadesim.h"
#include "fadesim_nn.h"
#include "rb_core.h"
#include "nn_core.h"

namespace fadesim {
namespace {

using nn::bf16x8;
using nn::f4;
using nn::zero8;

// ---------------------------------------------------------------- pooled features
// nn::pooled_features for one crossbar per workgroup, with the base field
// (p2, q2) written at full resolution and the scale S2.
using FeatGeo = nn::FeatGeo;

__global__ void __launch_bounds__(FeatGeo::NT) features_pooled_kernel(const float *__restrict__ g,
                                                                     const double *__restrict__ rw,
                                                                     const double *__restrict__ vin,
                                                                     __bf16 *__restrict__ feat,
                                                                     float *__restrict__ base,
                                                                     float *__restrict__ scale)
{
    using G = FeatGeo;
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, M = G::M, N = G::N, NT = G::NT;
    __shared__ double cbuf[2 * G::TY * N];
    __shared__ double cext[3 * N];
    __shared__ int dead;
    __shared__ float red[4 * (NT / 64)];
    rb::Shared sh{};
    sh.cbuf = cbuf;
    sh.cext = cext;
    sh.dead = &dead;
    const int b = blockIdx.x, t = threadIdx.x, tx = t % TX, ty = t / TX;
    const int i0 = TR * ty, j0 = TC * tx;
    const size_t MN = (size_t)M * N;
    float p[TR][TC], q[TR][TC];
    const float S2 = nn::pooled_features(sh, g + (size_t)b * MN, rw[b], vin + (size_t)b * M,
                                         feat + (size_t)b * (M / 4) * (N / 4) * nn::kFeatC, red, p, q);
    if (base) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            rb::strow<TC>(base + (size_t)b * 2 * MN + (size_t)(i0 + ii) * N + j0, p[ii]);
            rb::strow<TC>(base + (size_t)b * 2 * MN + MN + (size_t)(i0 + ii) * N + j0, q[ii]);
        }
    }
    if (t == 0)
        scale[b] = S2;
}

// ---------------------------------------------------------------- features
// One workgroup per crossbar (M, N <= 128).  First-order drops from the ideal
// cell currents I = g c (c = vin_i, bitlines at 0 V):
//   p1_ij = r sum_k (min(j,k)+1) I_ik      (wordline, source at the left)
//   q1_ij = r sum_k (n - max(i,k)) I_kj    (bitline, sink at the bottom)
// second order from I = g (c - p1 - q1).  Features (bf16, 8 channels):
//   p1/s, q1/s, (p2-p1)/s2, (q2-q1)/s2, g/gmax, vin/vmax, r/5, 1
// with s = max(p1, q1) and s2 = max|p2-p1|, |q2-q1| of the instance; base =
// (p2, q2) in FP32 and the output scale s2 per instance.
//
// Data flow: g and the row scans live in LDS (rows padded to N+1 floats, so a
// thread per row and a thread per column both access it conflict-free); the
// column thread computes q1 and, from the same column, q2 -- it needs p1 of
// its column (LDS) -- and leaves the second-order cell current in the g plane;
// the row threads then turn that into p2 in place.  Global traffic (g in, q1 /
// q2 via the base and scratch planes, features out) is coalesced.  Scans
// accumulate in FP64.
constexpr int kFeatMax = 128;
constexpr int kFeatThreads = 256;

__global__ void __launch_bounds__(kFeatThreads) features_kernel(const float *__restrict__ g,
                                                                const double *__restrict__ rw,
                                                                const double *__restrict__ vin, int M, int N,
                                                                float *__restrict__ tmp, __bf16 *__restrict__ feat,
                                                                float *__restrict__ base, float *__restrict__ scale)
{
    __shared__ float gl[kFeatMax * (kFeatMax + 1)]; // g, then the 2nd-order current, then p2
    __shared__ float pl[kFeatMax * (kFeatMax + 1)]; // p1
    __shared__ double vl[kFeatMax];
    __shared__ float red[4][kFeatThreads / 64];
    const int b = blockIdx.x, t = threadIdx.x;
    const int ld = N + 1;
    const size_t MN = (size_t)M * N;
    const float *gb = g + b * MN;
    const double *vb = vin + (size_t)b * M;
    const double r = rw[b];
    float *q1g = tmp + b * MN;             // q1 scratch plane
    float *bp = base + b * 2 * MN, *bq = bp + MN;
    for (size_t id = t; id < MN; id += kFeatThreads) {
        const int i = (int)(id / N), j = (int)(id % N);
        gl[i * ld + j] = gb[id];
    }
    for (int i = t; i < M; i += kFeatThreads)
        vl[i] = vb[i];
    __syncthreads();
    // row pass 1: p1
    if (t < M) {
        const int i = t;
        const double c = vl[i];
        double R = 0.0;
        for (int j = 0; j < N; ++j)
            R += (double)gl[i * ld + j] * c;
        double P0 = 0.0, P1 = 0.0;
        for (int j = 0; j < N; ++j) {
            const double cur = (double)gl[i * ld + j] * c;
            pl[i * ld + j] = (float)(r * (P1 + (double)(j + 1) * (R - P0)));
            P0 += cur;
            P1 += (double)(j + 1) * cur;
        }
    }
    __syncthreads();
    // column pass: q1, then the 2nd-order current of the column and q2
    if (t < N) {
        const int j = t;
        double T = 0.0;
        for (int k = 0; k < M; ++k)
            T += (double)gl[k * ld + j] * vl[k] * (double)(M - k);
        double A = 0.0, Bs = T;
        for (int i = 0; i < M; ++i) {
            const double cur = (double)gl[i * ld + j] * vl[i];
            A += cur;
            Bs -= cur * (double)(M - i);
            const float q1 = (float)(r * ((double)(M - i) * A + Bs));
            q1g[(size_t)i * N + j] = q1;
            // second-order cell current, kept in the g plane for the row pass
            gl[i * ld + j] = (float)((double)gl[i * ld + j] * (vl[i] - (double)pl[i * ld + j] - (double)q1));
        }
        T = 0.0;
        for (int k = 0; k < M; ++k)
            T += (double)gl[k * ld + j] * (double)(M - k);
        A = 0.0;
        Bs = T;
        for (int i = 0; i < M; ++i) {
            const double cur = (double)gl[i * ld + j];
            A += cur;
            Bs -= cur * (double)(M - i);
            bq[(size_t)i * N + j] = (float)(r * ((double)(M - i) * A + Bs));
        }
    }
    __syncthreads();
    // row pass 2: p2 in place of the 2nd-order current
    if (t < M) {
        const int i = t;
        double R = 0.0;
        for (int j = 0; j < N; ++j)
            R += (double)gl[i * ld + j];
        double P0 = 0.0, P1 = 0.0;
        for (int j = 0; j < N; ++j) {
            const double cur = (double)gl[i * ld + j];
            gl[i * ld + j] = (float)(r * (P1 + (double)(j + 1) * (R - P0)));
            P0 += cur;
            P1 += (double)(j + 1) * cur;
        }
    }
    __syncthreads();
    // scales
    float s1 = 0.f, s2 = 0.f, gm = 0.f, vm = 0.f;
    for (size_t id = t; id < MN; id += kFeatThreads) {
        const int i = (int)(id / N), j = (int)(id % N);
        const float p1 = pl[i * ld + j], p2 = gl[i * ld + j];
        const float q1 = q1g[id], q2 = bq[id];
        s1 = fmaxf(s1, fmaxf(fabsf(p1), fabsf(q1)));
        s2 = fmaxf(s2, fmaxf(fabsf(p2 - p1), fabsf(q2 - q1)));
        gm = fmaxf(gm, gb[id]);
    }
    for (int i = t; i < M; i += kFeatThreads)
        vm = fmaxf(vm, (float)fabs(vl[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 = fmaxf(s1, __shfl_xor(s1, off, 64));
        s2 = fmaxf(s2, __shfl_xor(s2, off, 64));
        gm = fmaxf(gm, __shfl_xor(gm, off, 64));
        vm = fmaxf(vm, __shfl_xor(vm, off, 64));
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = s1;
        red[1][t >> 6] = s2;
        red[2][t >> 6] = gm;
        red[3][t >> 6] = vm;
    }
    __syncthreads();
    float S1 = 1e-30f, S2 = 1e-30f, GM = 1e-30f, VM = 1e-30f;
#pragma unroll
    for (int w = 0; w < kFeatThreads / 64; ++w) {
        S1 = fmaxf(S1, red[0][w]);
        S2 = fmaxf(S2, red[1][w]);
        GM = fmaxf(GM, red[2][w]);
        VM = fmaxf(VM, red[3][w]);
    }
    if (t == 0)
        scale[b] = S2;
    for (size_t id = t; id < MN; id += kFeatThreads) {
        const int i = (int)(id / N), j = (int)(id % N);
        const float p1 = pl[i * ld + j], p2 = gl[i * ld + j];
        const float q1 = q1g[id], q2 = bq[id];
        bf16x8 f;
        f[0] = (__bf16)(p1 / S1);
        f[1] = (__bf16)(q1 / S1);
        f[2] = (__bf16)((p2 - p1) / S2);
        f[3] = (__bf16)((q2 - q1) / S2);
        f[4] = (__bf16)(gb[id] / GM);
        f[5] = (__bf16)((float)vl[i] / VM);
        f[6] = (__bf16)((float)(r / 5.0));
        f[7] = (__bf16)1.0f;
        *reinterpret_cast<bf16x8 *>(feat + (b * MN + id) * 8) = f;
        bp[id] = p2;
    }
}

// warm[b][0|1][i][j] = base + scale_b * out[b][i][j][0|1]
__global__ void combine_kernel(const float *__restrict__ base, const float *__restrict__ scale,
                               const __bf16 *__restrict__ out, int Cout, int B, int MN, float *__restrict__ warm)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)B * MN; k += (size_t)gridDim.x * blockDim.x) {
        const size_t b = k / MN, id = k % MN;
        const float s = scale[b];
        warm[b * 2 * MN + id] = base[b * 2 * MN + id] + s * (float)out[k * Cout + 0];
        warm[b * 2 * MN + MN + id] = base[b * 2 * MN + MN + id] + s * (float)out[k * Cout + 1];
    }
}

// warm[b][c][i][j] = base + scale_b * bilinear(out[b][.][.][c]) for a network
// output f = 2^level times coarser than the crossbar (M/f x N/f): sample point
// ((i + 1/2)/f - 1/2, (j + 1/2)/f - 1/2), clamped to the coarse grid.
__global__ void combine_up_kernel(const float *__restrict__ base, const float *__restrict__ scale,
                                  const __bf16 *__restrict__ out, int Cout, int B, int M, int N, int level,
                                  float *__restrict__ warm)
{
    const int f = 1 << level, Hc = M >> level, Wc = N >> level;
    const float inv = 1.0f / (float)f;
    const size_t MN = (size_t)M * N;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)B * MN; k += (size_t)gridDim.x * blockDim.x) {
        const size_t b = k / MN, id = k % MN;
        const int i = (int)(id / N), j = (int)(id % N);
        const float y = fminf(fmaxf(((float)i + 0.5f) * inv - 0.5f, 0.0f), (float)(Hc - 1));
        const float x = fminf(fmaxf(((float)j + 0.5f) * inv - 0.5f, 0.0f), (float)(Wc - 1));
        const int y0 = (int)y, x0 = (int)x;
        const int y1 = min(y0 + 1, Hc - 1), x1 = min(x0 + 1, Wc - 1);
        const float wy = y - (float)y0, wx = x - (float)x0;
        const __bf16 *o = out + b * (size_t)Hc * Wc * Cout;
        const float s = scale[b];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float v00 = (float)o[((size_t)y0 * Wc + x0) * Cout + c], v01 = (float)o[((size_t)y0 * Wc + x1) * Cout + c];
            const float v10 = (float)o[((size_t)y1 * Wc + x0) * Cout + c], v11 = (float)o[((size_t)y1 * Wc + x1) * Cout + c];
            const float top = v00 + wx * (v01 - v00), bot = v10 + wx * (v11 - v10);
            warm[b * 2 * MN + c * MN + id] = base[b * 2 * MN + c * MN + id] + s * (top + wy * (bot - top));
        }
    }
}

// The same combination with 32-bit cell indexing: block (cells, crossbar),
// one cell per thread (the 64-bit divisions of the flat form are the cost of
// that kernel, not its bytes).
__global__ void __launch_bounds__(256) combine_up2d_kernel(const float *__restrict__ base,
                                                          const float *__restrict__ scale,
                                                          const __bf16 *__restrict__ out, int Cout, int M, int N,
                                                          int level, float *__restrict__ warm)
{
    const int id = blockIdx.x * 256 + threadIdx.x, MN = M * N;
    if (id >= MN)
        return;
    const int f = 1 << level, Hc = M >> level, Wc = N >> level;
    const float inv = 1.0f / (float)f;
    const int i = id / N, j = id - i * N;
    const size_t b = blockIdx.y;
    const float y = fminf(fmaxf(((float)i + 0.5f) * inv - 0.5f, 0.0f), (float)(Hc - 1));
    const float x = fminf(fmaxf(((float)j + 0.5f) * inv - 0.5f, 0.0f), (float)(Wc - 1));
    const int y0 = (int)y, x0 = (int)x;
    const int y1 = min(y0 + 1, Hc - 1), x1 = min(x0 + 1, Wc - 1);
    const float wy = y - (float)y0, wx = x - (float)x0;
    const __bf16 *o = out + b * (size_t)Hc * Wc * Cout;
    const float s = scale[b];
    const float *bb = base + b * 2 * (size_t)MN;
    float *ww = warm + b * 2 * (size_t)MN;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const float v00 = (float)o[(y0 * Wc + x0) * Cout + c], v01 = (float)o[(y0 * Wc + x1) * Cout + c];
        const float v10 = (float)o[(y1 * Wc + x0) * Cout + c], v11 = (float)o[(y1 * Wc + x1) * Cout + c];
        const float top = v00 + wx * (v01 - v00), bot = v10 + wx * (v11 - v10);
        ww[c * MN + id] = bb[c * MN + id] + s * (top + wy * (bot - top));
    }
}

// ---------------------------------------------------------------- conv forward
struct ConvArgs {
    int B, H, W, Cin, Cout, Kp;
    const __bf16 *x;   // [B][H][W][Cin]
    const __bf16 *w;   // [Cout][Kp]
    __bf16 *y;         // [B][H][W][Cout]
    int relu;
};

// Each wave: 16 consecutive pixels of one row x all Cout (CT tiles of 16).
template <int CT>
__global__ void __launch_bounds__(256) conv3x3_fwd_kernel(ConvArgs A)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long npix = (long)A.B * A.H * A.W;
    const long p0 = ((long)blockIdx.x * 4 + wave) * 16;
    if (p0 >= npix)
        return;
    const int x0 = (int)(p0 % A.W);
    const long by = p0 / A.W;
    const int y = (int)(by % A.H);
    const long b = by / A.H;
    const int j = lane & 15, g = lane >> 4;
    const int K9 = 9 * A.Cin;
    f4 acc[CT];
#pragma unroll
    for (int t = 0; t < CT; ++t)
        acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < A.Kp; k0 += 32) {
        const int k = k0 + 8 * g;
        bf16x8 bv = zero8();
        if (k < K9) {
            const int tap = k / A.Cin, ci = k - tap * A.Cin;
            const int ky = tap / 3, kx = tap - 3 * ky;
            const int yy = y + ky - 1, xx = x0 + j + kx - 1;
            if (yy >= 0 && yy < A.H && xx >= 0 && xx < A.W)
                bv = *reinterpret_cast<const bf16x8 *>(A.x + (((b * A.H + yy) * A.W + xx) * A.Cin + ci));
        } else if (k == K9) {
            bv[0] = (__bf16)1.0f; // bias column
        }
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(A.w + (long)(16 * t + j) * A.Kp + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
    }
    __bf16 *yp = A.y + (p0 + j) * A.Cout + 4 * g;
#pragma unroll
    for (int t = 0; t < CT; ++t) {
        typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
        bf16x4 o;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = acc[t][r];
            if (A.relu)
                v = fmaxf(v, 0.f);
            o[r] = (__bf16)v;
        }
        *reinterpret_cast<bf16x4 *>(yp + 16 * t) = o;
    }
}

// LDS-tiled variant: a workgroup of 4 waves computes a 4-row x TW-pixel output
// tile (TW = 16 P, P <= 4 segments of 16 pixels per wave) for all Cout.  The
// input tile with its one-pixel halo, (4+2) x (TW+2) pixels x Cin, is staged
// in LDS once (zeros outside the image; pixel stride Cin + 8 bf16 to spread
// the banks); every wave then builds its B fragments (im2col) from LDS and
// reuses each weight (A) fragment for its P pixel segments.
constexpr int kConvRows = 4;

template <int CT, int P>
__global__ void __launch_bounds__(256) conv3x3_fwd_lds_kernel(ConvArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *tile = reinterpret_cast<__bf16 *>(smem);
    constexpr int TW = 16 * P;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = A.Cin, Cs = Cin + 8; // LDS pixel stride (bf16)
    const int tilesX = A.W / TW, tilesY = A.H / kConvRows;
    int blk = blockIdx.x;
    const int tx = blk % tilesX;
    blk /= tilesX;
    const int ty = blk % tilesY;
    const long b = blk / tilesY;
    const int x0 = tx * TW, y0 = ty * kConvRows;
    // stage the input tile (8 channels = 16 bytes per element)
    const int C8 = Cin / 8;
    const int nvec = (kConvRows + 2) * (TW + 2) * C8;
    for (int e = threadIdx.x; e < nvec; e += 256) {
        const int c8 = e % C8;
        const int px = (e / C8) % (TW + 2);
        const int py = e / (C8 * (TW + 2));
        const int yy = y0 + py - 1, xx = x0 + px - 1;
        bf16x8 v = zero8();
        if (yy >= 0 && yy < A.H && xx >= 0 && xx < A.W)
            v = *reinterpret_cast<const bf16x8 *>(A.x + (((b * A.H + yy) * A.W + xx) * Cin + 8 * c8));
        *reinterpret_cast<bf16x8 *>(tile + (py * (TW + 2) + px) * Cs + 8 * c8) = v;
    }
    __syncthreads();
    const int j = lane & 15, g = lane >> 4;
    const int K9 = 9 * Cin;
    f4 acc[P][CT];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int t = 0; t < CT; ++t)
            acc[p][t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < A.Kp; k0 += 32) {
        const int k = k0 + 8 * g;
        bf16x8 av[CT];
#pragma unroll
        for (int t = 0; t < CT; ++t)
            av[t] = *reinterpret_cast<const bf16x8 *>(A.w + (long)(16 * t + j) * A.Kp + k);
        int tap = 0, ci = 0;
        const bool real = k < K9;
        if (real) {
            tap = k / Cin;
            ci = k - tap * Cin;
        }
        const int ky = tap / 3, kx = tap - 3 * ky;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            bf16x8 bv = zero8();
            if (real)
                bv = *reinterpret_cast<const bf16x8 *>(tile + ((wave + ky) * (TW + 2) + 16 * p + j + kx) * Cs + ci);
            else if (k == K9)
                bv[0] = (__bf16)1.0f; // bias column
#pragma unroll
            for (int t = 0; t < CT; ++t)
                acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[p][t], 0, 0, 0);
        }
    }
    typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        __bf16 *yp = A.y + (((b * A.H + y0 + wave) * A.W + x0 + 16 * p + j) * A.Cout) + 4 * g;
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            bf16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = acc[p][t][r];
                if (A.relu)
                    v = fmaxf(v, 0.f);
                o[r] = (__bf16)v;
            }
            *reinterpret_cast<bf16x4 *>(yp + 16 * t) = o;
        }
    }
}

// ---------------------------------------------------------------- fused coarse network
// The whole coarse network (coarse32f, or coarse16f at half the width) for
// one crossbar per workgroup of 16 waves, every activation in LDS
// (nn::coarse_net_body): the same implicit GEMMs as conv3x3_fwd_kernel (the
// same A and B fragments, the same k order), so the result equals the
// layer-by-layer kernels'.
struct CoarseNetArgs {
    const __bf16 *x;   // [B][32][32][8]
    const __bf16 *w[7]; // e2a, e2, e3a, e3, t3, d2, out: fragment-major [Kp/32][Cout][32]
    int kp[7];
    __bf16 *out;       // [B][32][32][16]
};

constexpr int kNetThreads = 1024;
using nn::net_kpad;

template <int C1>
__global__ void __launch_bounds__(kNetThreads) coarse_net_kernel(CoarseNetArgs A)
{
    const long b = blockIdx.x;
    const int t = threadIdx.x;
    const __bf16 *const w[7] = {A.w[0], A.w[1], A.w[2], A.w[3], A.w[4], A.w[5], A.w[6]};
    if constexpr (C1 == 16) {
        // coarse16f on the zero-bordered layout (nn::coarse16_body_pad)
        using Lds = nn::Net16PadLds;
        __shared__ __align__(16) __bf16 bufA[Lds::kA];
        __shared__ __align__(16) __bf16 bufS[Lds::kS];
        __shared__ __align__(16) __bf16 bufX[Lds::kX];
        for (int e = t; e < 32 * 32; e += kNetThreads)
            *reinterpret_cast<bf16x8 *>(bufX + nn::Padded<32, 32, 8>::at(e / 32, e % 32)) =
                *reinterpret_cast<const bf16x8 *>(A.x + (b * 1024 + e) * 8);
        nn::zero_border<kNetThreads, 32, 32, 8>(bufX);
        __syncthreads();
        nn::coarse16_body_pad<kNetThreads / 64, nn::kOutPlain>(bufX, bufA, bufS, w, A.out + b * 1024 * nn::kOutC);
    } else {
        using Lds = nn::NetLds<C1>;
        __shared__ __align__(16) __bf16 bufA[Lds::kA];
        __shared__ __align__(16) __bf16 bufS[Lds::kS];
        __shared__ __align__(16) __bf16 bufX[Lds::kX];
        // input features (16 bytes per copy)
        for (int e = t; e < 32 * 32; e += kNetThreads)
            *reinterpret_cast<bf16x8 *>(bufX + 8 * e) = *reinterpret_cast<const bf16x8 *>(A.x + (b * 1024 + e) * 8);
        __syncthreads();
        nn::coarse_net_body<C1, kNetThreads / 64>(bufX, bufA, bufS, w, A.out + b * 1024 * nn::kOutC);
    }
}

// wT[ci][tap'*Cout + co] = w[co][(8-tap')*Cin + ci]; bias and pad columns 0.
__global__ void flip_weights_kernel(const __bf16 *__restrict__ w, int Cin, int Cout, int Kp, int KpT,
                                    __bf16 *__restrict__ wT)
{
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < Cin * KpT; id += gridDim.x * blockDim.x) {
        const int ci = id / KpT, k = id % KpT;
        float v = 0.f;
        if (k < 9 * Cout) {
            const int tapp = k / Cout, co = k % Cout;
            v = (float)w[(long)co * Kp + (8 - tapp) * Cin + ci];
        }
        wT[id] = (__bf16)v;
    }
}

// ---------------------------------------------------------------- conv weight gradient
// dW[co][k] += sum_p dz[p][co] * col[p][k] over the pixels of one image row
// block, for k-tiles [kt0, kt0 + KTG) and all Cout.  Pixel chunks of 32 are
// staged transposed in LDS: dzT[co][32] and colT[k][32] (rows padded to 40).
struct WgradArgs {
    int B, H, W, Cin, Cout, Kp;
    int rows_per_block, ktg;
    const __bf16 *x;  // [B][H][W][Cin]
    const __bf16 *dz; // [B][H][W][Cout]
    float *dw;        // [Cout][Kp] accumulated
};

constexpr int kLdsRow = 40; // bf16 per staged row (32 + 8 pad, 80 bytes: 16-byte aligned)

template <int CT>
__global__ void __launch_bounds__(256) conv3x3_wgrad_kernel(WgradArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *dzT = reinterpret_cast<__bf16 *>(smem);               // [CT*16][40]
    __bf16 *colT = dzT + CT * 16 * kLdsRow;                         // [ktg*16][40]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int b = blockIdx.x;
    const int y0 = blockIdx.y * A.rows_per_block;
    const int y1 = min(A.H, y0 + A.rows_per_block);
    const int kt0 = blockIdx.z * A.ktg;
    const int KT = A.Kp / 16;
    const int ktn = min(A.ktg, KT - kt0);
    const int ntiles = CT * ktn;
    const int K9 = 9 * A.Cin;
    const int j = lane & 15, g = lane >> 4;
    constexpr int kMaxTilesPerWave = 16;
    f4 acc[kMaxTilesPerWave];
#pragma unroll
    for (int q = 0; q < kMaxTilesPerWave; ++q)
        acc[q] = f4{0.f, 0.f, 0.f, 0.f};
    const long pbeg = (long)y0 * A.W, pend = (long)y1 * A.W;
    const long img = (long)b * A.H * A.W;
    for (long pc = pbeg; pc < pend; pc += 32) {
        // stage dz^T: 32 pixels x Cout
        for (int e = t; e < 32 * CT * 16; e += 256) {
            const int co = e % (CT * 16), pp = e / (CT * 16);
            const long p = pc + pp;
            dzT[co * kLdsRow + pp] = (p < pend) ? A.dz[(img + p) * A.Cout + co] : (__bf16)0.0f;
        }
        // stage col^T: 32 pixels x (ktn*16) k values
        for (int e = t; e < 32 * ktn * 16; e += 256) {
            const int kk = e % (ktn * 16), pp = e / (ktn * 16);
            const int k = kt0 * 16 + kk;
            const long p = pc + pp;
            float v = 0.f;
            if (p < pend) {
                if (k < K9) {
                    const int tap = k / A.Cin, ci = k - tap * A.Cin;
                    const int yy = (int)(p / A.W) + tap / 3 - 1, xx = (int)(p % A.W) + tap % 3 - 1;
                    if (yy >= 0 && yy < A.H && xx >= 0 && xx < A.W)
                        v = (float)A.x[(img + (long)yy * A.W + xx) * A.Cin + ci];
                } else if (k == K9) {
                    v = 1.f;
                }
            }
            colT[kk * kLdsRow + pp] = (__bf16)v;
        }
        __syncthreads();
        for (int q = 0; q < kMaxTilesPerWave; ++q) {
            const int tile = wave + 4 * q;
            if (tile >= ntiles)
                break;
            const int ct = tile % CT, kt = tile / CT;
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(dzT + (ct * 16 + j) * kLdsRow + 8 * g);
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(colT + (kt * 16 + j) * kLdsRow + 8 * g);
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[q], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int q = 0; q < kMaxTilesPerWave; ++q) {
        const int tile = wave + 4 * q;
        if (tile >= ntiles)
            break;
        const int ct = tile % CT, kt = tile / CT;
        const int k = (kt0 + kt) * 16 + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = ct * 16 + 4 * g + r;
            atomicAdd(A.dw + (long)co * A.Kp + k, acc[q][r]);
        }
    }
}

// ---------------------------------------------------------------- elementwise
__global__ void pool2_fwd_kernel(const __bf16 *__restrict__ x, int B, int H, int W, int C, __bf16 *__restrict__ y)
{
    // 8 channels (16 bytes) per thread; C is a multiple of 8
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t n = (size_t)B * Ho * Wo * C8;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C8) * 8;
        size_t r = id / C8;
        const int xo = (int)(r % Wo);
        r /= Wo;
        const int yo = (int)(r % Ho);
        const size_t b = r / Ho;
        const size_t i00 = ((b * H + 2 * yo) * W + 2 * xo) * C + c;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(x + i00);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(x + i00 + C);
        const bf16x8 a2 = *reinterpret_cast<const bf16x8 *>(x + i00 + (size_t)W * C);
        const bf16x8 a3 = *reinterpret_cast<const bf16x8 *>(x + i00 + (size_t)W * C + C);
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            o[e] = (__bf16)(0.25f * ((float)a[e] + (float)a1[e] + (float)a2[e] + (float)a3[e]));
        *reinterpret_cast<bf16x8 *>(y + (((b * Ho + yo) * Wo + xo) * C + c)) = o;
    }
}

// dx (H x W) = dy (H/2 x W/2) / 4, optionally accumulated into dx
__global__ void pool2_bwd_kernel(const __bf16 *__restrict__ dy, int B, int H, int W, int C, int accumulate,
                                 __bf16 *__restrict__ dx)
{
    const size_t n = (size_t)B * H * W * C;
    const int Ho = H / 2, Wo = W / 2;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C);
        size_t r = id / C;
        const int xx = (int)(r % W);
        r /= W;
        const int yy = (int)(r % H);
        const size_t b = r / H;
        float v = 0.25f * (float)dy[((b * Ho + yy / 2) * Wo + xx / 2) * C + c];
        if (accumulate)
            v += (float)dx[id];
        dx[id] = (__bf16)v;
    }
}

// u (2H x 2W) = nearest-upsample(t (H x W)) + e, 8 channels per thread
__global__ void up2_add_fwd_kernel(const __bf16 *__restrict__ t, const __bf16 *__restrict__ e, int B, int H, int W,
                                   int C, __bf16 *__restrict__ u)
{
    const int H2 = 2 * H, W2 = 2 * W, C8 = C / 8;
    const size_t n = (size_t)B * H2 * W2 * C8;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C8) * 8;
        size_t r = id / C8;
        const int xx = (int)(r % W2);
        r /= W2;
        const int yy = (int)(r % H2);
        const size_t b = r / H2;
        const size_t o = ((b * H2 + yy) * W2 + xx) * C + c;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(t + ((b * H + yy / 2) * W + xx / 2) * C + c);
        const bf16x8 k = *reinterpret_cast<const bf16x8 *>(e + o);
        bf16x8 v;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = (__bf16)((float)a[q] + (float)k[q]);
        *reinterpret_cast<bf16x8 *>(u + o) = v;
    }
}

// dt (H x W) = sum over the 2x2 block of du (2H x 2W)
__global__ void up2_bwd_kernel(const __bf16 *__restrict__ du, int B, int H, int W, int C, __bf16 *__restrict__ dt)
{
    const int W2 = 2 * W;
    const size_t n = (size_t)B * H * W * C;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C);
        size_t r = id / C;
        const int xo = (int)(r % W);
        r /= W;
        const int yo = (int)(r % H);
        const size_t b = r / H;
        const size_t i00 = ((b * 2 * H + 2 * yo) * W2 + 2 * xo) * C + c;
        dt[id] = (__bf16)((float)du[i00] + (float)du[i00 + C] + (float)du[i00 + (size_t)W2 * C] + (float)du[i00 + (size_t)W2 * C + C]);
    }
}

// dz = dy * (y > 0), in place on dy; or dz = (a + b) * (y > 0) when add != null
__global__ void relu_mask_kernel(__bf16 *__restrict__ dy, const __bf16 *__restrict__ add, const __bf16 *__restrict__ y,
                                 size_t n)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        float v = (float)dy[id];
        if (add)
            v += (float)add[id];
        dy[id] = (__bf16)((float)y[id] > 0.f ? v : 0.f);
    }
}

__global__ void add_kernel(const __bf16 *__restrict__ a, const __bf16 *__restrict__ b, __bf16 *__restrict__ c, size_t n)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x)
        c[id] = (__bf16)((float)a[id] + (float)b[id]);
}

// MSE on channels 0,1 of out [P][Cout] against target [B][2][MN]:
// loss += sum (out - t)^2 / (2 P), dout = (out - t) / P (other channels 0)
__global__ void mse_kernel(const __bf16 *__restrict__ out, int Cout, const float *__restrict__ target, int B, int MN,
                           __bf16 *__restrict__ dout, float *__restrict__ loss)
{
    const size_t P = (size_t)B * MN;
    float local = 0.f;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < P; k += (size_t)gridDim.x * blockDim.x) {
        const size_t b = k / MN, id = k % MN;
        for (int c = 0; c < Cout; ++c) {
            float d = 0.f;
            if (c < 2) {
                d = (float)out[k * Cout + c] - target[b * 2 * MN + (size_t)c * MN + id];
                local += d * d;
            }
            dout[k * Cout + c] = (__bf16)(d / (float)P);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        local += __shfl_xor(local, off, 64);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(loss, local / (2.f * (float)P));
}

// Adam on FP32 master weights; writes the bf16 copy.  Gradient clipped by gscale.
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ grad, __bf16 *__restrict__ wb, size_t n, float lr, float b1,
                            float b2, float eps, float bc1, float bc2, float wd, float gscale)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gr = grad[i] * gscale + wd * w[i];
        const float mi = b1 * m[i] + (1.f - b1) * gr;
        const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        w[i] = wi;
        wb[i] = (__bf16)wi;
    }
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ a, __bf16 *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (__bf16)a[i];
}

__global__ void bf16_to_f32_kernel(const __bf16 *__restrict__ a, float *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

inline int grid_for(size_t n, int block = 256)
{
    size_t g = (n + block - 1) / block;
    return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

} // namespace
} // namespace fadesim

using namespace fadesim;

#define NN_LAUNCH_CHECK() FADESIM_HIP_CHECK(hipGetLastError())

extern "C" int fadesim_nn_features_pooled(int B, int M, int N, int level, const float *d_g, const double *d_r,
                                          const double *d_vin, void *d_feat, float *d_base, float *d_scale,
                                          void *stream)
{
    if (B < 0 || !d_g || !d_r || !d_vin || !d_feat || !d_scale)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_features_pooled: bad arguments");
    if (M != FeatGeo::M || N != FeatGeo::N || level != 2)
        return set_error(FADESIM_E_UNSUPPORTED, "fadesim_nn_features_pooled: 128x128 crossbars, level 2 only");
    if (B == 0)
        return FADESIM_OK;
    hipLaunchKernelGGL(features_pooled_kernel, dim3(B), dim3(FeatGeo::NT), 0, (hipStream_t)stream, d_g, d_r, d_vin,
                       (__bf16 *)d_feat, d_base, d_scale);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_features(int B, int M, int N, const float *d_g, const double *d_r, const double *d_vin,
                                   void *d_feat, float *d_base, float *d_scale, void *stream)
{
    if (B < 0 || M < 2 || N < 2 || !d_g || !d_r || !d_vin || !d_feat || !d_base || !d_scale)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_features: bad arguments");
    if (B == 0)
        return FADESIM_OK;
    if (M > kFeatMax || N > kFeatMax)
        return set_error(FADESIM_E_UNSUPPORTED, "fadesim_nn_features: rows and cols must be <= 128");
    float *tmp = (float *)workspace((size_t)B * M * N * sizeof(float));
    if (!tmp)
        return FADESIM_E_NOMEM;
    hipLaunchKernelGGL(features_kernel, dim3(B), dim3(kFeatThreads), 0, (hipStream_t)stream, d_g, d_r, d_vin, M, N, tmp,
                       (__bf16 *)d_feat, d_base, d_scale);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_combine(int B, int M, int N, const float *d_base, const float *d_scale, const void *d_out,
                                  int cout, float *d_warm, void *stream)
{
    if (B < 0 || cout < 2 || !d_base || !d_scale || !d_out || !d_warm)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_combine: bad arguments");
    const size_t n = (size_t)B * M * N;
    hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_base, d_scale,
                       (const __bf16 *)d_out, cout, B, M * N, d_warm);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_combine_up(int B, int M, int N, const float *d_base, const float *d_scale,
                                     const void *d_out, int cout, int level, float *d_warm, void *stream)
{
    if (B < 0 || cout < 2 || level < 0 || level > 6 || !d_base || !d_scale || !d_out || !d_warm ||
        (M >> level) < 1 || (N >> level) < 1 || ((M >> level) << level) != M || ((N >> level) << level) != N)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_combine_up: bad arguments");
    if (B == 0)
        return FADESIM_OK;
    if ((size_t)M * N * 2 < (1u << 31) && B <= 65535) {
        hipLaunchKernelGGL(combine_up2d_kernel, dim3((M * N + 255) / 256, B), dim3(256), 0, (hipStream_t)stream,
                           d_base, d_scale, (const __bf16 *)d_out, cout, M, N, level, d_warm);
        NN_LAUNCH_CHECK();
        return FADESIM_OK;
    }
    const size_t n = (size_t)B * M * N;
    hipLaunchKernelGGL(combine_up_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_base, d_scale,
                       (const __bf16 *)d_out, cout, B, M, N, level, d_warm);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

template <int CT>
static int launch_conv_lds(const ConvArgs &a, hipStream_t s)
{
    const int P = a.W >= 64 ? 4 : a.W / 16;
    const int TW = 16 * P;
    const long blocks = (long)a.B * (a.H / kConvRows) * (a.W / TW);
    const size_t lds = (size_t)(kConvRows + 2) * (TW + 2) * (a.Cin + 8) * sizeof(__bf16);
    const dim3 grid((unsigned)blocks);
    switch (P) {
    case 1: hipLaunchKernelGGL((conv3x3_fwd_lds_kernel<CT, 1>), grid, dim3(256), lds, s, a); break;
    case 2: hipLaunchKernelGGL((conv3x3_fwd_lds_kernel<CT, 2>), grid, dim3(256), lds, s, a); break;
    default: hipLaunchKernelGGL((conv3x3_fwd_lds_kernel<CT, 4>), grid, dim3(256), lds, s, a); break;
    }
    return FADESIM_OK;
}

extern "C" int fadesim_nn_coarse_forward(int B, const void *d_x, const void *const *d_w, const int *kp, void *d_out,
                                         void *stream)
{
    if (B < 0 || !d_x || !d_w || !kp || !d_out)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_coarse_forward: bad arguments");
    // the first width, from the depth of the second layer (C1 -> C1)
    const int c1 = kp[1] == net_kpad(32) ? 32 : kp[1] == net_kpad(16) ? 16 : 0;
    if (!c1)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_coarse_forward: layer 2 depth fits neither width 32 nor 16");
    const int cin[7] = {8, c1, c1, 2 * c1, 2 * c1, c1, c1};
    CoarseNetArgs a{};
    a.x = (const __bf16 *)d_x;
    a.out = (__bf16 *)d_out;
    for (int l = 0; l < 7; ++l) {
        if (!d_w[l] || kp[l] != net_kpad(cin[l]))
            return set_error(FADESIM_E_INVALID, "fadesim_nn_coarse_forward: layer weights / Kp do not fit");
        a.w[l] = (const __bf16 *)d_w[l];
        a.kp[l] = kp[l];
    }
    if (B == 0)
        return FADESIM_OK;
    if (c1 == 32)
        hipLaunchKernelGGL(coarse_net_kernel<32>, dim3(B), dim3(kNetThreads), 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(coarse_net_kernel<16>, dim3(B), dim3(kNetThreads), 0, (hipStream_t)stream, a);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_conv3x3_fwd(int B, int H, int W, int Cin, int Cout, int Kp, const void *d_x, const void *d_w,
                                      void *d_y, int relu, void *stream)
{
    if (B < 1 || H < 1 || W % 16 != 0 || Cin % 8 != 0 || Cout % 16 != 0 || Kp % 32 != 0 || Kp < 9 * Cin + 1)
        return set_error(FADESIM_E_INVALID, "conv3x3_fwd: need W%16==0, Cin%8==0, Cout%16==0, Kp%32==0, Kp>9*Cin");
    ConvArgs a{B, H, W, Cin, Cout, Kp, (const __bf16 *)d_x, (const __bf16 *)d_w, (__bf16 *)d_y, relu};
    hipStream_t s = (hipStream_t)stream;
    const int TWt = W >= 64 ? 64 : W;
    const size_t lds_bytes = (size_t)(kConvRows + 2) * (TWt + 2) * (Cin + 8) * sizeof(__bf16);
    const bool tiled = H % kConvRows == 0 && (W == 16 || W == 32 || W % 64 == 0) && lds_bytes <= 65536;
    if (tiled) {
        switch (Cout / 16) {
        case 1: launch_conv_lds<1>(a, s); break;
        case 2: launch_conv_lds<2>(a, s); break;
        case 4: launch_conv_lds<4>(a, s); break;
        case 6: launch_conv_lds<6>(a, s); break;
        case 8: launch_conv_lds<8>(a, s); break;
        default: return set_error(FADESIM_E_UNSUPPORTED, "conv3x3_fwd: Cout must be 16, 32, 64, 96 or 128");
        }
        NN_LAUNCH_CHECK();
        return FADESIM_OK;
    }
    const long npix = (long)B * H * W;
    const dim3 grid((unsigned)((npix + 63) / 64));
    switch (Cout / 16) {
    case 1: hipLaunchKernelGGL(conv3x3_fwd_kernel<1>, grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(conv3x3_fwd_kernel<2>, grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(conv3x3_fwd_kernel<4>, grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(conv3x3_fwd_kernel<6>, grid, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(conv3x3_fwd_kernel<8>, grid, dim3(256), 0, s, a); break;
    default: return set_error(FADESIM_E_UNSUPPORTED, "conv3x3_fwd: Cout must be 16, 32, 64, 96 or 128");
    }
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_flip_weights(int Cin, int Cout, int Kp, int KpT, const void *d_w, void *d_wT, void *stream)
{
    if (KpT < 9 * Cout + 1 || Kp < 9 * Cin + 1)
        return set_error(FADESIM_E_INVALID, "flip_weights: bad padded widths");
    hipLaunchKernelGGL(flip_weights_kernel, dim3(grid_for((size_t)Cin * KpT)), dim3(256), 0, (hipStream_t)stream,
                       (const __bf16 *)d_w, Cin, Cout, Kp, KpT, (__bf16 *)d_wT);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_conv3x3_wgrad(int B, int H, int W, int Cin, int Cout, int Kp, const void *d_x,
                                        const void *d_dz, float *d_dw, void *stream)
{
    if (B < 1 || H < 1 || W < 1 || Cin % 8 != 0 || Cout % 16 != 0 || Kp % 32 != 0 || Kp < 9 * Cin + 1)
        return set_error(FADESIM_E_INVALID, "conv3x3_wgrad: bad shape");
    const int CT = Cout / 16;
    const int KT = Kp / 16;
    int ktg = 64 / CT;
    if (ktg > KT)
        ktg = KT;
    // rows per block: about 2048 pixels
    int rpb = 2048 / W;
    if (rpb < 1)
        rpb = 1;
    if (rpb > H)
        rpb = H;
    WgradArgs a{B, H, W, Cin, Cout, Kp, rpb, ktg, (const __bf16 *)d_x, (const __bf16 *)d_dz, d_dw};
    const dim3 grid(B, (H + rpb - 1) / rpb, (KT + ktg - 1) / ktg);
    const size_t lds = (size_t)(CT * 16 + ktg * 16) * kLdsRow * sizeof(__bf16);
    hipStream_t s = (hipStream_t)stream;
    switch (CT) {
    case 1: hipLaunchKernelGGL(conv3x3_wgrad_kernel<1>, grid, dim3(256), lds, s, a); break;
    case 2: hipLaunchKernelGGL(conv3x3_wgrad_kernel<2>, grid, dim3(256), lds, s, a); break;
    case 4: hipLaunchKernelGGL(conv3x3_wgrad_kernel<4>, grid, dim3(256), lds, s, a); break;
    case 6: hipLaunchKernelGGL(conv3x3_wgrad_kernel<6>, grid, dim3(256), lds, s, a); break;
    case 8: hipLaunchKernelGGL(conv3x3_wgrad_kernel<8>, grid, dim3(256), lds, s, a); break;
    default: return set_error(FADESIM_E_UNSUPPORTED, "conv3x3_wgrad: Cout must be 16, 32, 64, 96 or 128");
    }
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_pool2(int B, int H, int W, int C, const void *d_x, void *d_y, void *stream)
{
    if (C % 8 != 0 || H % 2 != 0 || W % 2 != 0)
        return set_error(FADESIM_E_INVALID, "pool2: need C % 8 == 0 and even H, W");
    const size_t n = (size_t)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(pool2_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_x, B,
                       H, W, C, (__bf16 *)d_y);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_pool2_bwd(int B, int H, int W, int C, const void *d_dy, int accumulate, void *d_dx,
                                    void *stream)
{
    const size_t n = (size_t)B * H * W * C;
    hipLaunchKernelGGL(pool2_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_dy,
                       B, H, W, C, accumulate, (__bf16 *)d_dx);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_up2_add(int B, int H, int W, int C, const void *d_t, const void *d_e, void *d_u, void *stream)
{
    if (C % 8 != 0)
        return set_error(FADESIM_E_INVALID, "up2_add: need C % 8 == 0");
    const size_t n = (size_t)B * 4 * H * W * (C / 8);
    hipLaunchKernelGGL(up2_add_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_t,
                       (const __bf16 *)d_e, B, H, W, C, (__bf16 *)d_u);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_up2_bwd(int B, int H, int W, int C, const void *d_du, void *d_dt, void *stream)
{
    const size_t n = (size_t)B * H * W * C;
    hipLaunchKernelGGL(up2_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_du, B,
                       H, W, C, (__bf16 *)d_dt);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_relu_mask(size_t n, void *d_dy, const void *d_add, const void *d_y, void *stream)
{
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (__bf16 *)d_dy,
                       (const __bf16 *)d_add, (const __bf16 *)d_y, n);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_add(size_t n, const void *d_a, const void *d_b, void *d_c, void *stream)
{
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_a,
                       (const __bf16 *)d_b, (__bf16 *)d_c, n);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_mse(int B, int M, int N, const void *d_out, int cout, const float *d_target, void *d_dout,
                              float *d_loss, void *stream)
{
    const size_t n = (size_t)B * M * N;
    hipLaunchKernelGGL(mse_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)d_out, cout,
                       d_target, B, M * N, (__bf16 *)d_dout, d_loss);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_adam(size_t n, float *d_w, float *d_m, float *d_v, const float *d_grad, void *d_wb, float lr,
                               float beta1, float beta2, float eps, int step, float weight_decay, float grad_scale,
                               void *stream)
{
    const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_w, d_m, d_v, d_grad,
                       (__bf16 *)d_wb, n, lr, beta1, beta2, eps, bc1, bc2, weight_decay, grad_scale);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}

extern "C" int fadesim_nn_cast(size_t n, const void *d_src, int src_bf16, void *d_dst, void *stream)
{
    if (src_bf16)
        hipLaunchKernelGGL(bf16_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (const __bf16 *)d_src, (float *)d_dst, n);
    else
        hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                           (const float *)d_src, (__bf16 *)d_dst, n);
    NN_LAUNCH_CHECK();
    return FADESIM_OK;
}
