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

namespace fadesim {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e)
        z[e] = (__bf16)0.0f;
    return z;
}

// ---------------------------------------------------------------- features
// One workgroup per crossbar, one thread per row (pass over rows) then per
// column.  First-order drops from the ideal cell currents I = g c:
//   p1_ij = r sum_k (min(j,k)+1) I_ik      (wordline, source at the left)
//   q1_ij = r sum_k (n - max(i,k)) I_kj    (bitline, sink at the bottom)
// second order from I = g (c - p1 - q1).  Features (bf16, 8 channels):
//   p1/s, q1/s, (p2-p1)/s2, (q2-q1)/s2, g/gmax, vin/vmax, r/5, 1
// with s = max(p1, q1) and s2 = max|p2-p1|, |q2-q1| of the instance; base =
// (p2, q2) in FP32 and the output scale s2 per instance.
__global__ void __launch_bounds__(128) features_kernel(const float *__restrict__ g, const double *__restrict__ rw,
                                                       const double *__restrict__ vin, int M, int N,
                                                       float *__restrict__ tmp, __bf16 *__restrict__ feat,
                                                       float *__restrict__ base, float *__restrict__ scale)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const size_t MN = (size_t)M * N;
    const float *gb = g + b * MN;
    const double *vb = vin + (size_t)b * M;
    const double r = rw[b];
    float *p1 = tmp + b * 4 * MN, *q1 = p1 + MN, *p2 = p1 + 2 * MN, *q2 = p1 + 3 * MN;
    __shared__ float red[4][128];
    // order 1 and 2 (order 2 uses the order-1 fields)
    for (int order = 0; order < 2; ++order) {
        float *P = order ? p2 : p1, *Q = order ? q2 : q1;
        for (int i = t; i < M; i += blockDim.x) {
            const double c = vb[i];
            double R = 0.0;
            for (int k = 0; k < N; ++k) {
                const size_t id = (size_t)i * N + k;
                const double cur = (double)gb[id] * (order ? c - (double)p1[id] - (double)q1[id] : c);
                R += cur;
            }
            double P0 = 0.0, P1 = 0.0;
            for (int j = 0; j < N; ++j) {
                const size_t id = (size_t)i * N + j;
                const double cur = (double)gb[id] * (order ? c - (double)p1[id] - (double)q1[id] : c);
                P[id] = (float)(r * (P1 + (double)(j + 1) * (R - P0)));
                P0 += cur;
                P1 += (double)(j + 1) * cur;
            }
        }
        for (int j = t; j < N; j += blockDim.x) {
            double T = 0.0;
            for (int k = 0; k < M; ++k) {
                const size_t id = (size_t)k * N + j;
                const double cur = (double)gb[id] * (order ? vb[k] - (double)p1[id] - (double)q1[id] : vb[k]);
                T += cur * (double)(M - k);
            }
            double A = 0.0, Bs = T;
            for (int i = 0; i < M; ++i) {
                const size_t id = (size_t)i * N + j;
                const double cur = (double)gb[id] * (order ? vb[i] - (double)p1[id] - (double)q1[id] : vb[i]);
                A += cur;
                Bs -= cur * (double)(M - i);
                Q[id] = (float)(r * ((double)(M - i) * A + Bs));
            }
        }
        __syncthreads();
    }
    // scales
    float s1 = 0.f, s2 = 0.f, gm = 0.f, vm = 0.f;
    for (size_t id = t; id < MN; id += blockDim.x) {
        s1 = fmaxf(s1, fmaxf(fabsf(p1[id]), fabsf(q1[id])));
        s2 = fmaxf(s2, fmaxf(fabsf(p2[id] - p1[id]), fabsf(q2[id] - q1[id])));
        gm = fmaxf(gm, gb[id]);
    }
    for (int i = t; i < M; i += blockDim.x)
        vm = fmaxf(vm, (float)fabs(vb[i]));
    red[0][t] = s1;
    red[1][t] = s2;
    red[2][t] = gm;
    red[3][t] = vm;
    __syncthreads();
    for (int off = 64; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < 4; ++q)
                red[q][t] = fmaxf(red[q][t], red[q][t + off]);
        __syncthreads();
    }
    const float S1 = fmaxf(red[0][0], 1e-30f), S2 = fmaxf(red[1][0], 1e-30f);
    const float GM = fmaxf(red[2][0], 1e-30f), VM = fmaxf(red[3][0], 1e-30f);
    if (t == 0)
        scale[b] = S2;
    for (size_t id = t; id < MN; id += blockDim.x) {
        const int i = (int)(id / N);
        bf16x8 f;
        f[0] = (__bf16)(p1[id] / S1);
        f[1] = (__bf16)(q1[id] / S1);
        f[2] = (__bf16)((p2[id] - p1[id]) / S2);
        f[3] = (__bf16)((q2[id] - q1[id]) / S2);
        f[4] = (__bf16)(gb[id] / GM);
        f[5] = (__bf16)((float)vb[i] / VM);
        f[6] = (__bf16)((float)(r / 5.0));
        f[7] = (__bf16)1.0f;
        *reinterpret_cast<bf16x8 *>(feat + (b * MN + id) * 8) = f;
        base[b * 2 * MN + id] = p2[id];
        base[b * 2 * MN + MN + id] = q2[id];
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
    const int Ho = H / 2, Wo = W / 2;
    const size_t n = (size_t)B * Ho * Wo * C;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C);
        size_t r = id / C;
        const int xo = (int)(r % Wo);
        r /= Wo;
        const int yo = (int)(r % Ho);
        const size_t b = r / Ho;
        const size_t i00 = ((b * H + 2 * yo) * W + 2 * xo) * C + c;
        const float s = (float)x[i00] + (float)x[i00 + C] + (float)x[i00 + (size_t)W * C] + (float)x[i00 + (size_t)W * C + C];
        y[id] = (__bf16)(0.25f * s);
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

// u (2H x 2W) = nearest-upsample(t (H x W)) + e
__global__ void up2_add_fwd_kernel(const __bf16 *__restrict__ t, const __bf16 *__restrict__ e, int B, int H, int W,
                                   int C, __bf16 *__restrict__ u)
{
    const int H2 = 2 * H, W2 = 2 * W;
    const size_t n = (size_t)B * H2 * W2 * C;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(id % C);
        size_t r = id / C;
        const int xx = (int)(r % W2);
        r /= W2;
        const int yy = (int)(r % H2);
        const size_t b = r / H2;
        u[id] = (__bf16)((float)t[((b * H + yy / 2) * W + xx / 2) * C + c] + (float)e[id]);
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

extern "C" int fadesim_nn_features(int B, int M, int N, const float *d_g, const double *d_r, const double *d_vin,
                                   void *d_feat, float *d_base, float *d_scale, void *stream)
{
    if (B < 0 || M < 2 || N < 2 || !d_g || !d_r || !d_vin || !d_feat || !d_base || !d_scale)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_features: bad arguments");
    if (B == 0)
        return FADESIM_OK;
    float *tmp = (float *)workspace((size_t)B * 4 * M * N * sizeof(float));
    if (!tmp)
        return FADESIM_E_NOMEM;
    hipLaunchKernelGGL(features_kernel, dim3(B), dim3(128), 0, (hipStream_t)stream, d_g, d_r, d_vin, M, N, tmp,
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

extern "C" int fadesim_nn_conv3x3_fwd(int B, int H, int W, int Cin, int Cout, int Kp, const void *d_x, const void *d_w,
                                      void *d_y, int relu, void *stream)
{
    if (B < 1 || H < 1 || W % 16 != 0 || Cin % 8 != 0 || Cout % 16 != 0 || Kp % 32 != 0 || Kp < 9 * Cin + 1)
        return set_error(FADESIM_E_INVALID, "conv3x3_fwd: need W%16==0, Cin%8==0, Cout%16==0, Kp%32==0, Kp>9*Cin");
    ConvArgs a{B, H, W, Cin, Cout, Kp, (const __bf16 *)d_x, (const __bf16 *)d_w, (__bf16 *)d_y, relu};
    const long npix = (long)B * H * W;
    const dim3 grid((unsigned)((npix + 63) / 64));
    hipStream_t s = (hipStream_t)stream;
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
    const size_t n = (size_t)B * (H / 2) * (W / 2) * C;
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
    const size_t n = (size_t)B * 4 * H * W * C;
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
