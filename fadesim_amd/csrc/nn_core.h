// Device code of the learned warm start shared by its own kernels (nn.hip)
// and by the batched solver's guess kernel that runs it in-kernel
// (rbsor.hip): the pooled physics features of a 128x128 crossbar and the
// fused coarse network (coarse16f / coarse32f) on one workgroup's waves.
//
// A 3x3 "same" convolution with weights w[Cout][Kp] (k = tap*Cin + ci,
// tap = 3ky+kx, bias in column 9*Cin, zero columns up to Kp = roundup(9*Cin+1,
// 32)) is the GEMM  y[p][co] = sum_k w[co][k] * col[p][k]  with col[p][k] the
// input pixel p + (ky-1, kx-1), channel ci (0 outside the image) and
// col[p][9*Cin] = 1, on v_mfma_f32_16x16x32_bf16: A = 16 output channels x 32
// k, B = 32 k x 16 pixels, D = 16 channels x 16 pixels, FP32 accumulate.
// Lane layout: A[i][k] in lane i + 16*(k/8) element k%8, B[k][j] in lane
// j + 16*(k/8) element k%8, D[i][j] in lane j + 16*(i/4) register i%4.
// Every output tile is accumulated by one wave in a fixed k order, so the
// result does not depend on how many waves share a layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "rb_core.h"

namespace fadesim {
namespace nn {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e)
        z[e] = (__bf16)0.0f;
    return z;
}

// The thread index, opaque to the optimizer: addresses derived from it are
// computed where they are used, not hoisted out of a caller's loop over
// crossbars and held in registers (or spilled) for the kernel's lifetime.
__device__ __forceinline__ int opaque_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

constexpr int kFeatC = 8;   // feature channels
constexpr int kOutC = 16;   // network output channels (0: wordline drops, 1: bitline drops)

// ---------------------------------------------------------------- pooled features
// Features of one 128x128 crossbar averaged over 4x4 cells (the network's
// 32x32 input), with the drop fields computed the solver's way: each thread
// holds a 4x8 register tile; (p1, q1) and (p2, q2) are two undamped
// wire-solve iterations from zero (rb::wire_iteration: line scans by lane
// shuffles, column sums across thread rows in FP64), for grounded bitlines.
// A thread's tile is two 4x4 pooling cells, so the pooled features come out
// of registers.  Channels (bf16):
//   p1/S1, q1/S1, (p2-p1)/S2, (q2-q1)/S2, g/gmax, vin/vmax, r/5, 1
// averaged over the cell, with S1 = max(|p1|, |q1|), S2 = max(|p2-p1|,
// |q2-q1|) over the crossbar.  Leaves (p2, q2) -- the base field -- in p, q,
// writes the features to feat ([32][32][8], global memory or LDS) and
// returns S2 (the network output's scale), the same in every thread.
// red: 4 x (NT/64) floats of LDS.  Whole workgroup; ends with a barrier.
using FeatGeo = rb::Geo<4, 8, 32, 16>;

// PAD: feat is a zero-bordered [34][34][8] tensor (the padded layout of the
// coarse16f layers below) and the features go to its interior.
template <bool PAD = false>
__device__ __forceinline__ float pooled_features(const rb::Shared &sh, const float *gp, double r, const double *vin,
                                                 __bf16 *feat, float *red, float (&p)[FeatGeo::TR][FeatGeo::TC],
                                                 float (&q)[FeatGeo::TR][FeatGeo::TC])
{
    using G = FeatGeo;
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, M = G::M, N = G::N, NT = G::NT;
    constexpr int PW = N / 4; // pooled width
    const rb::StripCtx X{};
    const int t = opaque_tid(), tx = t % TX, ty = t / TX;
    const int i0 = TR * ty, j0 = TC * tx;
    float vi[TR];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii)
        vi[ii] = (float)rb::ldg(vin + i0 + ii);
    auto cur = [&](int ii, const float (&pr)[TC], const float (&qr)[TC], float (&a)[TC], float (&c)[TC]) {
        float gr[TC];
        rb::ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, gr);
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            a[jj] = c[jj] = gr[jj] * ((vi[ii] - pr[jj]) - qr[jj]);
    };
    float p1[TR][TC], q1[TR][TC];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii)
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            p[ii][jj] = q[ii][jj] = 0.f;
    int nbar = 0;
    rb::wire_iteration<G, false>(sh, X, cur, r, 1.0, M, 0, i0, j0, tx, ty, t, nbar, p, q);
    float s1 = 0.f, gm = 0.f, vm = 0.f;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        vm = fmaxf(vm, fabsf(vi[ii]));
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            p1[ii][jj] = p[ii][jj];
            q1[ii][jj] = q[ii][jj];
            s1 = fmaxf(s1, fmaxf(fabsf(p[ii][jj]), fabsf(q[ii][jj])));
        }
    }
    rb::wire_iteration<G, false>(sh, X, cur, r, 1.0, M, 0, i0, j0, tx, ty, t, nbar, p, q);
    float s2 = 0.f;
    float pool[2][5]; // per pooling cell: sums of p1, q1, p2 - p1, q2 - q1, g
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 5; ++e)
            pool[h][e] = 0.f;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        float gr[TC];
        rb::ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, gr);
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const float dp = p[ii][jj] - p1[ii][jj], dq = q[ii][jj] - q1[ii][jj];
            s2 = fmaxf(s2, fmaxf(fabsf(dp), fabsf(dq)));
            gm = fmaxf(gm, gr[jj]);
            float *pc = pool[jj / 4];
            pc[0] += p1[ii][jj];
            pc[1] += q1[ii][jj];
            pc[2] += dp;
            pc[3] += dq;
            pc[4] += gr[jj];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 = fmaxf(s1, __shfl_xor(s1, off, 64));
        s2 = fmaxf(s2, __shfl_xor(s2, off, 64));
        gm = fmaxf(gm, __shfl_xor(gm, off, 64));
        vm = fmaxf(vm, __shfl_xor(vm, off, 64));
    }
    constexpr int NW = NT / 64;
    if ((t & 63) == 0) {
        red[0 * NW + (t >> 6)] = s1;
        red[1 * NW + (t >> 6)] = s2;
        red[2 * NW + (t >> 6)] = gm;
        red[3 * NW + (t >> 6)] = vm;
    }
    __syncthreads();
    float S1 = 1e-30f, S2 = 1e-30f, GM = 1e-30f, VM = 1e-30f;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        S1 = fmaxf(S1, red[0 * NW + w]);
        S2 = fmaxf(S2, red[1 * NW + w]);
        GM = fmaxf(GM, red[2 * NW + w]);
        VM = fmaxf(VM, red[3 * NW + w]);
    }
    const float vavg = 0.25f * (((vi[0] + vi[1]) + vi[2]) + vi[3]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float *pc = pool[h];
        bf16x8 f;
        f[0] = (__bf16)(pc[0] * (1.f / 16.f) / S1);
        f[1] = (__bf16)(pc[1] * (1.f / 16.f) / S1);
        f[2] = (__bf16)(pc[2] * (1.f / 16.f) / S2);
        f[3] = (__bf16)(pc[3] * (1.f / 16.f) / S2);
        f[4] = (__bf16)(pc[4] * (1.f / 16.f) / GM);
        f[5] = (__bf16)(vavg / VM);
        f[6] = (__bf16)((float)(r / 5.0));
        f[7] = (__bf16)1.0f;
        const size_t px = PAD ? (size_t)(ty + 1) * (PW + 2) + 2 * tx + h + 1 : (size_t)ty * PW + 2 * tx + h;
        *reinterpret_cast<bf16x8 *>(feat + px * kFeatC) = f;
    }
    __syncthreads(); // red and the features are complete
    return S2;
}

// ---------------------------------------------------------------- fused coarse network
constexpr int net_kpad(int cin) { return ((9 * cin + 1 + 31) / 32) * 32; }

// One layer on the NW waves of the workgroup: wave (cg, pg) computes CTW
// output-channel tiles (16 channels each) of P pixel tiles (16 pixels each),
// so a weight fragment serves P tiles and an input fragment CTW tiles; the
// depth is a compile-time constant (the k loop unrolls by 2 so that the next
// fragments load while the present ones multiply).
template <int NW, int CIN, int COUT, int H, int W, bool RELU, int CTW, int P>
__device__ __forceinline__ void net_conv(const __bf16 *src, const __bf16 *__restrict__ w, __bf16 *dst)
{
    constexpr int CT = COUT / 16, PT = H * W / 16, K9 = 9 * CIN, KP = net_kpad(CIN);
    constexpr int NPG = PT / P; // pixel groups
    static_assert(W % 16 == 0 && CIN % 8 == 0 && COUT % 16 == 0, "tile shapes");
    static_assert(CT % CTW == 0 && PT % P == 0 && (CT / CTW) * NPG == NW, "waves x tiles");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int pg = wave % NPG, cg = wave / NPG;
    f4 acc[P][CTW];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < KP; k0 += 32) {
        const int k = k0 + 8 * g;
        bf16x8 av[CTW];
#pragma unroll
        for (int t = 0; t < CTW; ++t) // fragment-major weights: one contiguous 1 KB per wave
            av[t] = *reinterpret_cast<const bf16x8 *>(w + ((k0 / 32) * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
        int tap = 0, ci = 0;
        const bool real = k < K9;
        if (real) {
            tap = k / CIN;
            ci = k - tap * CIN;
        }
        const int ky = tap / 3, kx = tap - 3 * ky;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int p0 = 16 * (pg * P + q), y = p0 / W, x0 = p0 % W;
            bf16x8 bv = zero8();
            if (real) {
                const int yy = y + ky - 1, xx = x0 + j + kx - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    bv = *reinterpret_cast<const bf16x8 *>(src + ((yy * W + xx) * CIN + ci));
            } else if (k == K9) {
                bv[0] = (__bf16)1.0f; // bias column
            }
#pragma unroll
            for (int t = 0; t < CTW; ++t)
                acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[q][t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int p0 = 16 * (pg * P + q);
#pragma unroll
        for (int t = 0; t < CTW; ++t) {
            bf16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = acc[q][t][r];
                if (RELU)
                    v = fmaxf(v, 0.f);
                o[r] = (__bf16)v;
            }
            *reinterpret_cast<bf16x4 *>(dst + (p0 + j) * COUT + 16 * (cg * CTW + t) + 4 * g) = o;
        }
    }
}

// Lane shifts within rows of 16 lanes (DPP row_shr:1 / row_shl:1): lane j
// receives lane j-1's (j+1's) fragment, 0 where that lane is outside the row.
__device__ __forceinline__ bf16x8 frag_from_left(bf16x8 v)
{
    typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i4;
    i4 x = __builtin_bit_cast(i4, v), y;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        y[e] = __builtin_amdgcn_update_dpp(0, x[e], 0x111, 0xF, 0xF, true);
    return __builtin_bit_cast(bf16x8, y);
}
__device__ __forceinline__ bf16x8 frag_from_right(bf16x8 v)
{
    typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i4;
    i4 x = __builtin_bit_cast(i4, v), y;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        y[e] = __builtin_amdgcn_update_dpp(0, x[e], 0x101, 0xF, 0xF, true);
    return __builtin_bit_cast(bf16x8, y);
}

// net_conv for CIN a multiple of 32: the k blocks of one kernel row (taps
// kx = 0, 1, 2 of row ky, each CIN/32 blocks) are built from ONE load of the
// centre tap: the left and right taps are the centre fragments of the
// neighbouring lanes (DPP row shifts), only the two edge pixels of the 16-pixel
// segment are loaded again.  Same fragments, same k order as net_conv.
template <int NW, int CIN, int COUT, int H, int W, bool RELU, int CTW, int P>
__device__ __forceinline__ void net_conv_rows(const __bf16 *src, const __bf16 *__restrict__ w, __bf16 *dst)
{
    constexpr int CT = COUT / 16, PT = H * W / 16, K9 = 9 * CIN, KP = net_kpad(CIN);
    constexpr int NPG = PT / P, CB = CIN / 32; // pixel groups, k blocks per tap
    static_assert(W % 16 == 0 && CIN % 32 == 0 && COUT % 16 == 0 && KP == K9 + 32, "tile shapes");
    static_assert(CT % CTW == 0 && PT % P == 0 && (CT / CTW) * NPG == NW, "waves x tiles");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int pg = wave % NPG, cg = wave / NPG;
    f4 acc[P][CTW];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int ky = 0; ky < 3; ++ky) {
        bf16x8 av[3][CB][CTW]; // weights of the three taps of this kernel row
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int t = 0; t < CTW; ++t)
                    av[kx][cb][t] = *reinterpret_cast<const bf16x8 *>(
                        w + (((3 * ky + kx) * CB + cb) * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int p0 = 16 * (pg * P + q), y = p0 / W, x0 = p0 % W, yy = y + ky - 1;
            if (yy < 0 || yy >= H)
                continue; // the whole segment reads padding: the fragments are zero
            const __bf16 *rowp = src + (yy * W) * CIN + 8 * g;
            bf16x8 b[3][CB];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const bf16x8 c = *reinterpret_cast<const bf16x8 *>(rowp + (x0 + j) * CIN + 32 * cb);
                bf16x8 l = frag_from_left(c), r = frag_from_right(c);
                if (j == 0 && x0 > 0)
                    l = *reinterpret_cast<const bf16x8 *>(rowp + (x0 - 1) * CIN + 32 * cb);
                if (j == 15 && x0 + 16 < W)
                    r = *reinterpret_cast<const bf16x8 *>(rowp + (x0 + 16) * CIN + 32 * cb);
                b[0][cb] = l;
                b[1][cb] = c;
                b[2][cb] = r;
            }
#pragma unroll
            for (int kx = 0; kx < 3; ++kx)
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                    for (int t = 0; t < CTW; ++t)
                        acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[kx][cb][t], b[kx][cb], acc[q][t], 0,
                                                                           0, 0);
        }
    }
    // the bias block: k = K9 is the ones column
    {
        bf16x8 av[CTW];
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            av[t] = *reinterpret_cast<const bf16x8 *>(w + ((K9 / 32) * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
        bf16x8 bv = zero8();
        if (g == 0)
            bv[0] = (__bf16)1.0f;
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int t = 0; t < CTW; ++t)
                acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[q][t], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int p0 = 16 * (pg * P + q);
#pragma unroll
        for (int t = 0; t < CTW; ++t) {
            bf16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = acc[q][t][r];
                if (RELU)
                    v = fmaxf(v, 0.f);
                o[r] = (__bf16)v;
            }
            *reinterpret_cast<bf16x4 *>(dst + (p0 + j) * COUT + 16 * (cg * CTW + t) + 4 * g) = o;
        }
    }
}

// The fused network's layers for first width C1 (C1 channels at 32x32, 2 C1
// at 16x16) on NW waves (16: the network kernel's 1024 threads, 8: the guess
// kernel's 512): the implicit-GEMM form and the split of the waves over
// output-channel tiles and pixel tiles per layer.
template <int C1, int NW>
struct CoarseLayers;
template <>
struct CoarseLayers<32, 16> { // coarse32f
    static __device__ __forceinline__ void e2a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 8, 32, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void e2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 32, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void e3a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 64, 16, 16, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void e3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 64, 64, 16, 16, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void t3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 64, 32, 16, 16, true, 1, 2>(x, w, y); }
    static __device__ __forceinline__ void d2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 32, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void out(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 16, 32, 32, false, 1, 4>(x, w, y); }
};
template <>
struct CoarseLayers<16, 16> { // coarse16f
    static __device__ __forceinline__ void e2a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 8, 16, 32, 32, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void e2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 16, 16, 32, 32, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void e3a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 16, 32, 16, 16, true, 1, 2>(x, w, y); }
    static __device__ __forceinline__ void e3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 32, 16, 16, true, 1, 2>(x, w, y); }
    static __device__ __forceinline__ void t3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<16, 32, 16, 16, 16, true, 1, 1>(x, w, y); }
    static __device__ __forceinline__ void d2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 16, 16, 32, 32, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void out(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<16, 16, 16, 32, 32, false, 1, 4>(x, w, y); }
};
template <>
struct CoarseLayers<16, 8> { // coarse16f on 8 waves
    static __device__ __forceinline__ void e2a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<8, 8, 16, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void e2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<8, 16, 16, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void e3a(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<8, 16, 32, 16, 16, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void e3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<8, 32, 32, 16, 16, true, 1, 4>(x, w, y); }
    static __device__ __forceinline__ void t3(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv_rows<8, 32, 16, 16, 16, true, 1, 2>(x, w, y); }
    static __device__ __forceinline__ void d2(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<8, 16, 16, 32, 32, true, 1, 8>(x, w, y); }
    static __device__ __forceinline__ void out(const __bf16 *x, const __bf16 *w, __bf16 *y) { net_conv<8, 16, 16, 32, 32, false, 1, 8>(x, w, y); }
};

// LDS of the fused network of width C1: bufA (e2a / d2, or e3a and e3), bufS
// (e2 and u2), bufX (the input, p3 and t3), in bf16 elements.
template <int C1>
struct NetLds {
    static constexpr int kA = 32 * 32 * C1, kS = 32 * 32 * C1;
    static constexpr int kX = 32 * 32 * kFeatC > 16 * 16 * C1 ? 32 * 32 * kFeatC : 16 * 16 * C1;
};

// The whole coarse network of one crossbar on the workgroup's NW waves, the
// input features already in bufX (and a barrier since): e2a, e2 (C1
// channels, 32x32) -> 2x2 mean -> e3a, e3 (2 C1), t3 (C1) at 16x16 ->
// nearest 2x upsample + e2 -> d2 (C1) -> out (16 channels) into `out`
// ([32][32][16], global memory or LDS; no barrier after it).
template <int C1, int NW>
__device__ __forceinline__ void coarse_net_body(__bf16 *bufX, __bf16 *bufA, __bf16 *bufS,
                                                const __bf16 *const (&w)[7], __bf16 *out)
{
    using L = CoarseLayers<C1, NW>;
    constexpr int C2 = 2 * C1, G1 = C1 / 8, NT = NW * 64; // 8-channel groups of a C1-channel pixel
    const int t = threadIdx.x;
    L::e2a(bufX, w[0], bufA);
    __syncthreads();
    L::e2(bufA, w[1], bufS); // kept for the skip
    __syncthreads();
    // p3: 2x2 mean of e2 -> 16x16xC1 in bufX
    for (int e = t; e < 16 * 16 * G1; e += NT) {
        const int c = (e % G1) * 8, po = e / G1, yo = po / 16, xo = po % 16;
        const int i00 = ((2 * yo) * 32 + 2 * xo) * C1 + c;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(bufS + i00);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + C1);
        const bf16x8 a2 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + 32 * C1);
        const bf16x8 a3 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + 32 * C1 + C1);
        bf16x8 o;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            o[q] = (__bf16)(0.25f * ((float)a[q] + (float)a1[q] + (float)a2[q] + (float)a3[q]));
        *reinterpret_cast<bf16x8 *>(bufX + po * C1 + c) = o;
    }
    __syncthreads();
    L::e3a(bufX, w[2], bufA); // -> bufA[0 : 16x16xC2)
    __syncthreads();
    L::e3(bufA, w[3], bufA + 16 * 16 * C2); // -> bufA[16x16xC2 : 2 x 16x16xC2)
    __syncthreads();
    L::t3(bufA + 16 * 16 * C2, w[4], bufX); // -> bufX
    __syncthreads();
    // u2 = nearest upsample of t3 + e2, in place in bufS
    for (int e = t; e < 32 * 32 * G1; e += NT) {
        const int c = (e % G1) * 8, p = e / G1, yy = p / 32, xx = p % 32;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(bufX + ((yy / 2) * 16 + xx / 2) * C1 + c);
        bf16x8 k = *reinterpret_cast<const bf16x8 *>(bufS + p * C1 + c);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            k[q] = (__bf16)((float)a[q] + (float)k[q]);
        *reinterpret_cast<bf16x8 *>(bufS + p * C1 + c) = k;
    }
    __syncthreads();
    L::d2(bufS, w[5], bufA);
    __syncthreads();
    L::out(bufA, w[6], out);
}


// ---------------------------------------------------------------- padded coarse16f network
// The coarse16f layers on activations stored with a one-pixel zero border,
// [H+2][W+2][C] with pixel (y, x) at ((y+1)(W+2) + x+1) C: every B fragment
// of the implicit GEMM is then an unconditional LDS load -- a pixel outside
// the image reads the border's zeros -- at a lane address computed once per
// k step plus a compile-time offset per pixel tile, so a wave issues all of
// a step's loads before its MFMAs (no per-fragment bounds test, exec-mask
// branch and wait).  The fragments, their k order and so the results are
// those of net_conv / net_conv_rows.
template <int H, int W, int C>
struct Padded {
    static constexpr int W2 = W + 2, kElems = (H + 2) * (W + 2) * C;
    static __device__ __forceinline__ int at(int y, int x) { return ((y + 1) * W2 + (x + 1)) * C; }
};

// Zeros the border pixels of a padded [H+2][W+2][C] tensor (the whole
// workgroup of NT threads; no barrier).
template <int NT, int H, int W, int C>
__device__ __forceinline__ void zero_border(__bf16 *buf)
{
    constexpr int W2 = W + 2, PER = 2 * W2 + 2 * H, G = C / 8;
    static_assert(C % 8 == 0, "16-byte channel groups");
    const int t = opaque_tid();
    for (int e = t; e < PER * G; e += NT) {
        const int pix = e / G, c = (e % G) * 8;
        int y, x;
        if (pix < W2) {
            y = 0;
            x = pix;
        } else if (pix < 2 * W2) {
            y = H + 1;
            x = pix - W2;
        } else {
            const int rr = pix - 2 * W2;
            y = 1 + (rr >> 1);
            x = (rr & 1) ? W + 1 : 0;
        }
        *reinterpret_cast<bf16x8 *>(buf + (y * W2 + x) * C + c) = zero8();
    }
}

// Output of a padded layer: the interior of a padded tensor (kOutPad), a
// plain [H][W][COUT] tensor (kOutPlain, global memory or LDS), or channels 0
// and 1 rounded to bf16 as FP32 pairs [H*W] (kOutStage: the stage of the
// combine at load, the values fadesim_nn_combine_up reads from a bf16 output).
constexpr int kOutPad = 0, kOutPlain = 1, kOutStage = 2;

template <int OUT, int COUT, int H, int W, bool RELU, int CTW, int P>
__device__ __forceinline__ void store_tiles(const f4 (&acc)[P][CTW], int pg, int cg, int j, int g, void *dst)
{
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int p = 16 * (pg * P + q) + j, y = p / W, x = p % W;
#pragma unroll
        for (int t = 0; t < CTW; ++t) {
            bf16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = acc[q][t][r];
                if (RELU)
                    v = fmaxf(v, 0.f);
                o[r] = (__bf16)v;
            }
            if constexpr (OUT == kOutStage) {
                if (g == 0 && cg * CTW + t == 0)
                    reinterpret_cast<float2 *>(dst)[p] = make_float2((float)o[0], (float)o[1]);
            } else {
                const int px = OUT == kOutPad ? Padded<H, W, COUT>::at(y, x) : p * COUT;
                *reinterpret_cast<bf16x4 *>(reinterpret_cast<__bf16 *>(dst) + px + 16 * (cg * CTW + t) + 4 * g) = o;
            }
        }
    }
}

// net_conv on a padded input (CIN 8 or 16: a 32-deep k step spans 4 or 2
// taps).  Each wave's P pixel tiles start at a row start (W = 16, or P even
// at W = 32), so tile q sits at a fixed offset from tile 0.
template <int NW, int CIN, int COUT, int H, int W, bool RELU, int CTW, int P, int OUT>
__device__ __forceinline__ void net_conv_pad(const __bf16 *src, const __bf16 *__restrict__ w, void *dst)
{
    constexpr int CT = COUT / 16, PT = H * W / 16, K9 = 9 * CIN, KP = net_kpad(CIN), NS = KP / 32, W2 = W + 2;
    constexpr int NPG = PT / P;
    static_assert(W % 16 == 0 && (CIN == 8 || CIN == 16) && COUT % 16 == 0, "tile shapes");
    static_assert(CT % CTW == 0 && PT % P == 0 && (CT / CTW) * NPG == NW, "waves x tiles");
    static_assert(W == 16 || P % 2 == 0, "a wave's tiles start at a row start");
    const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int pg = wave % NPG, cg = wave / NPG;
    const int y00 = 16 * pg * P / W; // first row of the wave's tiles
    f4 acc[P][CTW];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
    bf16x8 one = zero8();
    one[0] = (__bf16)1.0f;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 32 * s + 8 * g;
        bf16x8 av[CTW];
#pragma unroll
        for (int t = 0; t < CTW; ++t) // fragment-major weights: one contiguous 1 KB per wave
            av[t] = *reinterpret_cast<const bf16x8 *>(w + (s * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
        const bool real = k < K9;
        const int tap = real ? k / CIN : 4; // a lane past the taps reads the centre, then selects
        const int ci = real ? k - tap * CIN : 0;
        const int ky = tap / 3, kx = tap - 3 * ky;
        // padded pixel (y00 + ky - 1, j + kx - 1) of tile 0
        const __bf16 *a0 = src + ((y00 + ky) * W2 + j + kx) * CIN + ci;
        bf16x8 b[P];
#pragma unroll
        for (int q = 0; q < P; ++q)
            b[q] = *reinterpret_cast<const bf16x8 *>(a0 + ((16 * q / W) * W2 + (16 * q % W)) * CIN);
        if (s == NS - 1) { // the bias column k = K9 and the zero columns past it
#pragma unroll
            for (int q = 0; q < P; ++q)
                b[q] = real ? b[q] : (k == K9 ? one : zero8());
        }
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int t = 0; t < CTW; ++t)
                acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], b[q], acc[q][t], 0, 0, 0);
    }
    store_tiles<OUT, COUT, H, W, RELU, CTW, P>(acc, pg, cg, j, g, dst);
}

// net_conv_rows on a padded input (CIN a multiple of 32): the left and right
// taps from the centre fragments of the neighbouring lanes, the two edge
// pixels of a 16-pixel segment loaded (the border's zeros at the image edge).
template <int NW, int CIN, int COUT, int H, int W, bool RELU, int CTW, int P, int OUT>
__device__ __forceinline__ void net_conv_rows_pad(const __bf16 *src, const __bf16 *__restrict__ w, void *dst)
{
    constexpr int CT = COUT / 16, PT = H * W / 16, K9 = 9 * CIN, KP = net_kpad(CIN), W2 = W + 2;
    constexpr int NPG = PT / P, CB = CIN / 32;
    static_assert(W % 16 == 0 && CIN % 32 == 0 && COUT % 16 == 0 && KP == K9 + 32, "tile shapes");
    static_assert(CT % CTW == 0 && PT % P == 0 && (CT / CTW) * NPG == NW, "waves x tiles");
    const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int pg = wave % NPG, cg = wave / NPG;
    f4 acc[P][CTW];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            acc[q][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int ky = 0; ky < 3; ++ky) {
        bf16x8 av[3][CB][CTW];
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int t = 0; t < CTW; ++t)
                    av[kx][cb][t] = *reinterpret_cast<const bf16x8 *>(
                        w + (((3 * ky + kx) * CB + cb) * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int p0 = 16 * (pg * P + q), y = p0 / W, x0 = p0 % W, yy = y + ky - 1;
            if (yy < 0 || yy >= H)
                continue; // the whole segment reads the border: the fragments are zero
            const __bf16 *rowp = src + ((yy + 1) * W2 + 1) * CIN + 8 * g;
            bf16x8 b[3][CB];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const bf16x8 c = *reinterpret_cast<const bf16x8 *>(rowp + (x0 + j) * CIN + 32 * cb);
                bf16x8 l = frag_from_left(c), r = frag_from_right(c);
                if (j == 0)
                    l = *reinterpret_cast<const bf16x8 *>(rowp + (x0 - 1) * CIN + 32 * cb);
                if (j == 15)
                    r = *reinterpret_cast<const bf16x8 *>(rowp + (x0 + 16) * CIN + 32 * cb);
                b[0][cb] = l;
                b[1][cb] = c;
                b[2][cb] = r;
            }
#pragma unroll
            for (int kx = 0; kx < 3; ++kx)
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                    for (int t = 0; t < CTW; ++t)
                        acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[kx][cb][t], b[kx][cb], acc[q][t], 0,
                                                                           0, 0);
        }
    }
    {
        bf16x8 av[CTW];
#pragma unroll
        for (int t = 0; t < CTW; ++t)
            av[t] = *reinterpret_cast<const bf16x8 *>(w + ((K9 / 32) * COUT + 16 * (cg * CTW + t) + j) * 32 + 8 * g);
        bf16x8 bv = zero8();
        if (g == 0)
            bv[0] = (__bf16)1.0f;
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int t = 0; t < CTW; ++t)
                acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[q][t], 0, 0, 0);
    }
    store_tiles<OUT, COUT, H, W, RELU, CTW, P>(acc, pg, cg, j, g, dst);
}

// The coarse16f layers on the padded layout, on NW = 16 or 8 waves (the same
// splits of the waves over tiles as CoarseLayers<16, NW>); OUT: how the last
// layer stores.
template <int NW, int OUT>
struct Coarse16Pad;
template <int OUT>
struct Coarse16Pad<16, OUT> {
    static __device__ __forceinline__ void e2a(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<16, 8, 16, 32, 32, true, 1, 4, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e2(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<16, 16, 16, 32, 32, true, 1, 4, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e3a(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<16, 16, 32, 16, 16, true, 1, 2, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e3(const __bf16 *x, const __bf16 *w, void *y) { net_conv_rows_pad<16, 32, 32, 16, 16, true, 1, 2, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void t3(const __bf16 *x, const __bf16 *w, void *y) { net_conv_rows_pad<16, 32, 16, 16, 16, true, 1, 1, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void d2(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<16, 16, 16, 32, 32, true, 1, 4, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void out(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<16, 16, 16, 32, 32, false, 1, 4, OUT>(x, w, y); }
};
template <int OUT>
struct Coarse16Pad<8, OUT> {
    static __device__ __forceinline__ void e2a(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<8, 8, 16, 32, 32, true, 1, 8, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e2(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<8, 16, 16, 32, 32, true, 1, 8, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e3a(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<8, 16, 32, 16, 16, true, 1, 4, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void e3(const __bf16 *x, const __bf16 *w, void *y) { net_conv_rows_pad<8, 32, 32, 16, 16, true, 1, 4, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void t3(const __bf16 *x, const __bf16 *w, void *y) { net_conv_rows_pad<8, 32, 16, 16, 16, true, 1, 2, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void d2(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<8, 16, 16, 32, 32, true, 1, 8, kOutPad>(x, w, y); }
    static __device__ __forceinline__ void out(const __bf16 *x, const __bf16 *w, void *y) { net_conv_pad<8, 16, 16, 32, 32, false, 1, 8, OUT>(x, w, y); }
};

// LDS of the padded coarse16f network, in bf16 elements: bufA (e2a / d2 at
// 34x34x16, or e3a and e3 at 18x18x32 each), bufS (e2 and u2, 34x34x16),
// bufX (the input 34x34x8, p3 and t3 at 18x18x16).
struct Net16PadLds {
    static constexpr int kA = Padded<32, 32, 16>::kElems > 2 * Padded<16, 16, 32>::kElems
                                  ? Padded<32, 32, 16>::kElems
                                  : 2 * Padded<16, 16, 32>::kElems;
    static constexpr int kS = Padded<32, 32, 16>::kElems;
    static constexpr int kX = Padded<32, 32, kFeatC>::kElems > Padded<16, 16, 16>::kElems
                                  ? Padded<32, 32, kFeatC>::kElems
                                  : Padded<16, 16, 16>::kElems;
};

// coarse_net_body for coarse16f on the padded layout: the input features in
// bufX's interior ([34][34][8]), its border zero, and a barrier since; the
// last layer stores as OUT says (kOutPlain: [32][32][16] bf16 to `out`;
// kOutStage: channels 0, 1 as FP32 pairs to `out`, which may overlay bufS).
template <int NW, int OUT>
__device__ __forceinline__ void coarse16_body_pad(__bf16 *bufX, __bf16 *bufA, __bf16 *bufS,
                                                  const __bf16 *const (&w)[7], void *out)
{
    using L = Coarse16Pad<NW, OUT>;
    constexpr int C1 = 16, C2 = 32, G1 = C1 / 8, NT = NW * 64;
    using P32 = Padded<32, 32, C1>;
    using P16 = Padded<16, 16, C1>;
    const int t = opaque_tid();
    L::e2a(bufX, w[0], bufA);
    zero_border<NT, 32, 32, C1>(bufA);
    __syncthreads();
    L::e2(bufA, w[1], bufS); // kept for the skip
    zero_border<NT, 32, 32, C1>(bufS);
    __syncthreads();
    // p3: 2x2 mean of e2 -> 16x16xC1 in bufX
    for (int e = t; e < 16 * 16 * G1; e += NT) {
        const int c = (e % G1) * 8, po = e / G1, yo = po / 16, xo = po % 16;
        const int i00 = P32::at(2 * yo, 2 * xo) + c;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(bufS + i00);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + C1);
        const bf16x8 a2 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + P32::W2 * C1);
        const bf16x8 a3 = *reinterpret_cast<const bf16x8 *>(bufS + i00 + P32::W2 * C1 + C1);
        bf16x8 o;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            o[q] = (__bf16)(0.25f * ((float)a[q] + (float)a1[q] + (float)a2[q] + (float)a3[q]));
        *reinterpret_cast<bf16x8 *>(bufX + P16::at(yo, xo) + c) = o;
    }
    zero_border<NT, 16, 16, C1>(bufX);
    __syncthreads();
    __bf16 *e3o = bufA + Padded<16, 16, C2>::kElems;
    L::e3a(bufX, w[2], bufA);
    zero_border<NT, 16, 16, C2>(bufA);
    __syncthreads();
    L::e3(bufA, w[3], e3o);
    zero_border<NT, 16, 16, C2>(e3o);
    __syncthreads();
    L::t3(e3o, w[4], bufX);
    zero_border<NT, 16, 16, C1>(bufX);
    __syncthreads();
    // u2 = nearest upsample of t3 + e2, in place in bufS's interior
    for (int e = t; e < 32 * 32 * G1; e += NT) {
        const int c = (e % G1) * 8, p = e / G1, yy = p / 32, xx = p % 32;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(bufX + P16::at(yy / 2, xx / 2) + c);
        bf16x8 k = *reinterpret_cast<const bf16x8 *>(bufS + P32::at(yy, xx) + c);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            k[q] = (__bf16)((float)a[q] + (float)k[q]);
        *reinterpret_cast<bf16x8 *>(bufS + P32::at(yy, xx) + c) = k;
    }
    __syncthreads();
    L::d2(bufS, w[5], bufA);
    zero_border<NT, 32, 32, C1>(bufA);
    __syncthreads();
    L::out(bufA, w[6], out);
}

} // namespace nn
} // namespace fadesim
