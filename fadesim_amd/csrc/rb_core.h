// Device core of the red-black SOR crossbar solvers (uniform batches in
// rbsor.hip, mixed sizes in ragged.hip): FP32 sweeps on drop variables, FP64
// adjoint-weighted residual test, one FP64 iterative-refinement step, FP64
// readout.
//
// Network and update rule follow NodebasedGSMethod::nodeBaseGS
// (crossbarHspice/V2/concrete_command.h:1218-1253): the same KCL equations
// relaxed by SOR, here in red-black order (colour of the wordline node u_ij is
// (i+j)&1, of the bitline node d_ij (i+j+1)&1; every coupling joins opposite
// colours, so one colour is updated in parallel while the other is frozen).
//
// Variables are the drops p = vin_i - u_ij >= 0 and q = d_ij - vout_j >= 0
// with c_ij = vin_i - vout_j.  With wire conductance a = 1/r the KCL rows are
//   (g + a_j + a_{j+1}) p - a p_{j-1} - a p_{j+1} + g q = g c   (p_{-1} = 0)
//   (g + b_i + b_{i+1}) q - b q_{i-1} - b q_{i+1} + g p = g c   (q_n = 0)
// (no a_{j+1} term at the open right end, no b_i term at the open top).
// SOR with the diagonal taken as 2a (a at the open ends) -- the cell term is
// ~1e-6 of it -- gives for a node x with line neighbours x1, x2:
//   tt = fma(-h, x + x_other, rho),  x' = fma(w/2, x1 + x2, fma(1-w, x, tt))
//   h = w g / (2a),  rho = w rhs / (2a)  (rhs = g c, or the FP64 residual
//   after refinement).  At an open end (D~ = a) the missing neighbour is the
//   ghost x2 = x1 + (2/w) tt, which turns the update into the exact one
//   (1-w) x + w x1 + 2 tt.
//
// Layout: one workgroup per crossbar -- or, for crossbars too large for one
// CU, per strip of full rows (a team of workgroups solves the crossbar, see
// StripCtx).  Thread (ty, tx) owns the TR x TC cell tile rows TR*ty..,
// columns TC*tx.. with p, q, h, rho_p, rho_q in VGPRs.  Per half-sweep a
// thread reads (TR + TC) halo values from LDS (the neighbours' edge values of
// the other colour), updates its TR*TC cells and publishes the TR + TC edge
// values it changed.  Conductances come from HBM once per sweep batch
// (coefficient reload after each convergence test).
//
// Convergence test (every few sweeps, FP64): residual r = b - A x at every
// node, then the first-order error of the readout
//   D_j = sum z_j . r,   I_j(x) - I_j* ~= D_j
// with the adjoint z_j taken from the wire Green's functions (bitline:
// r (n - max(i,i')) weighted by g of column j; wordline: r (min(j,j')+1)
// weighted by g_ij).  Stop when max_j |D_j| / |I_j| * safety <= tol.
//
// FP32 sweeps stall near 1e-6 relative error, so once the estimate falls
// below switch_tol the solve takes one iterative-refinement step: the exact
// FP64 residual becomes the right-hand side of a correction solve, whose
// initial guess is corr_iters wire Green's-function (IR2S) iterations on the
// correction (phase_green) rather than zero.  A solve that meets the
// tolerance at a test stops there, refined or not.
//
// Initial guess: zero drops or the caller's field, then ir2s_iters IR2S
// iterations (phase_ir2s: every wire solved exactly for the present cell
// currents, optionally damped).  The iteration contracts by ~0.8 r gbar n m
// per step, so a few of them bring a 128x128 crossbar to the tolerance before
// the first sweep.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fadesim.h"

namespace fadesim {
namespace rb {

constexpr int kWsPlanes = 5; // workspace planes per slot: zq | refinement rho p, q | snapshot p, q
constexpr int kMaxStrips = 16;
constexpr int kRowsLd = 512;                      // row stride of the strip exchange buffers
constexpr unsigned long long kWaitTicks = 500000000ull; // 5 s of the 100 MHz constant clock

// Solve controls shared by every instance of a size class.
struct Ctl {
    float omega;
    double tol_eff;    // tol / safety
    double switch_tol; // refinement switch point
    double dwr_scale;  // DWR: stop at est <= tol_eff * dwr_scale
    int max_sweeps;
    unsigned flags;
    int check_first;   // sweeps before the first test
    int check_min, check_max; // bounds of the predicted test interval
    int ir2s_iters;    // wire-solve (IR2S) iterations applied to the initial guess
    int corr_iters;    // IR2S iterations of the refinement correction's guess (>= 1)
    double ir2s_omega; // damping of the IR2S iterations, (0, 1]; < 0: Chebyshev schedule (ir2s_step)
    double ir2s_geo;   // the schedule's geometry factor, ir2s_geo(rows, cols)
};

// One crossbar.  rows is the full row count (all strips).
struct Inst {
    const float *g;    // [rows][N]
    double rwire;
    const double *vin; // [rows]
    const double *vout; // [N]; never null (zeros supplied by the host)
    const float *warm; // [2][rows][N] or null
    float *field;      // [2][rows][N] or null
    double *iout;      // [N]
    int *sweeps;       // one element or null
    int *status;
    double *est;
    int *rejected;     // one element or null: 1 when the initial guess was discarded for the zero guess
    int rows;
    // optional: warm = base + cscale * bilinear(coarse[.][.][0..1]), the
    // coarse grid 2^clevel times coarser (rbsor only; null elsewhere)
    const __bf16 *coarse;
    float cscale;
    int cch, clevel;
};

// Team of S workgroups solving one crossbar, strip s = rows [s*M, (s+1)*M).
// Neighbouring strips exchange their boundary bitline rows after every
// half-sweep through `rows` (double-buffered by publication parity) and
// per-strip publication counters `flag` (device-coherent stores and loads,
// team_st / team_ld, with no cache maintenance);
// column sums of the convergence test and of the set-up go through `part` /
// `setup` and the counter barrier `cnt`.  A wait that does not end within
// kWaitTicks raises `abort`, which every wait checks, so no workgroup spins
// forever.
struct StripCtx {
    int S, s;
    float *rows;   // [2 parity][S][2 top/bottom][kRowsLd]
    int *flag;     // [S][2] publications of the top / bottom row
    int *cnt;      // barrier arrivals
    double *part;  // [2 parity][S][2][kRowsLd]
    double *setup; // [S][4][kRowsLd]
    int *abort;
};

// Global-memory views of plain pointers.  The phases below are functions of
// their own that receive their buffers as generic pointers; loads through a
// generic pointer are FLAT instructions, which count against the LDS wait
// counter as well, so every wait for an LDS result would also wait for the
// global loads in flight.  Every buffer passed to the solver is device
// global memory.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ T ldg(const T *p) { return *gmem(p); }

typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float *p)
{
    const f32x4 v = ldg(reinterpret_cast<const f32x4 *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d)
{
    *gmem(reinterpret_cast<f32x4 *>(p)) = f32x4{a, b, c, d};
}

// TC consecutive floats of a row (TC a multiple of 4).
template <int TC>
__device__ __forceinline__ void ldrow(const float *p, float (&v)[TC])
{
#pragma unroll
    for (int k = 0; k < TC; k += 4) {
        const float4 x = ld4(p + k);
        v[k] = x.x;
        v[k + 1] = x.y;
        v[k + 2] = x.z;
        v[k + 3] = x.w;
    }
}
template <int TC>
__device__ __forceinline__ void strow(float *p, const float (&v)[TC])
{
#pragma unroll
    for (int k = 0; k < TC; k += 4)
        st4(p + k, v[k], v[k + 1], v[k + 2], v[k + 3]);
}

// The same for a row in LDS (16-byte aligned).
template <int TC>
__device__ __forceinline__ void ldrow_lds(const float *p, float (&v)[TC])
{
#pragma unroll
    for (int k = 0; k < TC; k += 4) {
        const float4 x = *reinterpret_cast<const float4 *>(p + k);
        v[k] = x.x;
        v[k + 1] = x.y;
        v[k + 2] = x.z;
        v[k + 3] = x.w;
    }
}
template <int TC>
__device__ __forceinline__ void strow_lds(float *p, const float (&v)[TC])
{
#pragma unroll
    for (int k = 0; k < TC; k += 4)
        *reinterpret_cast<float4 *>(p + k) = make_float4(v[k], v[k + 1], v[k + 2], v[k + 3]);
}

// Geometry: TR x TC cells per thread, TY x TX threads, M x N cells per
// workgroup (a whole crossbar, or one strip of full rows).
template <int TR_, int TC_, int TY_, int TX_>
struct Geo {
    static constexpr int TR = TR_, TC = TC_, TY = TY_, TX = TX_;
    static constexpr int NT = TY * TX;
    static constexpr int M = TR * TY, N = TC * TX;
    static constexpr int NW = NT / 64;
    // LDS halo planes: p left/right columns [TR][NT]; q top/bottom rows
    // [TC][NT + 2 TX] with thread t at TX + t -- the virtual thread rows -1 and
    // TY hold the neighbouring strips' boundary rows.
    static constexpr int QLD = NT + 2 * TX;
    static constexpr int PL = 0, PR = TR * NT, QT = 2 * TR * NT, QB = 2 * TR * NT + TC * QLD;
    static constexpr int HALO = 2 * TR * NT + 2 * TC * QLD;
    static_assert(TR % 2 == 0 && TC % 4 == 0, "tile sides: TR even, TC multiple of 4");
    static_assert((NT % 64 == 0 && 64 % TX == 0) || TX % 64 == 0, "tile rows must not straddle waves");
    static_assert(NT >= M && NT >= N, "one thread per row and per column in the test");
    static_assert(N <= kRowsLd, "strip rows exceed the exchange buffer");
};

template <int COL, class G>
__device__ __forceinline__ void half_sweep(float (&yp)[G::TR][G::TC], float (&yq)[G::TR][G::TC],
                                           const float (&hn)[G::TR][G::TC],
                                           const float (&rp)[G::TR][G::TC],
                                           const float (&rq)[G::TR][G::TC], float *halo, int t,
                                           bool has_l, bool has_r, bool has_u, bool has_d,
                                           float cA, float cB, float c2)
{
    constexpr int TR = G::TR, TC = G::TC, NT = G::NT, TX = G::TX, QLD = G::QLD;
    float hl[TR], hr[TR], hu[TC], hd[TC];
#pragma unroll
    for (int k = 0; k < TR; ++k) {
        hl[k] = hr[k] = 0.f;
        if ((k & 1) == COL) { // left neighbour of p at (k,0); the source at j = 0
            const float v = halo[G::PR + k * NT + (has_l ? t - 1 : t)];
            hl[k] = has_l ? v : 0.f;
        } else { // right neighbour of p at (k,TC-1)
            hr[k] = halo[G::PL + k * NT + (has_r ? t + 1 : t)];
        }
    }
#pragma unroll
    for (int k = 0; k < TC; ++k) {
        hu[k] = hd[k] = 0.f;
        if ((k & 1) != COL) { // upper neighbour of q at (0,k)
            hu[k] = halo[G::QB + k * QLD + t]; // thread row ty-1 (virtual row -1: upper strip)
        } else { // lower neighbour of q at (TR-1,k); the sink below the last row
            const float v = halo[G::QT + k * QLD + 2 * TX + t]; // thread row ty+1
            hd[k] = has_d ? v : 0.f;
        }
    }
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const float s = yp[ii][jj] + yq[ii][jj];
            if (((ii + jj) & 1) == COL) {
                const float tt = fmaf(hn[ii][jj], s, rp[ii][jj]);
                const float n1 = jj > 0 ? yp[ii][jj - 1] : hl[ii];
                float n2;
                if (jj < TC - 1)
                    n2 = yp[ii][jj + 1];
                else
                    n2 = has_r ? hr[ii] : fmaf(c2, tt, n1); // open right end
                yp[ii][jj] = fmaf(cB, n1 + n2, fmaf(cA, yp[ii][jj], tt));
            } else {
                const float tt = fmaf(hn[ii][jj], s, rq[ii][jj]);
                const float n2 = ii < TR - 1 ? yq[ii + 1][jj] : hd[jj];
                float n1;
                if (ii > 0)
                    n1 = yq[ii - 1][jj];
                else
                    n1 = has_u ? hu[jj] : fmaf(c2, tt, n2); // open top end
                yq[ii][jj] = fmaf(cB, n1 + n2, fmaf(cA, yq[ii][jj], tt));
            }
        }
    }
#pragma unroll
    for (int k = 0; k < TR; ++k) {
        if ((k & 1) == COL)
            halo[G::PL + k * NT + t] = yp[k][0];
        else
            halo[G::PR + k * NT + t] = yp[k][TC - 1];
    }
#pragma unroll
    for (int k = 0; k < TC; ++k) {
        if ((k & 1) != COL)
            halo[G::QT + k * QLD + TX + t] = yq[0][k];
        else
            halo[G::QB + k * QLD + TX + t] = yq[TR - 1][k];
    }
    __syncthreads();
}

template <class G>
__device__ __forceinline__ void publish_all(const float (&yp)[G::TR][G::TC],
                                            const float (&yq)[G::TR][G::TC], float *halo, int t)
{
    constexpr int NT = G::NT, QLD = G::QLD, TX = G::TX;
#pragma unroll
    for (int k = 0; k < G::TR; ++k) {
        halo[G::PL + k * NT + t] = yp[k][0];
        halo[G::PR + k * NT + t] = yp[k][G::TC - 1];
    }
#pragma unroll
    for (int k = 0; k < G::TC; ++k) {
        halo[G::QT + k * QLD + TX + t] = yq[0][k];
        halo[G::QB + k * QLD + TX + t] = yq[G::TR - 1][k];
    }
}

// Adds scale x bilinear(coarse) to the loaded warm-start drops: the network
// output of a grid 2^level times coarser than the crossbar (channel 0 to the
// wordline drops, 1 to the bitline drops; an even channel count, so the pair
// is one 4-byte load), staged as FP32 pairs in `stage` (LDS, 2 (rows >>
// level) (N >> level) floats).  Sample point ((i + 1/2)/f - 1/2,
// (j + 1/2)/f - 1/2) clamped to the coarse grid, the arithmetic of
// fadesim_nn_combine_up term by term, so the drops equal that kernel's field
// bit for bit.  add_coarse_staged: the sampling from a filled stage (level
// lvl, scale s); add_coarse: staged from I.coarse, then sampled, ending with
// a workgroup barrier (the stage is free again).
template <class G>
__device__ __forceinline__ void add_coarse_staged(const float *stage, int lvl, float s, int rows, int gi0, int i0,
                                                  int j0, float (&yp)[G::TR][G::TC], float (&yq)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, N = G::N;
    const int f = 1 << lvl, Hc = rows >> lvl, Wc = N >> lvl;
    const float2 *st2 = reinterpret_cast<const float2 *>(stage);
    const float inv = 1.0f / (float)f;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        const int i = gi0 + i0 + ii;
        const float y = fminf(fmaxf(((float)i + 0.5f) * inv - 0.5f, 0.0f), (float)(Hc - 1));
        const int y0 = (int)y, y1 = min(y0 + 1, Hc - 1);
        const float wy = y - (float)y0;
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const int j = j0 + jj;
            const float x = fminf(fmaxf(((float)j + 0.5f) * inv - 0.5f, 0.0f), (float)(Wc - 1));
            const int x0 = (int)x, x1 = min(x0 + 1, Wc - 1);
            const float wx = x - (float)x0;
            const float2 a00 = st2[y0 * Wc + x0], a01 = st2[y0 * Wc + x1];
            const float2 a10 = st2[y1 * Wc + x0], a11 = st2[y1 * Wc + x1];
            {
                const float top = a00.x + wx * (a01.x - a00.x), bot = a10.x + wx * (a11.x - a10.x);
                yp[ii][jj] = yp[ii][jj] + s * (top + wy * (bot - top));
            }
            {
                const float top = a00.y + wx * (a01.y - a00.y), bot = a10.y + wx * (a11.y - a10.y);
                yq[ii][jj] = yq[ii][jj] + s * (top + wy * (bot - top));
            }
        }
    }
}

// Stages channels 0 and 1 of a coarse grid of hw pixels with cch channels
// per pixel (bf16, an even count, so the pair is one 4-byte word) as FP32
// pairs in `stage` (bf16 -> FP32 exactly: the bits shifted up).  src is
// global memory or LDS.
__device__ __forceinline__ void stage_coarse_pair(float2 &dst, unsigned v)
{
    dst = make_float2(__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u));
}

template <class G>
__device__ __forceinline__ void add_coarse(const Inst &I, float *stage, int t, int gi0, int i0, int j0,
                                           float (&yp)[G::TR][G::TC], float (&yq)[G::TR][G::TC])
{
    const int lvl = I.clevel, Hc = I.rows >> lvl, Wc = G::N >> lvl, HW = Hc * Wc;
    // opaque to the optimizer: the sample positions below are computed here,
    // not hoisted out of a caller's loop over crossbars as a register tile
    // (which would spill)
    asm volatile("" : "+v"(i0), "+v"(j0));
    __syncthreads();
    const unsigned *src = reinterpret_cast<const unsigned *>(I.coarse);
    float2 *st2 = reinterpret_cast<float2 *>(stage);
    for (int k = t; k < HW; k += G::NT)
        stage_coarse_pair(st2[k], ldg(src + (size_t)k * (I.cch >> 1)));
    __syncthreads();
    add_coarse_staged<G>(stage, lvl, I.cscale, I.rows, gi0, i0, j0, yp, yq);
    __syncthreads();
}

// ---------------------------------------------------------------- strip exchange

// Spin until *p >= target (relaxed: no cache maintenance).  All active
// lanes of the wave call it together; false when the team aborted or the wait
// timed out.
__device__ __forceinline__ bool poll_geq(const int *p, int target, int *abort)
{
    if (__hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(gmem(abort), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                return false;
            if (wall_clock64() - t0 > kWaitTicks) {
                __hip_atomic_store(gmem(abort), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    return true;
}

// poll_geq, then an agent-scope acquire: ordinary loads after it see what the
// releasing side stored before its release.
__device__ __forceinline__ bool wait_geq(const int *p, int target, int *abort)
{
    if (!poll_geq(p, target, abort))
        return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// Team exchanges carry no whole-cache maintenance (an agent-scope release
// writes back the XCD's L2, an acquire invalidates it: ~15 us per exchange
// on a busy chip).  Instead every value a strip publishes is stored with a
// device-coherent store (team_st), the wave waits until those stores have
// completed (team_commit) before the publishing counter moves, and readers,
// once they have seen the counter, read with device-coherent loads (team_ld),
// which do not return stale cached lines.
__device__ __forceinline__ void team_st(double *p, double v)
{
    __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void team_st(float *p, float v)
{
    __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double team_ld(const double *p)
{
    return __hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float team_ld(const float *p)
{
    return __hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A pair of floats as one 8-byte access.
__device__ __forceinline__ void team_st2(unsigned long long *p, float lo, float hi)
{
    const unsigned long long v = (unsigned long long)__float_as_uint(lo) | ((unsigned long long)__float_as_uint(hi) << 32);
    __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void team_ld2(const unsigned long long *p, float &lo, float &hi)
{
    const unsigned long long v = __hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    lo = __uint_as_float((unsigned)v);
    hi = __uint_as_float((unsigned)(v >> 32));
}
// This wave's earlier vector-memory operations have completed (a store: it
// is visible at device scope); also a compiler barrier.
__device__ __forceinline__ void team_commit() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Publication `pub` (0-based) of this strip's boundary bitline rows: the
// top-row threads write row 0, the bottom-row threads row M-1, each group
// then moves its counter to pub + 1 (a thread row lies within one wave, so
// the wave's team_commit covers all of the row's stores).
template <class G>
__device__ __forceinline__ void strip_publish(const StripCtx &X, const float (&yq)[G::TR][G::TC], int ty,
                                              int j0, int lane, int pub)
{
    const int par = pub & 1;
    if (ty == 0 && X.s > 0) {
        float *dst = X.rows + (((size_t)par * X.S + X.s) * 2 + 0) * kRowsLd + j0;
#pragma unroll
        for (int k = 0; k < G::TC; ++k)
            team_st(dst + k, yq[0][k]);
        team_commit();
        if ((lane % G::TX) == 0)
            __hip_atomic_store(gmem(X.flag + 2 * X.s + 0), pub + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (ty == G::TY - 1 && X.s < X.S - 1) {
        float *dst = X.rows + (((size_t)par * X.S + X.s) * 2 + 1) * kRowsLd + j0;
#pragma unroll
        for (int k = 0; k < G::TC; ++k)
            team_st(dst + k, yq[G::TR - 1][k]);
        team_commit();
        if ((lane % G::TX) == 0)
            __hip_atomic_store(gmem(X.flag + 2 * X.s + 1), pub + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Loads the neighbours' publication `pub` into the virtual halo rows: the
// upper strip's bottom row above thread row 0, the lower strip's top row below
// thread row TY-1.  Ends with a workgroup barrier; returns false (uniformly)
// when the team aborted.
template <class G>
__device__ __forceinline__ bool strip_fetch(const StripCtx &X, float *halo, int *s_dead, int tx, int ty, int pub)
{
    constexpr int TC = G::TC, TX = G::TX, QLD = G::QLD;
    const int par = pub & 1;
    const int j0 = TC * tx;
    if (ty == 0 && X.s > 0) {
        if (poll_geq(X.flag + 2 * (X.s - 1) + 1, pub + 1, X.abort)) {
            const float *src = X.rows + (((size_t)par * X.S + X.s - 1) * 2 + 1) * kRowsLd + j0;
            float v[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k)
                v[k] = team_ld(src + k);
#pragma unroll
            for (int k = 0; k < TC; ++k)
                halo[G::QB + k * QLD + tx] = v[k]; // virtual thread row -1
        } else {
            *s_dead = 1;
        }
    }
    if (ty == G::TY - 1 && X.s < X.S - 1) {
        if (poll_geq(X.flag + 2 * (X.s + 1) + 0, pub + 1, X.abort)) {
            const float *src = X.rows + (((size_t)par * X.S + X.s + 1) * 2 + 0) * kRowsLd + j0;
            float v[TC];
#pragma unroll
            for (int k = 0; k < TC; ++k)
                v[k] = team_ld(src + k);
#pragma unroll
            for (int k = 0; k < TC; ++k)
                halo[G::QT + k * QLD + TX + G::NT + tx] = v[k]; // virtual thread row TY
        } else {
            *s_dead = 1;
        }
    }
    __syncthreads();
    return *s_dead == 0;
}

// Team sums of the two FP64 values every strip published for column t, at
// base + s stride + t and base + s stride + kRowsLd + t: over the strips
// above this one (E0, E1) and over the whole team (T0, T1), each accumulated
// in strip order.  The loads go out in batches of kBatch strips, each batch
// in flight together, so the exchange costs kMaxStrips / kBatch memory
// latencies instead of one per strip, at 4 kBatch VGPRs.
__device__ __forceinline__ void team_sums(const double *base, size_t stride, int S, int below, int t, double &E0,
                                          double &E1, double &T0, double &T1)
{
    constexpr int kBatch = 8;
    E0 = E1 = T0 = T1 = 0.0;
#pragma unroll
    for (int s0 = 0; s0 < kMaxStrips; s0 += kBatch) {
        if (s0 >= S)
            break;
        double a[kBatch], b[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const double *ps = base + (size_t)(s0 + k < S ? s0 + k : 0) * stride;
            a[k] = team_ld(ps + t);
            b[k] = team_ld(ps + kRowsLd + t);
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            if (s0 + k < below) {
                E0 += a[k];
                E1 += b[k];
            }
            if (s0 + k < S) {
                T0 += a[k];
                T1 += b[k];
            }
        }
    }
}

// team_sums over FP32 pairs (lo, hi) published at base + s stride + t, one
// 8-byte load per strip, accumulated in FP64 in strip order.
__device__ __forceinline__ void team_sums2(const unsigned long long *base, size_t stride, int S, int below, int t,
                                           double &E0, double &E1, double &T0, double &T1)
{
    constexpr int kBatch = 8;
    E0 = E1 = T0 = T1 = 0.0;
#pragma unroll
    for (int s0 = 0; s0 < kMaxStrips; s0 += kBatch) {
        if (s0 >= S)
            break;
        float a[kBatch], b[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k)
            team_ld2(base + (size_t)(s0 + k < S ? s0 + k : 0) * stride + t, a[k], b[k]);
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            if (s0 + k < below) {
                E0 += (double)a[k];
                E1 += (double)b[k];
            }
            if (s0 + k < S) {
                T0 += (double)a[k];
                T1 += (double)b[k];
            }
        }
    }
}

// Team barrier number `nbar` (0-based): every member's team_st stores before
// it are visible to every member's team_ld loads after it.  Ends with a workgroup barrier; false
// (uniformly) when the team aborted.
__device__ __forceinline__ bool team_barrier(const StripCtx &X, int *s_dead, int t, int nbar)
{
    team_commit();
    __syncthreads();
    if (t == 0) {
        __hip_atomic_fetch_add(gmem(X.cnt), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!poll_geq(X.cnt, X.S * (nbar + 1), X.abort))
            *s_dead = 1;
    }
    __syncthreads();
    return *s_dead == 0;
}

// The same barrier in two halves, for work that needs no team data in
// between: team_arrive counts this member in (after its team stores have
// completed), team_wait waits for the others.
__device__ __forceinline__ void team_arrive(const StripCtx &X, int t)
{
    team_commit();
    __syncthreads();
    if (t == 0)
        __hip_atomic_fetch_add(gmem(X.cnt), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool team_wait(const StripCtx &X, int *s_dead, int t, int nbar)
{
    if (t == 0 && !poll_geq(X.cnt, X.S * (nbar + 1), X.abort))
        *s_dead = 1;
    __syncthreads();
    return *s_dead == 0;
}

// ---------------------------------------------------------------- coefficients

// Sweep coefficients: hn = -(w/2a) g and rho.  Before the refinement
// rho = (w/2a) g c for both nodes of a cell; after it rho is the scaled FP64
// residual the refinement step stored in the workspace.
template <class G>
__device__ __forceinline__ void load_coeffs(const Inst &I, const float *wsb, const float *gp, int i0, int j0,
                                            int gi0, double hcoef, bool refined, float (&hn)[G::TR][G::TC],
                                            float (&rp)[G::TR][G::TC], float (&rq)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, M = G::M, N = G::N;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        float g[TC];
        ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, g);
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            hn[ii][jj] = (float)(-((double)g[jj] * hcoef));
        if (refined) {
            ldrow<TC>(wsb + 1 * M * N + (size_t)(i0 + ii) * N + j0, rp[ii]);
            ldrow<TC>(wsb + 2 * M * N + (size_t)(i0 + ii) * N + j0, rq[ii]);
        } else {
            const double vi = ldg(I.vin + (gi0 + i0 + ii));
#pragma unroll
            for (int jj = 0; jj < TC; ++jj) {
                const double c = vi - ldg(I.vout + (j0 + jj));
                rp[ii][jj] = rq[ii][jj] = (float)(((double)g[jj] * c) * hcoef);
            }
        }
    }
}

struct Shared {
    float *halo;
    double *S1, *Ifloor, *colI, *colD;
    float *wmax;
    int *dead;
    double *cbuf; // [2][TY][N] per-thread-row column partials
    double *cext; // [3][N] column values shared by the workgroup
    double *gsum; // [1] total conductance of the crossbar (all strips), from the set-up
};

// ---------------------------------------------------------------- tile reductions
// The convergence test, the set-up and the refinement guess work on each
// thread's register tile: sums along a row are shuffle scans over the TX
// lanes of a thread row, sums along a column are LDS scans over the TY thread
// rows (completed over the team for strips).  Accumulation is FP64.

// v of the lane `CTRL` selects (a DPP control, both 32-bit halves moved
// alike), +0 where that lane lies outside the row of 16 or the destination
// row is masked off by ROW_MASK.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Exclusive prefix over the TX lanes of a thread row, and the row total:
// the additions of a shift-by-1, 2, 4, ... scan, by DPP row shifts for rows
// of 16 lanes or more (a lane with nothing to its left adds +0), by lane
// shuffles for narrower rows.
template <int TX>
__device__ __forceinline__ void row_scan(double x, int tx, double &excl, double &total)
{
    double incl = x;
    if constexpr (TX < 16) {
#pragma unroll
        for (int d = 1; d < TX; d <<= 1) {
            const double y = __shfl_up(incl, (unsigned)d, TX);
            if (tx >= d)
                incl += y;
        }
    } else {
        incl += dpp_d<0x111>(incl); // row_shr:1
        incl += dpp_d<0x112>(incl); // row_shr:2
        incl += dpp_d<0x114>(incl); // row_shr:4
        incl += dpp_d<0x118>(incl); // row_shr:8
        if constexpr (TX >= 32)
            incl += dpp_d<0x142, 0xa>(incl); // row_bcast:15
        if constexpr (TX == 64)
            incl += dpp_d<0x143, 0xc>(incl); // row_bcast:31
    }
    total = __shfl(incl, TX - 1, TX);
    excl = incl - x;
}

// FP32 forms for the wire-solve iterations of the initial guess, which only
// need the drops to ~1e-5 relative (the drops are ~1e-3 of the cell voltage,
// and every result is certified by the FP64 test); products are fused
// (fmaf) here, unlike in the sweeps, whose CPU replays are bit-exact.
//
// v from the lane `CTRL` selects (a DPP control), 0 where that lane lies
// outside the row of 16 or the destination row is masked off by ROW_MASK.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

// Inclusive prefix over the TX lanes of a thread row.  TX = 16, 32 or 64
// (rows aligned to 16-lane DPP rows): shifts by 1, 2, 4, 8 within each row of
// 16, then row 15 -> rows 1, 3 and lane 31 -> rows 2, 3 (wave64 row
// broadcasts); narrower rows use lane shuffles.  The row total comes from the
// row's last lane.
template <int TX>
__device__ __forceinline__ void row_scan_f(float x, int tx, float &excl, float &total)
{
    static_assert(TX == 4 || TX == 8 || TX == 16 || TX == 32 || TX == 64, "thread rows of 4 to 64 lanes");
    float incl = x;
    if constexpr (TX < 16) {
#pragma unroll
        for (int d = 1; d < TX; d <<= 1) {
            const float y = __shfl_up(incl, (unsigned)d, TX);
            if (tx >= d)
                incl += y;
        }
        total = __shfl(incl, TX - 1, TX);
        excl = incl - x;
        return;
    }
    incl += dpp_f<0x111>(incl); // row_shr:1
    incl += dpp_f<0x112>(incl); // row_shr:2
    incl += dpp_f<0x114>(incl); // row_shr:4
    incl += dpp_f<0x118>(incl); // row_shr:8
    if constexpr (TX >= 32)
        incl += dpp_f<0x142, 0xa>(incl); // row_bcast:15
    if constexpr (TX == 64) {
        incl += dpp_f<0x143, 0xc>(incl); // row_bcast:31
        total = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(incl), 63));
    } else {
        total = __shfl(incl, TX - 1, TX);
    }
    excl = incl - x;
}

template <class G>
__device__ __forceinline__ void row_green_row_f(const float (&v)[G::TC], int j0, int tx, float scale,
                                                float (&out)[G::TC])
{
    constexpr int TC = G::TC;
    // the column indices j + 1 as floats: one conversion, then exact
    // additions of small integers
    const float j1 = (float)(j0 + 1);
    float L0 = 0.f, L1 = 0.f;
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        L0 += v[jj];
        L1 = fmaf(j1 + (float)jj, v[jj], L1);
    }
    float P0, P1, R, R1;
    row_scan_f<G::TX>(L0, tx, P0, R);
    row_scan_f<G::TX>(L1, tx, P1, R1);
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        const float jp1 = j1 + (float)jj;
        out[jj] = scale * fmaf(jp1, R - P0, P1);
        P0 += v[jj];
        P1 = fmaf(jp1, v[jj], P1);
    }
}

// Per-column values in LDS (S1, Ifloor, colI, colD, cext and the rows of
// cbuf) are kept in slot order: column j0 + jj of thread column tx sits in
// slot jj * TX + tx, so the TX threads of a thread row write one jj at
// consecutive addresses (no LDS bank conflicts), and the thread t < N that
// scans a column handles slot t, i.e. column col_of_slot(t).  The order of
// every sum is unchanged.
template <class G>
__device__ __forceinline__ int slot_of(int tx, int jj)
{
    return jj * G::TX + tx;
}
template <class G>
__device__ __forceinline__ int col_of_slot(int s)
{
    return (s % G::TX) * G::TC + s / G::TX;
}

// Column sums over the workgroup's rows of two per-thread partials (thread
// row ty, columns j0 + jj): outA, outB (LDS, N entries each, slot order).
template <class G>
__device__ __forceinline__ void col_reduce2(const Shared &sh, const double (&a)[G::TC], const double (&b)[G::TC], int t,
                                            int ty, int j0, double *outA, double *outB)
{
    constexpr int TC = G::TC, TY = G::TY, N = G::N;
    double *ba = sh.cbuf, *bb = sh.cbuf + TY * N;
    const int tx = j0 / TC;
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        ba[ty * N + slot_of<G>(tx, jj)] = a[jj];
        bb[ty * N + slot_of<G>(tx, jj)] = b[jj];
    }
    __syncthreads();
    if (t < N) {
        double sa = 0.0, sb = 0.0;
        for (int q = 0; q < TY; ++q) {
            sa += ba[q * N + t];
            sb += bb[q * N + t];
        }
        outA[t] = sa;
        outB[t] = sb;
    }
    __syncthreads();
}

// The middle of a bitline Green's sum: with every thread row's column partials
// (sum v, sum v (n - i)) in cbuf [2][TY][N], turn them in place into
// exclusive prefixes over the thread rows and put the prefix of the rows
// above this workgroup (strips: over the team, team barrier nbar++) and the
// crossbar total of sum v (n - i) in cext [3][N].  Returns false when the
// team aborted.
// F32X: the team exchange carries the workgroup totals as FP32 pairs (half
// the bytes of the FP64 exchange; for the FP32 wire-solve iterations).
template <class G, bool STRIP, bool F32X = false>
__device__ __forceinline__ double col_scan_begin(const Shared &sh, const StripCtx &X, int t, int nbar)
{
    constexpr int TY = G::TY, N = G::N;
    double *b0 = sh.cbuf, *b1 = sh.cbuf + TY * N;
    __syncthreads();
    // exclusive prefix over the thread rows, in place; totals of this workgroup
    double T0 = 0.0, T1 = 0.0;
    if (t < N) {
        for (int q = 0; q < TY; ++q) {
            const double x0 = b0[q * N + t], x1 = b1[q * N + t];
            b0[q * N + t] = T0;
            b1[q * N + t] = T1;
            T0 += x0;
            T1 += x1;
        }
    }
    if (STRIP) {
        double *mine = X.part + (((size_t)(nbar & 1) * X.S + X.s) * 2) * kRowsLd;
        if (t < N) {
            if (F32X) {
                team_st2(reinterpret_cast<unsigned long long *>(mine) + t, (float)T0, (float)T1);
            } else {
                team_st(mine + t, T0);
                team_st(mine + kRowsLd + t, T1);
            }
        }
        team_arrive(X, t);
    }
    return T1;
}

template <class G, bool STRIP, bool F32X = false>
__device__ __forceinline__ bool col_scan_end(const Shared &sh, const StripCtx &X, int t, int &nbar, double T1)
{
    constexpr int N = G::N;
    double E0 = 0.0, E1 = 0.0, Tall = T1;
    if (STRIP) {
        if (!team_wait(X, sh.dead, t, nbar++))
            return false;
        if (t < N) {
            double T0;
            const double *base = X.part + ((size_t)((nbar - 1) & 1) * X.S * 2) * kRowsLd;
            if (F32X)
                team_sums2(reinterpret_cast<const unsigned long long *>(base), 2 * kRowsLd, X.S, X.s, t, E0, E1, T0,
                           Tall);
            else
                team_sums(base, 2 * kRowsLd, X.S, X.s, t, E0, E1, T0, Tall);
        }
    }
    if (t < N) {
        sh.cext[t] = E0;
        sh.cext[N + t] = E1;
        sh.cext[2 * N + t] = Tall;
    }
    __syncthreads();
    return true;
}

// col_scan_begin then col_scan_end: for strips, work placed between the two
// overlaps the team barrier.
template <class G, bool STRIP>
__device__ __forceinline__ bool col_scan_exchange(const Shared &sh, const StripCtx &X, int t, int &nbar)
{
    const double T1 = col_scan_begin<G, STRIP>(sh, X, t, nbar);
    return col_scan_end<G, STRIP>(sh, X, t, nbar, T1);
}

// Bitline Green's sum of a float tile: out(ii, jj) = scale sum_i' (n -
// max(i,i')) v(i',j) over all n rows of the crossbar (i = gi0 + i0 + ii,
// j = j0 + jj).  The sums across thread rows are FP64 (strips complete them
// over the team, team barrier nbar++); within the tile FP32 from the FP64
// anchors -- the current entering the tile and the weighted sum below it --
// so no difference of large FP32 sums is formed.  Returns false when the
// team aborted.
template <class G, bool STRIP>
__device__ __forceinline__ bool col_green_tile_f(const Shared &sh, const StripCtx &X, const float (&v)[G::TR][G::TC],
                                                 float scale, int n, int gi0, int i0, int j0, int ty, int t, int &nbar,
                                                 float (&out)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, TY = G::TY, N = G::N;
    double *b0 = sh.cbuf, *b1 = sh.cbuf + TY * N;
    const int tx = j0 / TC;
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        float s0 = 0.f, s1 = 0.f;
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            s0 += v[ii][jj];
            s1 += v[ii][jj] * (float)(n - (gi0 + i0 + ii));
        }
        b0[ty * N + slot_of<G>(tx, jj)] = (double)s0;
        b1[ty * N + slot_of<G>(tx, jj)] = (double)s1;
    }
    if (!col_scan_exchange<G, STRIP>(sh, X, t, nbar))
        return false;
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        const int sj = slot_of<G>(tx, jj);
        const float A0 = (float)(sh.cext[sj] + b0[ty * N + sj]);
        const float Tb0 = (float)(sh.cext[2 * N + sj] - (sh.cext[N + sj] + b1[ty * N + sj]));
        float al = 0.f, cl = 0.f;
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            const float gi = (float)(n - (gi0 + i0 + ii));
            al += v[ii][jj];
            cl += v[ii][jj] * gi;
            out[ii][jj] = scale * (gi * (A0 + al) + (Tb0 - cl));
        }
    }
    __syncthreads();
    return true;
}

// One damped wire-solve iteration on a register tile (p, q):
//   p <- (1-w) p + w r sum_j' (min(j,j')+1) a(i,j')     (every wordline solved)
//   q <- (1-w) q + w r sum_i' (n - max(i,i')) b(i',j)   (every bitline solved)
// for the currents of the present (p, q): cur(ii, prow, qrow, a, b) gives
// tile row ii of a and b.  The currents are evaluated twice -- once for the
// column partials, once in the pass that writes the new p and q row by row --
// so no current tile is held: only p, q and the column accumulators are
// live.  FP32 arithmetic, with the bitline sums across thread rows in FP64.
// Returns false when the team aborted.
template <class G, bool STRIP, class CUR>
__device__ __forceinline__ bool wire_iteration(const Shared &sh, const StripCtx &X, CUR cur, double rwire, double w,
                                               int n, int gi0, int i0, int j0, int tx, int ty, int t, int &nbar,
                                               float (&p)[G::TR][G::TC], float (&q)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, TY = G::TY, N = G::N;
    double *b0 = sh.cbuf, *b1 = sh.cbuf + TY * N;
    const float wf = (float)w, w1f = (float)(1.0 - w), rf = (float)rwire;
    // opaque to the optimizer: the index arithmetic below is redone in every
    // iteration of the caller's loop instead of being hoisted out of it and
    // held in registers, which would spill
    asm volatile("" : "+v"(i0), "+v"(j0));
    {
        float s0[TC], s1[TC];
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            s0[jj] = s1[jj] = 0.f;
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            float a[TC], b[TC];
            cur(ii, p[ii], q[ii], a, b);
            const float gi = (float)(n - (gi0 + i0 + ii));
#pragma unroll
            for (int jj = 0; jj < TC; ++jj) {
                s0[jj] += b[jj];
                s1[jj] = fmaf(b[jj], gi, s1[jj]);
            }
        }
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            b0[ty * N + slot_of<G>(tx, jj)] = (double)s0[jj];
            b1[ty * N + slot_of<G>(tx, jj)] = (double)s1[jj];
        }
    }
    const double T1 = col_scan_begin<G, STRIP, true>(sh, X, t, nbar);
    // strips: the wordline half of the iteration runs while the team barrier
    // completes; the bitline inputs b are kept for the second half
    float sb[STRIP ? TR : 1][TC];
    if constexpr (STRIP) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            float a[TC], out[TC];
            cur(ii, p[ii], q[ii], a, sb[ii]);
            row_green_row_f<G>(a, j0, tx, rf, out);
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                p[ii][jj] = fmaf(wf, out[jj], w1f * p[ii][jj]);
        }
    }
    if (!col_scan_end<G, STRIP, true>(sh, X, t, nbar, T1))
        return false;
    // anchors of this thread row in FP64 -- the bitline current entering it
    // (sum over the rows above) and the remaining weighted sum below it --
    // then FP32 within the tile: q = r (gi (A0 + a) + (Tb0 - c)), no
    // difference of large sums in FP32
    float A0[TC], Tb0[TC], al[TC], cl[TC];
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        const int sj = slot_of<G>(tx, jj);
        const double A = sh.cext[sj] + b0[ty * N + sj];
        const double C = sh.cext[N + sj] + b1[ty * N + sj];
        A0[jj] = (float)A;
        Tb0[jj] = (float)(sh.cext[2 * N + sj] - C);
        al[jj] = cl[jj] = 0.f;
    }
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        float a[TC], b[TC], out[TC];
        if constexpr (STRIP) {
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                b[jj] = sb[ii][jj];
        } else {
            cur(ii, p[ii], q[ii], a, b);
        }
        const float gi = (float)(n - (gi0 + i0 + ii));
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            al[jj] += b[jj];
            cl[jj] = fmaf(b[jj], gi, cl[jj]);
            const float o = rf * fmaf(gi, A0[jj] + al[jj], Tb0[jj] - cl[jj]);
            q[ii][jj] = fmaf(wf, o, w1f * q[ii][jj]);
        }
        if constexpr (!STRIP) {
            row_green_row_f<G>(a, j0, tx, rf, out);
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                p[ii][jj] = fmaf(wf, out[jj], w1f * p[ii][jj]);
        }
    }
    __syncthreads();
    return true;
}

// Ideal readout S1, the per-column floor of the relative test and the bitline
// adjoint zq(i,j) = r sum_i' g(i',j) (n - max(i,i')) (this thread's tile, from
// its conductance tile gt) of this workgroup's rows of an n-row crossbar
// (strips: column sums over the team).  Returns false when the team aborted.
template <class G, bool STRIP>
__device__ __forceinline__ bool setup_tile(const Inst &I, const StripCtx &X, const float (&gt)[G::TR][G::TC],
                                           int gi0, int i0, int j0, int t, int ty, double rwire, const Shared &sh,
                                           int &nbar, float (&zq)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, N = G::N, NW = G::NW;
    double si[TC], sf[TC];
    float gs = 0.f; // this tile's conductance, for the wire-solve schedule's spectral estimate
#pragma unroll
    for (int jj = 0; jj < TC; ++jj)
        si[jj] = sf[jj] = 0.0;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        const double vi = ldg(I.vin + (gi0 + i0 + ii));
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const double c = vi - ldg(I.vout + (j0 + jj));
            si[jj] += (double)gt[ii][jj] * c;
            sf[jj] += (double)gt[ii][jj] * fabs(c);
            gs += gt[ii][jj];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        gs += __shfl_xor(gs, off, 64);
    if ((t & 63) == 0)
        sh.wmax[t >> 6] = gs;
    col_reduce2<G>(sh, si, sf, t, ty, j0, sh.S1, sh.Ifloor); // its barriers publish wmax
    double gsum = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k)
        gsum += (double)sh.wmax[k];
    if (STRIP) {
        double *mine = X.setup + ((size_t)X.s * 4) * kRowsLd;
        if (t < N) {
            team_st(mine + t, sh.S1[t]);
            team_st(mine + kRowsLd + t, sh.Ifloor[t]);
        }
        if (t == 0)
            team_st(mine + 2 * kRowsLd, gsum);
        if (!team_barrier(X, sh.dead, t, nbar++))
            return false;
        if (t < N) {
            double e0, e1, a, b;
            team_sums(X.setup, 4 * kRowsLd, X.S, 0, t, e0, e1, a, b);
            sh.S1[t] = a;
            sh.Ifloor[t] = b;
        }
        if (t == 0) {
            gsum = 0.0;
            for (int k = 0; k < X.S; ++k)
                gsum += team_ld(X.setup + ((size_t)k * 4 + 2) * kRowsLd);
        }
    }
    if (t == 0)
        *sh.gsum = gsum; // read after the barriers of the bitline adjoint below
    if (t < N)
        // a column carrying under 1e-9 of its driven current is judged absolutely
        sh.Ifloor[t] = fmax(1e-9 * sh.Ifloor[t], 1e-300);
    return col_green_tile_f<G, STRIP>(sh, X, gt, (float)rwire, I.rows, gi0, i0, j0, ty, t, nbar, zq);
}

// setup_tile on the conductance tile read from gp, the adjoint stored in
// workspace plane 0.
template <class G, bool STRIP>
__device__ __forceinline__ bool setup_columns(const Inst &I, const StripCtx &X, float *wsb, const float *gp, int gi0,
                                              int i0, int j0, int t, int ty, double rwire, const Shared &sh,
                                              int &nbar)
{
    constexpr int TR = G::TR, TC = G::TC, N = G::N;
    float gt[TR][TC], zq[TR][TC];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii)
        ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, gt[ii]);
    if (!setup_tile<G, STRIP>(I, X, gt, gi0, i0, j0, t, ty, rwire, sh, nbar, zq))
        return false;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii)
        strow<TC>(wsb + (size_t)(i0 + ii) * N + j0, zq[ii]);
    return true;
}

// Convergence test on the register tiles.  For every node the FP64 residual
// r = b - A x (b = g c before the refinement, the stored refinement residual
// after it; neighbours from the tile and the LDS halos, which hold every
// edge value after a full sweep); then
//   D_j = sum_i [ g_ij r W_i(j) + zq_ij r_q(i,j) ],  W_i(j) = sum_j' (min(j,j')+1) r_p(i,j')
//   I_j = S1_j - sum_i g_ij (p_ij + q_ij)
// with D_j the adjoint-weighted first-order error of I_j.  want_rho stores the
// scaled residual (planes 1, 2) as the refinement right-hand side instead.
// Strips first refresh the neighbours' rows (publication pub) and complete the
// column sums over the team.  Results: colI[j], colD[j]; return value =
// max_j |D_j|/|I_j| (negative when the team aborted).
//
// test_row: tile row ii (crossbar row gi0 + i0 + ii) -- its conductances
// grow and wordline drops prow, the wordline drops left and right of the tile
// (pl, pr; has_l / has_r false at an open end), the bitline drops of the row
// itself (qcur), of the row above (qup; none at the top: has_up false) and of
// the row below (qdown; none below the last row: has_down false), the stored
// refinement residual rows rhp, rhq (refined only) and the adjoint row zrow
// (not want_rho) -- adds its terms to Dp (adjoint-weighted residual) and Yp
// (sum g (p + q)); want_rho returns the scaled residuals in rpf, rqf instead.
template <class G>
__device__ __forceinline__ void test_row(const Inst &I, int gi0, int i0, int j0, int ii, int tx, double a,
                                         double rwire, double hcoef, bool refined, bool want_rho,
                                         const float (&grow)[G::TC], const float (&prow)[G::TC], float pl_edge,
                                         bool has_l, float pr_edge, bool has_r, const float (&qup)[G::TC], bool has_up,
                                         const float (&qcur)[G::TC], const float (&qdown)[G::TC], bool has_down,
                                         const float (&rhp)[G::TC], const float (&rhq)[G::TC],
                                         const float (&zrow)[G::TC], double (&Dp)[G::TC], double (&Yp)[G::TC],
                                         float (&rpf)[G::TC], float (&rqf)[G::TC])
{
    constexpr int TC = G::TC, TX = G::TX;
    const double vi = ldg(I.vin + (gi0 + i0 + ii));
    // residuals of the row: FP64 arithmetic, FP32 storage
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        const double x = (double)prow[jj], z = (double)qcur[jj], gd = (double)grow[jj];
        const double cell = gd * (x + z);
        double bp, bq;
        if (refined) {
            bp = (double)rhp[jj] / hcoef;
            bq = (double)rhq[jj] / hcoef;
        } else {
            bp = bq = gd * (vi - ldg(I.vout + (j0 + jj)));
        }
        double pl = 0.0;
        if (jj > 0)
            pl = (double)prow[jj - 1];
        else if (has_l)
            pl = (double)pl_edge;
        double wp;
        if (jj < TC - 1)
            wp = a * ((pl - x) + ((double)prow[jj + 1] - x));
        else if (has_r)
            wp = a * ((pl - x) + ((double)pr_edge - x));
        else
            wp = a * (pl - x); // open right end
        const double qd = has_down ? (double)qdown[jj] : 0.0; // sink below the last row
        double wq;
        if (has_up)
            wq = a * (((double)qup[jj] - z) + (qd - z));
        else
            wq = a * (qd - z); // open top
        const double rpv = (bp - cell) + wp, rqv = (bq - cell) + wq;
        if (want_rho) {
            rpf[jj] = (float)(rpv * hcoef);
            rqf[jj] = (float)(rqv * hcoef);
        } else {
            rpf[jj] = (float)rpv;
            Dp[jj] += (double)zrow[jj] * rqv; // bitline part of the adjoint
            Yp[jj] += cell;
        }
    }
    if (want_rho)
        return;
    // wordline part: g r W_i(j), W by a scan over the thread row
    double L0 = 0.0, L1 = 0.0;
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        L0 += (double)rpf[jj];
        L1 += (double)(j0 + jj + 1) * (double)rpf[jj];
    }
    double P0, P1, R, R1;
    row_scan<TX>(L0, tx, P0, R);
    row_scan<TX>(L1, tx, P1, R1);
#pragma unroll
    for (int jj = 0; jj < TC; ++jj) {
        const double jp1 = (double)(j0 + jj + 1);
        Dp[jj] += (double)grow[jj] * rwire * (P1 + jp1 * (R - P0));
        P0 += (double)rpf[jj];
        P1 += jp1 * (double)rpf[jj];
    }
}

// test_reduce: the column sums of every thread's Dp, Yp (strips: over the
// team), colD, colI = S1 - sum g (p + q), and the largest |D_j| / |I_j|,
// uniform over the workgroup (negative when the team aborted).
template <class G, bool STRIP>
__device__ __forceinline__ double test_reduce(const Shared &sh, const StripCtx &X, const double (&Dp)[G::TC],
                                              const double (&Yp)[G::TC], int t, int ty, int j0, int &nbar)
{
    constexpr int N = G::N, NW = G::NW;
    col_reduce2<G>(sh, Dp, Yp, t, ty, j0, sh.colD, sh.colI); // colI holds sum g (p + q) for now
    if (STRIP) {
        double *mine = X.part + (((size_t)(nbar & 1) * X.S + X.s) * 2) * kRowsLd;
        if (t < N) {
            team_st(mine + t, sh.colD[t]);
            team_st(mine + kRowsLd + t, sh.colI[t]);
        }
        if (!team_barrier(X, sh.dead, t, nbar++))
            return -1.0;
        if (t < N) {
            double e0, e1, D, Y;
            team_sums(X.part + ((size_t)((nbar - 1) & 1) * X.S * 2) * kRowsLd, 2 * kRowsLd, X.S, 0, t, e0, e1, D, Y);
            sh.colD[t] = D;
            sh.colI[t] = Y;
        }
    }
    float emax = 0.f;
    if (t < N) {
        const double Ij = sh.S1[t] - sh.colI[t];
        sh.colI[t] = Ij;
        const double e = fabs(sh.colD[t]) / fmax(fabs(Ij), sh.Ifloor[t]);
        emax = (e == e) ? (float)fmin(e, 3.0e38) : INFINITY;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        emax = fmaxf(emax, __shfl_xor(emax, off, 64));
    if ((t & 63) == 0)
        sh.wmax[t >> 6] = emax;
    __syncthreads();
    float em = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w)
        em = fmaxf(em, sh.wmax[w]);
    // identical in every lane: make it a scalar
    em = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, em)));
    __syncthreads(); // wmax / colI reads done before reuse
    return (double)em;
}

// conv_test: the drops from this thread's snapshot rows (planes 3, 4; the
// sweep state is not live across the test, which leaves it the register
// file), the adjoint from plane 0, the refinement residual from planes 1, 2.
template <class G, bool STRIP>
__device__ double conv_test(const Inst &I, const StripCtx &X, float *wsb, const float *gp, int gi0, int i0, int j0,
                            int t, int tx, int ty, double a, double rwire, double hcoef, bool refined, bool want_rho,
                            bool has_l, bool has_r, bool has_u, bool has_d, const Shared &sh, int pub, int &nbar)
{
    const float *snp = wsb + 3 * G::M * G::N, *snq = wsb + 4 * G::M * G::N;
    constexpr int TR = G::TR, TC = G::TC, N = G::N, M = G::M, NT = G::NT, TX = G::TX, QLD = G::QLD;
    const float *halo = sh.halo;
    if (STRIP && !strip_fetch<G>(X, sh.halo, sh.dead, tx, ty, pub))
        return -1.0;
    double Dp[TC], Yp[TC];
#pragma unroll
    for (int jj = 0; jj < TC; ++jj)
        Dp[jj] = Yp[jj] = 0.0;
    float qprev[TC], qcur[TC], qnext[TC];
    ldrow<TC>(snq + (size_t)i0 * N + j0, qcur);
#pragma unroll 1
    for (int ii = 0; ii < TR; ++ii) { // rolled: one tile row at a time keeps the test's registers low
        float grow[TC], rhp[TC], rhq[TC], prow[TC], zrow[TC], rpf[TC], rqf[TC];
        ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, grow);
        ldrow<TC>(snp + (size_t)(i0 + ii) * N + j0, prow);
        if (ii < TR - 1)
            ldrow<TC>(snq + (size_t)(i0 + ii + 1) * N + j0, qnext);
        if (refined) {
            ldrow<TC>(wsb + 1 * M * N + (size_t)(i0 + ii) * N + j0, rhp);
            ldrow<TC>(wsb + 2 * M * N + (size_t)(i0 + ii) * N + j0, rhq);
        }
        if (!want_rho)
            ldrow<TC>(wsb + (size_t)(i0 + ii) * N + j0, zrow);
        // the neighbours beyond the tile's edges from the LDS halo
        const float pl = has_l ? halo[G::PR + ii * NT + t - 1] : 0.f;
        const float pr = has_r ? halo[G::PL + ii * NT + t + 1] : 0.f;
        if (ii == 0 && has_u) {
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                qprev[jj] = halo[G::QB + jj * QLD + t];
        }
        if (ii == TR - 1 && has_d) {
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                qnext[jj] = halo[G::QT + jj * QLD + 2 * TX + t];
        }
        test_row<G>(I, gi0, i0, j0, ii, tx, a, rwire, hcoef, refined, want_rho, grow, prow, pl, has_l, pr, has_r,
                    qprev, ii > 0 || has_u, qcur, qnext, ii < TR - 1 || has_d, rhp, rhq, zrow, Dp, Yp, rpf, rqf);
        if (want_rho) {
            strow<TC>(wsb + 1 * M * N + (size_t)(i0 + ii) * N + j0, rpf);
            strow<TC>(wsb + 2 * M * N + (size_t)(i0 + ii) * N + j0, rqf);
        }
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            qprev[jj] = qcur[jj];
            qcur[jj] = qnext[jj];
        }
    }
    if (want_rho) {
        __syncthreads();
        return 0.0;
    }
    return test_reduce<G, STRIP>(sh, X, Dp, Yp, t, ty, j0, nbar);
}

// Estimated relative error of the zero-drop guess, from the test just run:
// that guess reads out S1_j (the ideal currents), and I_j - D_j (D_j being the
// estimated error of the present readout I_j) is the test's first-order
// estimate of the solution's current, so the zero guess is off by about
// |I_j - D_j - S1_j|.  The largest over the columns, uniform over the
// workgroup (strips: colI, colD and S1 are team totals, so over the team).
template <class G>
__device__ __forceinline__ double zero_guess_est(const Shared &sh, int t)
{
    constexpr int N = G::N, NW = G::NW;
    float e = 0.f;
    if (t < N) {
        const double It = sh.colI[t] - sh.colD[t];
        const double e0 = fabs(It - sh.S1[t]) / fmax(fabs(It), sh.Ifloor[t]);
        e = (e0 == e0) ? (float)fmin(e0, 3.0e38) : INFINITY;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        e = fmaxf(e, __shfl_xor(e, off, 64));
    if ((t & 63) == 0)
        sh.wmax[t >> 6] = e;
    __syncthreads();
    float em = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w)
        em = fmaxf(em, sh.wmax[w]);
    em = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, em)));
    __syncthreads();
    return (double)em;
}

// ---------------------------------------------------------------- phases
// The set-up, the test, the IR2S guess and the refinement guess run outside
// the sweep loop.  Each is compiled as a function of its own so that its
// register needs do not add to the pressure of the sweep loop (the tile state
// p, q, h, rho_p, rho_q in VGPRs); the tile state crosses them through the
// snapshot planes 3, 4 of the workspace.  The LDS pointers are passed one by
// one (a struct of LDS addresses passed by value would become a constant
// with address-space casts, which the backend cannot emit).
// Uniform values reach a device function in VGPRs; moving them to SGPRs
// (readfirstlane) lets the compiler keep addresses as scalar base + 32-bit
// per-lane offset instead of 64-bit per-lane pointers.
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}
// LDS arrays cross a function boundary as 32-bit LDS addresses: a pointer
// rebuilt from one is known to point into LDS, so the callee addresses it
// with ds_read / ds_write and a 32-bit offset instead of 64-bit flat addresses.
template <class T>
__device__ __forceinline__ unsigned lds_off(T *p)
{
    typedef __attribute__((address_space(3))) T lds_t;
    return (unsigned)(size_t)(lds_t *)p;
}
template <class T>
__device__ __forceinline__ T *lds_at(unsigned off)
{
    typedef __attribute__((address_space(3))) T lds_t;
    return (T *)(lds_t *)(size_t)__builtin_amdgcn_readfirstlane(off);
}
__device__ __forceinline__ float uniform_f32(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double uniform_f64(double x)
{
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int uniform_i32(int x) { return (int)__builtin_amdgcn_readfirstlane((unsigned)x); }
__device__ __forceinline__ void uniform_inst(Inst &I)
{
    I.g = uniform_ptr(I.g);
    I.vin = uniform_ptr(I.vin);
    I.vout = uniform_ptr(I.vout);
    I.rwire = uniform_f64(I.rwire);
    I.rows = uniform_i32(I.rows);
}

__device__ __forceinline__ void uniform_strip(StripCtx &X)
{
    X.S = uniform_i32(X.S);
    X.s = uniform_i32(X.s);
    X.rows = uniform_ptr(X.rows);
    X.flag = uniform_ptr(X.flag);
    X.cnt = uniform_ptr(X.cnt);
    X.part = uniform_ptr(X.part);
    X.setup = uniform_ptr(X.setup);
    X.abort = uniform_ptr(X.abort);
}

#define RB_SH_PARAMS                                                                                    \
    unsigned sh_halo, unsigned sh_S1, unsigned sh_Ifloor, unsigned sh_colI, unsigned sh_colD,                \
        unsigned sh_wmax, unsigned sh_dead, unsigned sh_cbuf, unsigned sh_cext, unsigned sh_gsum
#define RB_SH_ARGS(s)                                                                                      \
    lds_off(s.halo), lds_off(s.S1), lds_off(s.Ifloor), lds_off(s.colI), lds_off(s.colD), lds_off(s.wmax),    \
        lds_off(s.dead), lds_off(s.cbuf), lds_off(s.cext), lds_off(s.gsum)
#define RB_SH_LOCAL                                                                                     \
    Shared sh{lds_at<float>(sh_halo),  lds_at<double>(sh_S1),   lds_at<double>(sh_Ifloor),                 \
              lds_at<double>(sh_colI), lds_at<double>(sh_colD), lds_at<float>(sh_wmax),                    \
              lds_at<int>(sh_dead),    lds_at<double>(sh_cbuf), lds_at<double>(sh_cext),                   \
              lds_at<double>(sh_gsum)};                                                                  \
    uniform_inst(I);                                                                                         \
    uniform_strip(X);                                                                                        \
    wsb = uniform_ptr(wsb);                                                                                  \
    gp = uniform_ptr(gp);                                                                                    \
    gi0 = uniform_i32(gi0)
template <class G, bool STRIP>
__device__ __attribute__((noinline)) bool phase_setup(Inst I, StripCtx X, float *wsb, const float *gp, int gi0,
                                                      int i0, int j0, int t, int ty, double rwire, RB_SH_PARAMS,
                                                      int *nbar)
{
    RB_SH_LOCAL;
    return setup_columns<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, ty, rwire, sh, *nbar);
}

template <class G, bool STRIP>
__device__ __attribute__((noinline)) double phase_test(Inst I, StripCtx X, float *wsb, const float *gp, int gi0,
                                                       int i0, int j0, int t, int tx, int ty, double a, double rwire,
                                                       double hcoef, bool refined, bool want_rho, bool has_l,
                                                       bool has_r, bool has_u, bool has_d, RB_SH_PARAMS, int pub,
                                                       int *nbar)
{
    RB_SH_LOCAL;
    return conv_test<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, a, rwire, hcoef, refined, want_rho, has_l,
                               has_r, has_u, has_d, sh, pub, *nbar);
}

template <class G>
__device__ __forceinline__ void tile_load(const float *wsb, int i0, int j0, float (&yp)[G::TR][G::TC],
                                          float (&yq)[G::TR][G::TC])
{
#pragma unroll
    for (int ii = 0; ii < G::TR; ++ii) {
        ldrow<G::TC>(wsb + 3 * G::M * G::N + (size_t)(i0 + ii) * G::N + j0, yp[ii]);
        ldrow<G::TC>(wsb + 4 * G::M * G::N + (size_t)(i0 + ii) * G::N + j0, yq[ii]);
    }
}

template <class G>
__device__ __forceinline__ void tile_store(float *wsb, int i0, int j0, const float (&yp)[G::TR][G::TC],
                                           const float (&yq)[G::TR][G::TC])
{
#pragma unroll
    for (int ii = 0; ii < G::TR; ++ii) {
        strow<G::TC>(wsb + 3 * G::M * G::N + (size_t)(i0 + ii) * G::N + j0, yp[ii]);
        strow<G::TC>(wsb + 4 * G::M * G::N + (size_t)(i0 + ii) * G::N + j0, yq[ii]);
    }
}

// Wire-solve (IR2S) iterations on the drops, in place on the drops held in planes 3, 4:
//   s_ij = g_ij (c_ij - p_ij - q_ij)            (cell currents of the drops)
//   p_ij <- r sum_j' (min(j,j')+1) s_ij'       (wordline: source segment + wires)
//   q_ij <- r sum_i' (n - max(i,i')) s_i'j     (bitline: wires + sink segment)
// i.e. every line solved exactly for the present cell currents, both layers
// from the same currents: the iteration of ir_drop_fastSolve_cpu_singleMat
// (crossbarHspice/ir_drop_fast_solver.h:24-80) and of ir2s
// (pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/split_arr_matmul.cu:336-545)
// in drop variables, uniform wire resistance.  In operator form x <- M^-1 (b -
// K x) with M the wire Laplacians and K = [G G; G G] the cell coupling; the
// eigenvalues of M^-1 K lie in [0, rho], rho ~ 0.8 r gbar n m (0.03 at
// 128x128, r = 3 ohm, gbar ~ 7e-7 S; 0.15-0.8 at 512x512), so the iteration
// contracts by rho, and the damped one, x <- (1-w) x + w M^-1 (b - K x), by
// max(|1-w|, |1-w(1+rho)|) -- rho/(2+rho) at w = 2/(2+rho).  Returns false
// when the team aborted.
// The caller's sweep state is dead across the call (it is reloaded from
// planes 3, 4 afterwards), so the drops and the conductance tile stay in
// VGPRs for all iterations (wire_iteration).  start: the drops begin at zero
// (kStartZero), at the warm start (kStartWarm), or at the warm start plus the
// network's coarse correction combined while loading (kStartWarmCoarse, whole
// crossbars) -- read from HBM here, not from a workspace plane.  The result
// goes to planes 3, 4.  kStartBaseCoarse: the warm start's base field formed
// here -- two undamped wire-solve iterations from zero drops, the base of
// fadesim_nn_features_pooled when the bitlines are grounded -- plus the
// coarse correction (whole crossbars).
constexpr int kStartZero = 0, kStartWarm = 1, kStartWarmCoarse = 2;
// the learned warm start formed in-kernel (guess_solve with a START that
// provides it)
constexpr int kStartNet = 3;
constexpr int kStartBaseCoarse = 4;
// undamped wire-solve iterations of the base field of kStartBaseCoarse
constexpr int kBaseIters = 2;

// The starting drops of the wire-solve iterations (start as above) into the
// register tile p, q.
template <class G, bool STRIP>
__device__ __forceinline__ void load_start(const Inst &I, const Shared &sh, int t, int gi0, int i0, int j0, int start,
                                           float (&p)[G::TR][G::TC], float (&q)[G::TR][G::TC])
{
    constexpr int TC = G::TC, N = G::N;
    if (start == kStartZero) {
#pragma unroll
        for (int ii = 0; ii < G::TR; ++ii)
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                p[ii][jj] = q[ii][jj] = 0.f;
    } else {
        const size_t MNall = (size_t)I.rows * N;
#pragma unroll
        for (int ii = 0; ii < G::TR; ++ii) {
            ldrow<TC>(I.warm + (size_t)(gi0 + i0 + ii) * N + j0, p[ii]);
            ldrow<TC>(I.warm + MNall + (size_t)(gi0 + i0 + ii) * N + j0, q[ii]);
        }
        if (!STRIP && start == kStartWarmCoarse)
            add_coarse<G>(I, sh.halo, t, gi0, i0, j0, p, q);
    }
}

// Damping of wire-solve iteration k of K.  w in (0, 1]: that fixed damping.
// w < 0: the Chebyshev schedule for the eigenvalues of M^-1 K in [0, rho],
// rho = |w| x rho_est,
//   rho_est = r gbar (1 / (4 sin^2(pi / (4 cols + 2))) + 1 / (4 sin^2(pi / (4 rows + 2))))
// -- r gbar times the largest eigenvalues of the wordline and bitline Green's
// matrices min(j,j') + 1 (cols x cols) and n - max(i,i') (rows x rows), exact
// for uniform g, with gbar the crossbar's mean conductance: rho_est = r gsum
// geo, geo (Ctl::ir2s_geo) = the bracket / (rows cols), formed on the host.
// Step k is the Richardson step 1 / mu_k at the Chebyshev node
//   mu_k = 1 + rho c_k,  c_k = (1 + cos(pi (2k + 1) / (2K))) / 2
// of [1, 1 + rho], so after K steps the error is p(M^-1 K) e_0 with
// p(lambda) = T_K((2 lambda - rho) / rho) / T_K(-(2 + rho) / rho): at most
// 1 / T_K((2 + rho) / rho) on the interval, against (rho / (2 + rho))^K for
// the best fixed damping.  Every factor |1 - (1 + lambda) / mu_k| is below
// max(1, rho), so the order of the steps is immaterial at these K.  FP32,
// the cosine by the hardware's v_cos_f32 (argument in turns).
__device__ __forceinline__ double ir2s_step(double w, float rho, int k, int K)
{
    if (w > 0.0)
        return w;
    const float c = 0.5f * (1.0f + __builtin_amdgcn_cosf((float)(2 * k + 1) / (float)(4 * K)));
    return (double)(1.0f / fmaf(rho, c, 1.0f));
}
// geo of ir2s_step for a rows x cols crossbar (host)
inline double ir2s_geo(int rows, int cols)
{
    const double sc = std::sin(M_PI / (4.0 * cols + 2.0)), sr = std::sin(M_PI / (4.0 * rows + 2.0));
    return (0.25 / (sc * sc) + 0.25 / (sr * sr)) / ((double)rows * (double)cols);
}

// `iters` damped wire-solve iterations on the register tile p, q.  HOLD: the
// tile's conductances come from gt (held in registers by the caller), else
// every iteration reads them from gp (the cache) row by row and gt is unused.
template <class G, bool STRIP, bool HOLD>
__device__ __forceinline__ bool ir2s_iterate(const Inst &I, const StripCtx &X, const Shared &sh,
                                             const float (&gt)[G::TR][G::TC], const float *gp, int gi0,
                                             int i0, int j0, int t, int tx, int ty, double rwire, int &nbar, int iters,
                                             double w, double geo, float (&p)[G::TR][G::TC],
                                             float (&q)[G::TR][G::TC])
{
    constexpr int TC = G::TC;
    float vi[G::TR], vo[TC];
#pragma unroll
    for (int ii = 0; ii < G::TR; ++ii)
        vi[ii] = (float)ldg(I.vin + (gi0 + i0 + ii));
#pragma unroll
    for (int jj = 0; jj < TC; ++jj)
        vo[jj] = (float)ldg(I.vout + (j0 + jj));
    // s_ij = g_ij (c_ij - p_ij - q_ij) for both line solves
    auto cur = [&](int ii, const float (&pr)[TC], const float (&qr)[TC], float (&a)[TC], float (&b)[TC]) {
        float gr[TC];
        if constexpr (HOLD) {
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                gr[jj] = gt[ii][jj];
        } else {
            ldrow<TC>(gp + (size_t)(i0 + ii) * G::N + j0, gr);
        }
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            a[jj] = b[jj] = gr[jj] * (((vi[ii] - vo[jj]) - pr[jj]) - qr[jj]);
    };
    iters = uniform_i32(iters); // a loop with barriers: its trip count a scalar
    const float rho = uniform_f32(w < 0.0 ? (float)(-w * rwire * uniform_f64(*sh.gsum) * geo) : 0.f);
#pragma unroll 1
    for (int k = 0; k < iters; ++k) {
        // the cell voltages vi - vo are recomputed in every iteration rather
        // than hoisted out of the loop as a register tile (which would spill)
#pragma unroll
        for (int ii = 0; ii < G::TR; ++ii)
            asm volatile("" : "+v"(vi[ii]));
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            asm volatile("" : "+v"(vo[jj]));
        const double wk = uniform_f64(ir2s_step(w, rho, k, iters));
        if (!wire_iteration<G, STRIP>(sh, X, cur, rwire, wk, I.rows, gi0, i0, j0, tx, ty, t, nbar, p, q))
            return false;
    }
    return true;
}

// The starting drops of mode `start` into the register tile p, q: load_start,
// or for kStartBaseCoarse the base field's wire-solve iterations (HOLD as
// for ir2s_iterate) and the coarse correction.  Returns false when the team
// aborted.
template <class G, bool STRIP, bool HOLD>
__device__ __forceinline__ bool start_tile(const Inst &I, const StripCtx &X, const Shared &sh,
                                           const float (&gt)[G::TR][G::TC], const float *gp, int gi0, int i0, int j0,
                                           int t, int tx, int ty, double rwire, int &nbar, int start,
                                           float (&p)[G::TR][G::TC], float (&q)[G::TR][G::TC])
{
    if (start != kStartBaseCoarse) {
        load_start<G, STRIP>(I, sh, t, gi0, i0, j0, start, p, q);
        return true;
    }
#pragma unroll
    for (int ii = 0; ii < G::TR; ++ii)
#pragma unroll
        for (int jj = 0; jj < G::TC; ++jj)
            p[ii][jj] = q[ii][jj] = 0.f;
    if (!ir2s_iterate<G, STRIP, HOLD>(I, X, sh, gt, gp, gi0, i0, j0, t, tx, ty, rwire, nbar, kBaseIters, 1.0, 0.0, p,
                                      q))
        return false;
    if (!STRIP)
        add_coarse<G>(I, sh.halo, t, gi0, i0, j0, p, q);
    return true;
}

// Warm start combined at load (Inst::coarse): planes 3, 4 = base + scale x
// bilinear(coarse) over this thread's tile, the base read from I.warm or, when
// there is none, formed here (kStartBaseCoarse), out of line so that its
// registers do not weigh on the sweep loop.
template <class G, bool STRIP>
__device__ __attribute__((noinline)) bool phase_coarse(Inst I, StripCtx X, float *wsb, const float *gp, int gi0,
                                                       int i0, int j0, int t, int tx, int ty, double rwire,
                                                       RB_SH_PARAMS, int *nbar)
{
    RB_SH_LOCAL;
    constexpr int TR = G::TR, TC = G::TC;
    float yp[TR][TC], yq[TR][TC], gt[TR][TC];
    if (!start_tile<G, STRIP, false>(I, X, sh, gt, gp, gi0, i0, j0, t, tx, ty, rwire, *nbar,
                                     I.warm ? kStartWarmCoarse : kStartBaseCoarse, yp, yq))
        return false;
    tile_store<G>(wsb, i0, j0, yp, yq);
    return true;
}

template <class G, bool STRIP>
__device__ __attribute__((noinline)) bool phase_ir2s(Inst I, StripCtx X, float *wsb, const float *gp, int gi0,
                                                     int i0, int j0, int t, int tx, int ty, double rwire, RB_SH_PARAMS,
                                                     int *nbar, int iters, double w, double geo, int start)
{
    RB_SH_LOCAL;
    constexpr int TC = G::TC;
    float p[G::TR][TC], q[G::TR][TC];
    // whole crossbars keep the tile's conductances in registers for all
    // iterations (an iteration reads no memory but LDS); strips, whose
    // iterations also hold the bitline inputs across the team exchange,
    // reread them from the cache in each current evaluation instead, which
    // leaves fewer spills (measured: the uneven workload 8 % faster, the same
    // results bit for bit)
    constexpr bool HOLD = !STRIP;
    float gt[G::TR][TC];
    if constexpr (HOLD) {
#pragma unroll
        for (int ii = 0; ii < G::TR; ++ii)
            ldrow<TC>(gp + (size_t)(i0 + ii) * G::N + j0, gt[ii]);
    }
    if (!start_tile<G, STRIP, HOLD>(I, X, sh, gt, gp, gi0, i0, j0, t, tx, ty, rwire, *nbar, uniform_i32(start), p, q))
        return false;
    if (!ir2s_iterate<G, STRIP, HOLD>(I, X, sh, gt, gp, gi0, i0, j0, t, tx, ty, rwire, *nbar, iters, w, geo, p, q))
        return false;
    tile_store<G>(wsb, i0, j0, p, q);
    return true;
}

// Initial guess of the refinement correction y (A y = residual), into planes 3, 4:
// `iters` wire-solve iterations from zero, damped by w as in phase_ir2s,
//   y_{k+1} = (1-w) y_k + w (aL)^-1 (res - G (y_k,p + y_k,q))
// with (aL)^-1 the wordline (r (min(j,j')+1)) and bitline (r (n - max(i,i')))
// line inverses; the cell coupling G is ~1e-6 of the wire conductance, so
// two iterations already put y within a few 1e-3 of the correction and the
// correction converges in a few sweeps instead of restarting from zero.
// Reads the scaled residual (planes 1, 2); the correction stays in VGPRs.
template <class G, bool STRIP>
__device__ __attribute__((noinline)) bool phase_green(Inst I, StripCtx X, float *wsb, const float *gp, int gi0,
                                                      int i0, int j0, int t, int tx, int ty, double rwire,
                                                      double hcoef, RB_SH_PARAMS, int *nbar, int iters, double w,
                                                      double geo)
{
    RB_SH_LOCAL;
    constexpr int TR = G::TR, TC = G::TC, M = G::M, N = G::N;
    const float *RP = wsb + 1 * M * N, *RQ = wsb + 2 * M * N;
    const float ih = (float)(1.0 / hcoef);
    float p[TR][TC], q[TR][TC];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii)
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            p[ii][jj] = q[ii][jj] = 0.f;
    (void)I;
    // y starts at zero, so the first iteration's coupling term is exactly 0
    auto cur = [&](int ii, const float (&pr)[TC], const float (&qr)[TC], float (&a)[TC], float (&b)[TC]) {
        const size_t o = (size_t)(i0 + ii) * N + j0;
        float gr[TC], rp[TC], rq[TC];
        ldrow<TC>(RP + o, rp);
        ldrow<TC>(RQ + o, rq);
        ldrow<TC>(gp + o, gr);
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const float cpl = gr[jj] * (pr[jj] + qr[jj]);
            a[jj] = fmaf(rp[jj], ih, -cpl);
            b[jj] = fmaf(rq[jj], ih, -cpl);
        }
    };
    iters = uniform_i32(iters); // a loop with barriers: its trip count a scalar
    const float rho = uniform_f32(w < 0.0 ? (float)(-w * rwire * uniform_f64(*sh.gsum) * geo) : 0.f);
#pragma unroll 1
    for (int k = 0; k < iters; ++k)
        if (!wire_iteration<G, STRIP>(sh, X, cur, rwire, uniform_f64(ir2s_step(w, rho, k, iters)), I.rows, gi0, i0,
                                      j0, tx, ty, t, *nbar, p, q))
            return false;
    tile_store<G>(wsb, i0, j0, p, q);
    return true;
}

template <class G, bool STRIP>
__device__ __forceinline__ bool solve(const Inst &I, const StripCtx &X, const Ctl &C, float *wsb,
                                      const Shared &sh)
{
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, TY = G::TY;
    constexpr int M = G::M, N = G::N;
    float *halo = sh.halo;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int tx = t % TX, ty = t / TX;
    const int strip = STRIP ? X.s : 0;
    const int gi0 = strip * M; // first global row of this workgroup
    const float *gp = I.g + (size_t)gi0 * N;
    const int i0 = TR * ty, j0 = TC * tx;
    const bool top_ext = STRIP && X.s > 0, bot_ext = STRIP && X.s < X.S - 1;
    const bool has_l = tx > 0, has_r = tx < TX - 1;
    const bool has_u = ty > 0 || top_ext, has_d = ty < TY - 1 || bot_ext;
    const size_t MNall = (size_t)I.rows * N;

    const float wf = C.omega;
    const float cA = 1.0f - wf;
    const float cB = 0.5f * wf;
    const float c2 = 2.0f / wf;
    const double rwire = I.rwire;
    const double a = 1.0 / rwire;
    const double hcoef = 0.5 * (double)wf * rwire;
    const bool fixed = (C.flags & FADESIM_F_FIXED_SWEEPS) != 0;
    const bool dwr = (C.flags & FADESIM_F_DWR) != 0;
    int pub = 0, nbar = 0;
    if (t == 0)
        *sh.dead = 0;
    __syncthreads();

    if (!phase_setup<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, ty, rwire, RB_SH_ARGS(sh), &nbar))
        return false;

    float yp[TR][TC], yq[TR][TC];
    const bool coarse = !STRIP && I.coarse;
    // snap_valid: planes 3, 4 hold the present drops (no snapshot store
    // needed before a test)
    bool snap_valid = false;
    if (C.ir2s_iters > 0) {
        // the IR2S phase reads the initial guess itself (the coarse
        // correction combined while loading)
        const int start = coarse ? (I.warm ? kStartWarmCoarse : kStartBaseCoarse) : I.warm ? kStartWarm : kStartZero;
        if (!phase_ir2s<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, rwire, RB_SH_ARGS(sh), &nbar, C.ir2s_iters,
                                  C.ir2s_omega, C.ir2s_geo, start))
            return false;
        tile_load<G>(wsb, i0, j0, yp, yq);
        snap_valid = true;
    } else if (coarse) {
        // the warm start combined while loading, into planes 3, 4
        if (!phase_coarse<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, rwire, RB_SH_ARGS(sh), &nbar))
            return false;
        tile_load<G>(wsb, i0, j0, yp, yq);
        snap_valid = true;
    } else {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            if (I.warm) {
                ldrow<TC>(I.warm + (size_t)(gi0 + i0 + ii) * N + j0, yp[ii]);
                ldrow<TC>(I.warm + MNall + (size_t)(gi0 + i0 + ii) * N + j0, yq[ii]);
            } else {
#pragma unroll
                for (int jj = 0; jj < TC; ++jj)
                    yp[ii][jj] = yq[ii][jj] = 0.f;
            }
        }
    }
    publish_all<G>(yp, yq, halo, t);
    if (STRIP)
        strip_publish<G>(X, yq, ty, j0, lane, pub++);
    __syncthreads();

    // check scheduling only: single-precision hardware log2
    const float lrho = __log2f(fmaxf(wf - 1.0f, 1e-3f));
    bool refined = false;
    bool refine_pending = !(C.flags & FADESIM_F_NO_REFINE) && !fixed;
    int sweep = 0;
    int next_check = fixed ? C.max_sweeps : (C.max_sweeps < C.check_first ? C.max_sweeps : C.check_first);
    int st = FADESIM_ST_MAXIT;
    double est = INFINITY;
    const double stop_tol = dwr ? C.tol_eff * C.dwr_scale : C.tol_eff;
    // a given initial guess is judged at the first test against the zero guess
    bool judge_guess = (I.warm != nullptr || coarse) && !fixed;
    bool rejected = false;

    for (;;) {
        if (sweep < next_check) {
            // the sweep coefficients live only here: loaded for sweeps, dead
            // across the test
            float hn[TR][TC], rp[TR][TC], rq[TR][TC];
            load_coeffs<G>(I, wsb, gp, i0, j0, gi0, hcoef, refined, hn, rp, rq);
            snap_valid = false;
            while (sweep < next_check) {
                if (STRIP && !strip_fetch<G>(X, halo, sh.dead, tx, ty, pub - 1))
                    return false;
                half_sweep<0, G>(yp, yq, hn, rp, rq, halo, t, has_l, has_r, has_u, has_d, cA, cB, c2);
                if (STRIP) {
                    strip_publish<G>(X, yq, ty, j0, lane, pub++);
                    if (!strip_fetch<G>(X, halo, sh.dead, tx, ty, pub - 1))
                        return false;
                }
                half_sweep<1, G>(yp, yq, hn, rp, rq, halo, t, has_l, has_r, has_u, has_d, cA, cB, c2);
                if (STRIP)
                    strip_publish<G>(X, yq, ty, j0, lane, pub++);
                ++sweep;
            }
        }
        if (!snap_valid) {
#pragma unroll
            for (int ii = 0; ii < TR; ++ii) { // snapshot: the sweep state is not live across the test
                strow<TC>(wsb + 3 * M * N + (size_t)(i0 + ii) * N + j0, yp[ii]);
                strow<TC>(wsb + 4 * M * N + (size_t)(i0 + ii) * N + j0, yq[ii]);
            }
            snap_valid = true;
        }
        // the same in every lane, but returned in VGPRs: a scalar keeps the
        // decisions below uniform branches
        est = uniform_f64(phase_test<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, a, rwire, hcoef, refined, false,
                                               has_l, has_r, has_u, has_d, RB_SH_ARGS(sh), pub - 1, &nbar));
        if (est < 0.0)
            return false;
        if (judge_guess) {
            judge_guess = false;
            // a guess estimated worse than the zero guess (or not finite) is
            // dropped: the solve restarts from zero drops (and the IR2S
            // iterations on them, when asked for) and tests that at once
            if (!(est <= zero_guess_est<G>(sh, t))) {
                rejected = true;
                if (C.ir2s_iters > 0) {
                    if (!phase_ir2s<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, rwire, RB_SH_ARGS(sh), &nbar,
                                              C.ir2s_iters, C.ir2s_omega, C.ir2s_geo, kStartZero))
                        return false;
                    tile_load<G>(wsb, i0, j0, yp, yq);
                } else {
#pragma unroll
                    for (int ii = 0; ii < TR; ++ii)
#pragma unroll
                        for (int jj = 0; jj < TC; ++jj)
                            yp[ii][jj] = yq[ii][jj] = 0.f;
                    tile_store<G>(wsb, i0, j0, yp, yq);
                }
                publish_all<G>(yp, yq, halo, t);
                if (STRIP)
                    strip_publish<G>(X, yq, ty, j0, lane, pub++);
                __syncthreads();
                snap_valid = true;
                next_check = sweep;
                continue;
            }
        }
        bool reload = true, done = false, just_refined = false;
        if (fixed) {
            st = FADESIM_ST_CONVERGED;
            done = true;
        } else if (!(est < INFINITY)) {
            st = FADESIM_ST_NONFINITE;
            done = true;
        } else if (est <= stop_tol) {
            // certified with or without the refinement step: the estimate is
            // of the FP64 readout of the present field either way
            st = FADESIM_ST_CONVERGED;
            done = true;
        } else if (refine_pending && est <= C.switch_tol) {
            // ---- one iterative-refinement step: rhs <- exact FP64 residual ----
            phase_test<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, a, rwire, hcoef, false, true, has_l, has_r,
                                 has_u, has_d, RB_SH_ARGS(sh), pub - 1, &nbar);
            if (I.field) {
#pragma unroll
                for (int ii = 0; ii < TR; ++ii) {
                    float v[TC];
                    ldrow<TC>(wsb + 3 * M * N + (size_t)(i0 + ii) * N + j0, v);
                    strow<TC>(I.field + (size_t)(gi0 + i0 + ii) * N + j0, v);
                    ldrow<TC>(wsb + 4 * M * N + (size_t)(i0 + ii) * N + j0, v);
                    strow<TC>(I.field + MNall + (size_t)(gi0 + i0 + ii) * N + j0, v);
                }
            }
            reload = false;
            if (C.flags & FADESIM_F_ZERO_CORRECTION) {
#pragma unroll
                for (int ii = 0; ii < TR; ++ii)
#pragma unroll
                    for (int jj = 0; jj < TC; ++jj)
                        yp[ii][jj] = yq[ii][jj] = 0.f;
                snap_valid = false;
            } else {
                if (!phase_green<G, STRIP>(I, X, wsb, gp, gi0, i0, j0, t, tx, ty, rwire, hcoef, RB_SH_ARGS(sh), &nbar,
                                           C.corr_iters, C.ir2s_omega, C.ir2s_geo))
                    return false;
                tile_load<G>(wsb, i0, j0, yp, yq);
            }
            if (t < N)
                sh.S1[t] = sh.colI[t];
            publish_all<G>(yp, yq, halo, t);
            if (STRIP)
                strip_publish<G>(X, yq, ty, j0, lane, pub++);
            __syncthreads();
            refine_pending = false;
            refined = true;
            just_refined = true;
        }
        if (!done && sweep >= C.max_sweeps) {
            st = FADESIM_ST_MAXIT;
            done = true;
        }
        if (reload && (!done || I.field)) {
#pragma unroll
            for (int ii = 0; ii < TR; ++ii) {
                ldrow<TC>(wsb + 3 * M * N + (size_t)(i0 + ii) * N + j0, yp[ii]);
                ldrow<TC>(wsb + 4 * M * N + (size_t)(i0 + ii) * N + j0, yq[ii]);
            }
        }
        if (done)
            break;
        const double tgt = refine_pending ? C.switch_tol : stop_tol;
        const float k = (est > tgt) ? 0.8f * __log2f((float)(tgt / est)) / lrho : (float)C.check_min;
        int kk = (int)fminf(fmaxf(k, (float)C.check_min), (float)C.check_max);
        if (just_refined && (C.flags & FADESIM_F_TEST_FIRST))
            kk = 0; // test the correction's initial guess before sweeping
        next_check = sweep + kk;
        if (next_check > C.max_sweeps)
            next_check = C.max_sweeps;
    }

    // ---- outputs ---------------------------------------------------------------
    if (strip == 0) {
        if (t < N) {
            double Iv = sh.colI[t];
            if (dwr && st == FADESIM_ST_CONVERGED)
                Iv -= sh.colD[t];
            I.iout[col_of_slot<G>(t)] = Iv;
        }
        if (t == 0) {
            if (I.sweeps)
                *I.sweeps = sweep;
            if (I.status)
                *I.status = st;
            if (I.est)
                *I.est = fixed ? NAN : est;
            if (I.rejected)
                *I.rejected = rejected ? 1 : 0;
        }
    }
    if (I.field) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            float *P = I.field + (size_t)(gi0 + i0 + ii) * N + j0;
            float *Q = P + MNall;
            float pv[TC], qv[TC];
            if (refined) {
                ldrow<TC>(P, pv);
                ldrow<TC>(Q, qv);
            }
#pragma unroll
            for (int jj = 0; jj < TC; ++jj) {
                pv[jj] = refined ? pv[jj] + yp[ii][jj] : yp[ii][jj];
                qv[jj] = refined ? qv[jj] + yq[ii][jj] : yq[ii][jj];
            }
            strow<TC>(P, pv);
            strow<TC>(Q, qv);
        }
    }
    __syncthreads(); // colI / halo reads done before the next instance reuses them
    return true;
}

// The part of solve() before its first sweep, for whole crossbars whose
// guess is tested first (FADESIM_F_TEST_FIRST, not FADESIM_F_FIXED_SWEEPS):
// the set-up, the initial guess (warm start, coarse combine, IR2S
// iterations), its test, and -- when a given guess is judged worse than the
// zero guess -- the zero guess with its IR2S iterations and their test.  The
// same arithmetic in the same order as solve(), so whatever this writes
// equals what solve() writes bit for bit.  The drops go from the IR2S
// iterations' registers into fld (LDS, [2][M][N], overlaying the halo and
// the column buffer, both idle during the test), where the test reads them
// with their neighbours; only the adjoint crosses to the test through the
// workspace (plane 0).  HOLD: the iterations keep the conductance tile in
// registers, else they reread it from the cache row by row.  Returns true
// when a test ended the solve before any sweep (certified, or not finite),
// with every output written as solve() writes it; false when the solve needs
// sweeps (no output written but the field, which solve() then rewrites as it
// runs the crossbar from the start).  start: the initial guess of the
// first test (kStartZero, kStartWarm, kStartWarmCoarse as I holds them, or
// kStartNet); a guess other than kStartZero is judged against the zero
// guess.  START: start_fn(start, t, i0, j0, p, q) fills the register tile
// (thread t, first row i0, first column j0) with the drops of a start mode
// before the IR2S iterations (load_start, or a caller's that adds
// kStartNet).
template <class G, bool HOLD, class START>
__device__ __forceinline__ bool guess_solve(const Inst &I, const Ctl &C, float *wsb, const Shared &sh, float *fld,
                                            int first_start, START start_fn)
{
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, TY = G::TY;
    constexpr int M = G::M, N = G::N;
    const StripCtx X{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // opaque to the optimizer: addresses derived from the thread index are
    // computed for each crossbar, not hoisted out of the caller's loop over
    // crossbars and held (spilled) for the kernel's lifetime
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int tx = t % TX, ty = t / TX;
    const float *gp = I.g;
    const int i0 = TR * ty, j0 = TC * tx;
    const bool has_l = tx > 0, has_r = tx < TX - 1;
    const double rwire = I.rwire;
    const double a = 1.0 / rwire;
    const double hcoef = 0.5 * (double)C.omega * rwire;
    const bool dwr = (C.flags & FADESIM_F_DWR) != 0;
    float *fp = fld, *fq = fld + M * N;
    int nbar = 0;
    if (t == 0)
        *sh.dead = 0;
    __syncthreads();

    setup_columns<G, false>(I, X, wsb, gp, 0, i0, j0, t, ty, rwire, sh, nbar);

    const int iters = uniform_i32(C.ir2s_iters);
    // the drops of the guess into fld (and the field output)
    auto guess = [&](int start) {
        float p[TR][TC], q[TR][TC];
        start_fn(start, t, i0, j0, p, q);
        if (iters > 0) {
            float gt[TR][TC];
            if constexpr (HOLD) {
#pragma unroll
                for (int ii = 0; ii < TR; ++ii)
                    ldrow<TC>(gp + (size_t)(i0 + ii) * N + j0, gt[ii]);
            }
            ir2s_iterate<G, false, HOLD>(I, X, sh, gt, gp, 0, i0, j0, t, tx, ty, rwire, nbar, iters, C.ir2s_omega,
                                         C.ir2s_geo, p, q);
        }
        __syncthreads(); // the column buffer and the halo stage are idle
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            strow_lds<TC>(fp + (i0 + ii) * N + j0, p[ii]);
            strow_lds<TC>(fq + (i0 + ii) * N + j0, q[ii]);
        }
        if (I.field) {
#pragma unroll
            for (int ii = 0; ii < TR; ++ii) {
                strow<TC>(I.field + (size_t)(i0 + ii) * N + j0, p[ii]);
                strow<TC>(I.field + (size_t)M * N + (size_t)(i0 + ii) * N + j0, q[ii]);
            }
        }
        __syncthreads();
    };
    // the test, one tile row at a time from fld; its value is the same in
    // every lane, made a scalar so that every decision below is a uniform
    // branch
    auto test = [&]() {
        double Dp[TC], Yp[TC];
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            Dp[jj] = Yp[jj] = 0.0;
#pragma unroll 1
        for (int ii = 0; ii < TR; ++ii) {
            const int i = i0 + ii;
            float grow[TC], prow[TC], qup[TC], qcur[TC], qdown[TC], zrow[TC], rpf[TC], rqf[TC];
            ldrow<TC>(gp + (size_t)i * N + j0, grow);
            ldrow<TC>(wsb + (size_t)i * N + j0, zrow);
            ldrow_lds<TC>(fp + i * N + j0, prow);
            ldrow_lds<TC>(fq + i * N + j0, qcur);
            const bool up = i > 0, down = i < M - 1;
            if (up)
                ldrow_lds<TC>(fq + (i - 1) * N + j0, qup);
            if (down)
                ldrow_lds<TC>(fq + (i + 1) * N + j0, qdown);
            const float pl = has_l ? fp[i * N + j0 - 1] : 0.f;
            const float pr = has_r ? fp[i * N + j0 + TC] : 0.f;
            test_row<G>(I, 0, i0, j0, ii, tx, a, rwire, hcoef, false, false, grow, prow, pl, has_l, pr, has_r, qup,
                        up, qcur, qdown, down, zrow, zrow, zrow, Dp, Yp, rpf, rqf);
        }
        __syncthreads(); // fld reads done: the column sums reuse its memory
        return uniform_f64(test_reduce<G, false>(sh, X, Dp, Yp, t, ty, j0, nbar));
    };
    guess(first_start);
    double est = test();
    bool rejected = false;
    if (first_start != kStartZero && !(est <= zero_guess_est<G>(sh, t))) {
        rejected = true;
        guess(kStartZero);
        est = test();
    }
    const bool nonfinite = !(est < INFINITY);
    if (!nonfinite && !(est <= (dwr ? C.tol_eff * C.dwr_scale : C.tol_eff)))
        return false;
    const int st = nonfinite ? FADESIM_ST_NONFINITE : FADESIM_ST_CONVERGED;
    if (t < N) {
        double Iv = sh.colI[t];
        if (dwr && st == FADESIM_ST_CONVERGED)
            Iv -= sh.colD[t];
        I.iout[col_of_slot<G>(t)] = Iv;
    }
    if (t == 0) {
        if (I.sweeps)
            *I.sweeps = 0;
        if (I.status)
            *I.status = st;
        if (I.est)
            *I.est = est;
        if (I.rejected)
            *I.rejected = rejected ? 1 : 0;
    }
    __syncthreads(); // colI reads done before the next instance reuses it
    return true;
}

} // namespace rb
} // namespace fadesim
