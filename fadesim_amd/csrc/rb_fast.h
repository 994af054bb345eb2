// Guess-first fast path of the batched solve (rbsor.hip).
//
// When a solve tests its initial guess before sweeping (FADESIM_F_TEST_FIRST
// with IR2S iterations or a warm start), most crossbars are certified at that
// first test and never sweep.  For them the sweeping kernel's structure --
// out-of-line phases that hand the drops over through workspace planes, so
// that the sweep loop keeps the register file -- costs more than the work:
// the set-up, the guess and the test re-read and re-write the tile from
// memory between phases.  fast_solve does the same arithmetic with the tile
// held in registers from the conductance load to the readout: the ideal
// readout and the bitline adjoint of the set-up, the warm start formed at
// load, the IR2S iterations and the FP64 test.  An instance the test
// certifies gets its outputs here (0 sweeps); any other goes to a list that
// the sweeping kernel then solves from the start, on its own path.
//
// Every expression is the sweeping kernel's (setup_columns, add_coarse,
// wire_iteration, conv_test, zero_guess_est), in the same order and
// precision, so a certified instance's outputs equal the sweeping kernel's
// bit for bit.
#pragma once
#include "rb_core.h"

namespace fadesim {
namespace rb {

// The FP64 convergence test of conv_test (no refinement, one workgroup per
// crossbar) on the register tiles of the drops p, q; conductances and the
// bitline adjoint (zqp, [M][N]) are read row by row.  Leaves colI, colD in
// LDS; returns max_j |D_j| / |I_j|.
template <class G>
__device__ __forceinline__ double reg_test(const Inst &I, const float (&p)[G::TR][G::TC],
                                           const float (&q)[G::TR][G::TC], const float *zqp, int i0, int j0, int t,
                                           int tx, int ty, double a, double rwire, bool has_l, bool has_r, bool has_u,
                                           bool has_d, const Shared &sh)
{
    constexpr int TR = G::TR, TC = G::TC, N = G::N, NT = G::NT, NW = G::NW, TX = G::TX, QLD = G::QLD;
    const float *halo = sh.halo;
    double Dp[TC], Yp[TC];
#pragma unroll
    for (int jj = 0; jj < TC; ++jj)
        Dp[jj] = Yp[jj] = 0.0;
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        const double vi = I.vin[i0 + ii];
        float gt[TC], zq[TC], rpf[TC];
        ldrow<TC>(I.g + (size_t)(i0 + ii) * N + j0, gt);
        ldrow<TC>(zqp + (size_t)(i0 + ii) * N + j0, zq);
#pragma unroll
        for (int jj = 0; jj < TC; ++jj) {
            const double x = (double)p[ii][jj], z = (double)q[ii][jj], gd = (double)gt[jj];
            const double cell = gd * (x + z);
            const double bp = gd * (vi - I.vout[j0 + jj]), bq = bp;
            double pl = 0.0;
            if (jj > 0)
                pl = (double)p[ii][jj - 1];
            else if (has_l)
                pl = (double)halo[G::PR + ii * NT + t - 1];
            double wp;
            if (jj < TC - 1)
                wp = a * ((pl - x) + ((double)p[ii][jj + 1] - x));
            else if (has_r)
                wp = a * ((pl - x) + ((double)halo[G::PL + ii * NT + t + 1] - x));
            else
                wp = a * (pl - x); // open right end
            double qd = 0.0;       // sink below the last row
            if (ii < TR - 1)
                qd = (double)q[ii + 1][jj];
            else if (has_d)
                qd = (double)halo[G::QT + jj * QLD + 2 * TX + t];
            double wq;
            if (ii > 0)
                wq = a * (((double)q[ii - 1][jj] - z) + (qd - z));
            else if (has_u)
                wq = a * (((double)halo[G::QB + jj * QLD + t] - z) + (qd - z));
            else
                wq = a * (qd - z); // open top
            const double rpv = (bp - cell) + wp, rqv = (bq - cell) + wq;
            rpf[jj] = (float)rpv;
            Dp[jj] += (double)zq[jj] * rqv; // bitline part of the adjoint
            Yp[jj] += cell;
        }
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
            Dp[jj] += (double)gt[jj] * rwire * (P1 + jp1 * (R - P0));
            P0 += (double)rpf[jj];
            P1 += jp1 * (double)rpf[jj];
        }
    }
    col_reduce2<G>(sh, Dp, Yp, t, ty, j0, sh.colD, sh.colI); // colI holds sum g (p + q) for now
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
    em = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, em)));
    __syncthreads();
    return (double)em;
}

// The initial guess (warm start formed at load, or zero drops) and the IR2S
// iterations on it, in registers.
template <class G>
__device__ __forceinline__ void fast_guess(const Inst &I, const Ctl &C, const Shared &sh, bool use_warm,
                                           double rwire, int i0, int j0, int t, int tx, int ty,
                                           float (&p)[G::TR][G::TC], float (&q)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, N = G::N, M = G::M;
    if (use_warm) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            ldrow<TC>(I.warm + (size_t)(i0 + ii) * N + j0, p[ii]);
            ldrow<TC>(I.warm + (size_t)M * N + (size_t)(i0 + ii) * N + j0, q[ii]);
        }
        if (I.coarse)
            add_coarse<G>(I, sh.halo, t, 0, i0, j0, p, q);
    } else {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii)
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                p[ii][jj] = q[ii][jj] = 0.f;
    }
    if (C.ir2s_iters == 0)
        return;
    // the tile's conductances and line voltages stay in registers for the
    // iterations (as in phase_ir2s)
    float gt[TR][TC], vi[TR], vo[TC];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        ldrow<TC>(I.g + (size_t)(i0 + ii) * N + j0, gt[ii]);
        vi[ii] = (float)I.vin[i0 + ii];
    }
#pragma unroll
    for (int jj = 0; jj < TC; ++jj)
        vo[jj] = (float)I.vout[j0 + jj];
    auto cur = [&](int ii, const float (&pr)[TC], const float (&qr)[TC], float (&a)[TC], float (&b)[TC]) {
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            a[jj] = b[jj] = gt[ii][jj] * (((vi[ii] - vo[jj]) - pr[jj]) - qr[jj]);
    };
    const StripCtx none{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nbar = 0;
#pragma unroll 1
    for (int k = 0; k < C.ir2s_iters; ++k) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii)
            asm volatile("" : "+v"(vi[ii]));
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            asm volatile("" : "+v"(vo[jj]));
        wire_iteration<G, false>(sh, none, cur, rwire, C.ir2s_omega, M, 0, i0, j0, tx, ty, t, nbar, p, q);
    }
}

// One crossbar on the fast path.  Returns true when it was certified (outputs
// written), false when it must go to the sweeping kernel.
// One crossbar on the fast path; zqp: this workgroup's [M][N] plane for the
// bitline adjoint.  Returns true when it was certified (outputs written),
// false when it must go to the sweeping kernel.
template <class G>
__device__ bool fast_solve(const Inst &I, const Ctl &C, const Shared &sh, float *zqp)
{
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, TY = G::TY, N = G::N, M = G::M;
    const int t = threadIdx.x, tx = t % TX, ty = t / TX, i0 = ty * TR, j0 = tx * TC;
    const bool has_l = tx > 0, has_r = tx < TX - 1, has_u = ty > 0, has_d = ty < TY - 1;
    const double rwire = I.rwire, a = 1.0 / rwire;
    // ---- set-up (setup_columns): ideal readout S1, floor, bitline adjoint zq
    {
        float gt[TR][TC];
#pragma unroll
        for (int ii = 0; ii < TR; ++ii)
            ldrow<TC>(I.g + (size_t)(i0 + ii) * N + j0, gt[ii]);
        double si[TC], sf[TC];
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            si[jj] = sf[jj] = 0.0;
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            const double vi = I.vin[i0 + ii];
#pragma unroll
            for (int jj = 0; jj < TC; ++jj) {
                const double c = vi - I.vout[j0 + jj];
                si[jj] += (double)gt[ii][jj] * c;
                sf[jj] += (double)gt[ii][jj] * fabs(c);
            }
        }
        col_reduce2<G>(sh, si, sf, t, ty, j0, sh.S1, sh.Ifloor);
        if (t < N)
            sh.Ifloor[t] = fmax(1e-9 * sh.Ifloor[t], 1e-300);
        float zq[TR][TC];
        const StripCtx none{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        int nbar = 0;
        col_green_tile_f<G, false>(sh, none, gt, (float)rwire, M, 0, i0, j0, ty, t, nbar, zq);
#pragma unroll
        for (int ii = 0; ii < TR; ++ii)
            strow<TC>(zqp + (size_t)(i0 + ii) * N + j0, zq[ii]);
    }
    // ---- guess, IR2S, test; a guess worse than the zero guess is replaced
    const bool warm = I.warm != nullptr;
    const bool dwr = (C.flags & FADESIM_F_DWR) != 0;
    const double stop_tol = dwr ? C.tol_eff * C.dwr_scale : C.tol_eff;
    float p[TR][TC], q[TR][TC];
    bool rejected = false;
    double est;
    for (int attempt = 0;; ++attempt) {
        fast_guess<G>(I, C, sh, warm && attempt == 0, rwire, i0, j0, t, tx, ty, p, q);
        publish_all<G>(p, q, sh.halo, t);
        __syncthreads();
        est = reg_test<G>(I, p, q, zqp, i0, j0, t, tx, ty, a, rwire, has_l, has_r, has_u, has_d, sh);
        if (!(warm && attempt == 0) || est <= zero_guess_est<G>(sh, t))
            break;
        rejected = true;
    }
    const bool nonfinite = !(est < INFINITY);
    if (!nonfinite && !(est <= stop_tol))
        return false;
    // ---- outputs (as the sweeping kernel writes them)
    if (t < N) {
        double Iv = sh.colI[t];
        if (dwr && !nonfinite)
            Iv -= sh.colD[t];
        I.iout[t] = Iv;
    }
    if (t == 0) {
        if (I.sweeps)
            *I.sweeps = 0;
        if (I.status)
            *I.status = nonfinite ? FADESIM_ST_NONFINITE : FADESIM_ST_CONVERGED;
        if (I.est)
            *I.est = est;
        if (I.rejected)
            *I.rejected = rejected ? 1 : 0;
    }
    if (I.field) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            strow<TC>(I.field + (size_t)(i0 + ii) * N + j0, p[ii]);
            strow<TC>(I.field + (size_t)M * N + (size_t)(i0 + ii) * N + j0, q[ii]);
        }
    }
    __syncthreads(); // colI / halo reads done before the next instance reuses them
    return true;
}

} // namespace rb
} // namespace fadesim
