// Batched IR-drop solver for gfx950: FP32 red-black SOR on drop variables,
// FP64 adjoint-weighted residual convergence test, one FP64 iterative
// refinement step and an FP64 bitline-current readout.
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
//   tt = fma(-h, x + x_other, rho),  x' = fma(w/2, x1 + x2, (1-w) x) + tt
//   h = w g / (2a),  rho = w rhs / (2a)  (rhs = g c, or the FP64 residual
//   after refinement).  At an open end (D~ = a) the missing neighbour is the
//   ghost x2 = x1 + (2/w) tt, which turns the update into the exact one
//   (1-w) x + w x1 + 2 tt.
//
// Layout: one workgroup per crossbar; thread (ty, tx) owns the TR x TC cell
// tile rows TR*ty.., columns TC*tx.. with p, q, hh, rho_p, rho_q in VGPRs.
// Per half-sweep a thread reads (TR + TC) halo values from LDS (the
// neighbours' edge values of the other colour), updates its TR*TC cells and
// publishes the TR + TC edge values it changed.  Conductances come from HBM
// once per sweep batch (coefficient reload after each convergence test).
//
// Convergence test (every few sweeps, FP64): residual r = b - A x at every
// node, then the first-order error of the readout
//   D_j = sum z_j . r,   I_j(x) - I_j* ~= D_j
// with the adjoint z_j taken from the wire Green's functions (bitline:
// r (n - max(i,i')) weighted by g of column j; wordline: r (min(j,j')+1)
// weighted by g_ij).  Stop when max_j |D_j| / |I_j| * safety <= tol.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "common.h"
#include "fadesim.h"

namespace fadesim {
namespace {

struct RbArgs {
    int batch;
    const float *g;
    const int *gidx;   // optional: instance b uses conductance slot gidx[b]
    int *queue;        // work counter (zeroed before the launch)
    const double *rw;
    const double *vin;
    const double *vout; // never null (zeros supplied by the host)
    const float *warm;
    float *ws; // per workgroup slot [7][M][N]: zq | rho p,q | snapshot p,q | residual p,q
    float omega;
    double tol_eff;    // tol / safety
    double switch_tol; // refinement switch point
    double dwr_scale;  // DWR: stop at est <= tol_eff * dwr_scale
    int max_sweeps;
    unsigned flags;
    double *iout;
    int *sweeps;
    int *status;
    double *est;
    float *field;
};

constexpr int kWsPlanes = 7;

__global__ void fill_zero_f64(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void zero_counter(int *c) { *c = 0; }

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d)
{
    *reinterpret_cast<float4 *>(p) = make_float4(a, b, c, d);
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

// Geometry of one template instance: TR x TC cells per thread, TY x TX threads.
template <int TR_, int TC_, int TY_, int TX_>
struct Geo {
    static constexpr int TR = TR_, TC = TC_, TY = TY_, TX = TX_;
    static constexpr int NT = TY * TX;
    static constexpr int M = TR * TY, N = TC * TX;
    static constexpr int NW = NT / 64;
    // LDS halo planes: p left/right columns [TR][NT], q top/bottom rows [TC][NT]
    static constexpr int PL = 0, PR = TR * NT, QT = 2 * TR * NT, QB = (2 * TR + TC) * NT;
    static constexpr int HALO = (2 * TR + 2 * TC) * NT;
    static_assert(TR % 2 == 0 && TC % 4 == 0, "tile sides: TR even, TC multiple of 4");
    static_assert(NT % 64 == 0 && 64 % TX == 0, "tile rows must not straddle waves");
    static_assert(NT >= M && NT >= N, "one thread per row and per column in the test");
};

template <int COL, class G>
__device__ __forceinline__ void half_sweep(float (&yp)[G::TR][G::TC], float (&yq)[G::TR][G::TC],
                                           const float (&hn)[G::TR][G::TC],
                                           const float (&rp)[G::TR][G::TC],
                                           const float (&rq)[G::TR][G::TC], float *halo, int t,
                                           bool has_l, bool has_r, bool has_u, bool has_d,
                                           float cA, float cB, float c2)
{
    constexpr int TR = G::TR, TC = G::TC, NT = G::NT, TX = G::TX;
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
            hu[k] = halo[G::QB + k * NT + (has_u ? t - TX : t)];
        } else { // lower neighbour of q at (TR-1,k); the sink at i = n-1
            const float v = halo[G::QT + k * NT + (has_d ? t + TX : t)];
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
                float t2 = cA * yp[ii][jj];
                t2 = fmaf(cB, n1 + n2, t2);
                yp[ii][jj] = t2 + tt;
            } else {
                const float tt = fmaf(hn[ii][jj], s, rq[ii][jj]);
                const float n2 = ii < TR - 1 ? yq[ii + 1][jj] : hd[jj];
                float n1;
                if (ii > 0)
                    n1 = yq[ii - 1][jj];
                else
                    n1 = has_u ? hu[jj] : fmaf(c2, tt, n2); // open top end
                float t2 = cA * yq[ii][jj];
                t2 = fmaf(cB, n1 + n2, t2);
                yq[ii][jj] = t2 + tt;
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
            halo[G::QT + k * NT + t] = yq[0][k];
        else
            halo[G::QB + k * NT + t] = yq[TR - 1][k];
    }
    __syncthreads();
}

template <class G>
__device__ __forceinline__ void publish_all(const float (&yp)[G::TR][G::TC],
                                            const float (&yq)[G::TR][G::TC], float *halo, int t)
{
    constexpr int NT = G::NT;
#pragma unroll
    for (int k = 0; k < G::TR; ++k) {
        halo[G::PL + k * NT + t] = yp[k][0];
        halo[G::PR + k * NT + t] = yp[k][G::TC - 1];
    }
#pragma unroll
    for (int k = 0; k < G::TC; ++k) {
        halo[G::QT + k * NT + t] = yq[0][k];
        halo[G::QB + k * NT + t] = yq[G::TR - 1][k];
    }
}

// Sweep coefficients: hn = -(w/2a) g and rho.  Before the refinement
// rho = (w/2a) g c for both nodes of a cell; after it rho is the scaled FP64
// residual the refinement step stored in the workspace.
template <class G>
__device__ __forceinline__ void load_coeffs(const RbArgs &A, const float *wsb, const float *gp, int b,
                                            int i0, int j0, double hcoef, bool refined,
                                            float (&hn)[G::TR][G::TC], float (&rp)[G::TR][G::TC],
                                            float (&rq)[G::TR][G::TC])
{
    constexpr int TR = G::TR, TC = G::TC, M = G::M, N = G::N;
    float vo[TC];
    if (!refined) {
#pragma unroll
        for (int jj = 0; jj < TC; ++jj)
            vo[jj] = 0.f; // placeholder, the FP64 value is reloaded below
    }
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
            const double vi = A.vin[(size_t)b * M + i0 + ii];
#pragma unroll
            for (int jj = 0; jj < TC; ++jj) {
                const double c = vi - A.vout[(size_t)b * N + j0 + jj];
                rp[ii][jj] = rq[ii][jj] = (float)(((double)g[jj] * c) * hcoef);
            }
        }
    }
    (void)vo;
}

// Convergence test, run on the snapshot of the drops in the workspace (only a
// few values per thread are live).  For every node the FP64 residual
// r = b - A x (b = g c before the refinement, the stored refinement residual
// after it); then
//   D_j = sum_i [ g_ij r W_i(j) + zq_ij r_q(i,j) ],  W_i(j) = sum_j' (min(j,j')+1) r_p(i,j')
//   I_j = S1_j - sum_i g_ij (p_ij + q_ij)
// with D_j the adjoint-weighted first-order error of I_j.  want_rho stores the
// scaled residual as the refinement right-hand side instead of testing.
// Results: colI[j], colD[j] (threads t < N), return value = max_j |D_j|/|I_j|.
template <class G>
__device__ double conv_test(const RbArgs &A, float *wsb, const float *gp, int b, int t, double a,
                            double rwire, double hcoef, bool refined, bool want_rho,
                            const double *S1, const double *Ifloor, double *colI, double *colD,
                            float *wmax)
{
    constexpr int M = G::M, N = G::N, NT = G::NT, NW = G::NW;
    const float *zq = wsb;
    const float *rho_p = wsb + 1 * M * N, *rho_q = wsb + 2 * M * N;
    const float *fp = wsb + 3 * M * N, *fq = wsb + 4 * M * N;
    float *resp = wsb + 5 * M * N, *resq = wsb + 6 * M * N;
    const float *g = gp;
    // 1. residuals, one node pair per thread and step
    for (int k = t; k < M * N; k += NT) {
        const int i = k / N, j = k - i * N;
        const double gd = (double)g[k];
        const double x = (double)fp[k], z = (double)fq[k];
        const double cell = gd * (x + z);
        double bp, bq;
        if (refined) {
            bp = (double)rho_p[k] / hcoef;
            bq = (double)rho_q[k] / hcoef;
        } else {
            bp = bq = gd * (A.vin[(size_t)b * M + i] - A.vout[(size_t)b * N + j]);
        }
        const double pl = j > 0 ? (double)fp[k - 1] : 0.0;
        const double wp = (j < N - 1) ? a * ((pl - x) + ((double)fp[k + 1] - x)) : a * (pl - x);
        const double qd = i < M - 1 ? (double)fq[k + N] : 0.0;
        const double wq = (i > 0) ? a * (((double)fq[k - N] - z) + (qd - z)) : a * (qd - z);
        const double rpv = (bp - cell) + wp, rqv = (bq - cell) + wq;
        if (want_rho) {
            wsb[1 * M * N + k] = (float)(rpv * hcoef);
            wsb[2 * M * N + k] = (float)(rqv * hcoef);
        } else {
            resp[k] = (float)rpv;
            resq[k] = (float)rqv;
        }
    }
    __syncthreads();
    if (want_rho)
        return 0.0;
    // 2. wordline adjoint: one thread per row, v = g r W + zq r_q into resp
    if (t < M) {
        const int i = t;
        double R = 0.0;
        for (int j = 0; j < N; ++j)
            R += (double)resp[i * N + j];
        double P0 = 0.0, P1 = 0.0;
        for (int j = 0; j < N; ++j) {
            const int k = i * N + j;
            const double r = (double)resp[k];
            const double jp1 = (double)(j + 1);
            const double W = P1 + jp1 * (R - P0);
            resp[k] = (float)((double)g[k] * rwire * W + (double)zq[k] * (double)resq[k]);
            P0 += r;
            P1 += jp1 * r;
        }
    }
    __syncthreads();
    // 3. column sums, one thread per column
    float emax = 0.f;
    if (t < N) {
        const int j = t;
        double D = 0.0, Y = 0.0;
        for (int i = 0; i < M; ++i) {
            const int k = i * N + j;
            D += (double)resp[k];
            Y += (double)g[k] * ((double)fp[k] + (double)fq[k]);
        }
        const double I = S1[j] - Y;
        colI[j] = I;
        colD[j] = D;
        const double e = fabs(D) / fmax(fabs(I), Ifloor[j]);
        emax = (e == e) ? (float)fmin(e, 3.0e38) : INFINITY;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        emax = fmaxf(emax, __shfl_xor(emax, off, 64));
    if ((t & 63) == 0)
        wmax[t >> 6] = emax;
    __syncthreads();
    float em = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w)
        em = fmaxf(em, wmax[w]);
    // identical in every lane: make it a scalar
    em = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, em)));
    return (double)em;
}

template <class G>
__device__ __forceinline__ void snapshot(float *wsb, const float (&yp)[G::TR][G::TC],
                                         const float (&yq)[G::TR][G::TC], int i0, int j0)
{
    constexpr int M = G::M, N = G::N;
#pragma unroll
    for (int ii = 0; ii < G::TR; ++ii) {
        strow<G::TC>(wsb + 3 * M * N + (size_t)(i0 + ii) * N + j0, yp[ii]);
        strow<G::TC>(wsb + 4 * M * N + (size_t)(i0 + ii) * N + j0, yq[ii]);
    }
}

struct RbShared {
    float *halo;
    double *S1, *Ifloor, *colI, *colD;
    float *wmax;
};

// Solve instance b with this workgroup's workspace slot wsb.
template <class G>
__device__ __forceinline__ void solve_instance(const RbArgs &A, int b, float *wsb, const RbShared &sh)
{
    constexpr int TR = G::TR, TC = G::TC, TX = G::TX, TY = G::TY;
    constexpr int M = G::M, N = G::N;
    float *halo = sh.halo;
    double *S1 = sh.S1, *Ifloor = sh.Ifloor, *colI = sh.colI, *colD = sh.colD;
    float *wmax = sh.wmax;

    const int t = threadIdx.x;
    const int tx = t % TX, ty = t / TX;
    const size_t base = (size_t)b * M * N;
    const float *gp = A.g + (size_t)(A.gidx ? A.gidx[b] : b) * M * N;
    const int i0 = TR * ty, j0 = TC * tx;
    const bool has_l = tx > 0, has_r = tx < TX - 1, has_u = ty > 0, has_d = ty < TY - 1;

    const float wf = A.omega;
    const float cA = 1.0f - wf;
    const float cB = 0.5f * wf;
    const float c2 = 2.0f / wf;
    const double rwire = A.rw[b];
    const double a = 1.0 / rwire;
    const double hcoef = 0.5 * (double)wf * rwire;
    const bool fixed = (A.flags & FADESIM_F_FIXED_SWEEPS) != 0;
    const bool dwr = (A.flags & FADESIM_F_DWR) != 0;

    // ideal readout, the per-column floor of the relative test and the
    // bitline adjoint zq(i,j) = r sum_i' g(i',j) (n - max(i,i'))
    if (t < N) {
        double si = 0.0, sf = 0.0, T = 0.0;
        const double vo = A.vout[(size_t)b * N + t];
        for (int i = 0; i < M; ++i) {
            const double gd = (double)gp[(size_t)i * N + t];
            const double c = A.vin[(size_t)b * M + i] - vo;
            si += gd * c;
            sf += gd * fabs(c);
            T += gd * (double)(M - i);
        }
        double Ac = 0.0, Bs = T;
        for (int i = 0; i < M; ++i) {
            const double gd = (double)gp[(size_t)i * N + t];
            Ac += gd;
            Bs -= gd * (double)(M - i);
            wsb[(size_t)i * N + t] = (float)(rwire * ((double)(M - i) * Ac + Bs));
        }
        S1[t] = si;
        // a column carrying under 1e-9 of its driven current is judged absolutely
        Ifloor[t] = fmax(1e-9 * sf, 1e-300);
    }

    float yp[TR][TC], yq[TR][TC], hn[TR][TC], rp[TR][TC], rq[TR][TC];
#pragma unroll
    for (int ii = 0; ii < TR; ++ii) {
        if (A.warm) {
            ldrow<TC>(A.warm + 2 * base + (size_t)(i0 + ii) * N + j0, yp[ii]);
            ldrow<TC>(A.warm + 2 * base + (size_t)M * N + (size_t)(i0 + ii) * N + j0, yq[ii]);
        } else {
#pragma unroll
            for (int jj = 0; jj < TC; ++jj)
                yp[ii][jj] = yq[ii][jj] = 0.f;
        }
    }
    publish_all<G>(yp, yq, halo, t);
    __syncthreads();

    // check scheduling only: single-precision hardware log2
    const float lrho = __log2f(fmaxf(wf - 1.0f, 1e-3f));
    bool refined = false;
    bool refine_pending = !(A.flags & FADESIM_F_NO_REFINE) && !fixed;
    int sweep = 0;
    int next_check = fixed ? A.max_sweeps : (A.max_sweeps < 32 ? A.max_sweeps : 32);
    int st = FADESIM_ST_MAXIT;
    double est = INFINITY;
    const double stop_tol = dwr ? A.tol_eff * A.dwr_scale : A.tol_eff;

    for (;;) {
        load_coeffs<G>(A, wsb, gp, b, i0, j0, hcoef, refined, hn, rp, rq);
        while (sweep < next_check) {
            half_sweep<0, G>(yp, yq, hn, rp, rq, halo, t, has_l, has_r, has_u, has_d, cA, cB, c2);
            half_sweep<1, G>(yp, yq, hn, rp, rq, halo, t, has_l, has_r, has_u, has_d, cA, cB, c2);
            ++sweep;
        }
        snapshot<G>(wsb, yp, yq, i0, j0);
        __syncthreads();
        est = conv_test<G>(A, wsb, gp, b, t, a, rwire, hcoef, refined, false, S1, Ifloor, colI,
                           colD, wmax);
        if (fixed) {
            st = FADESIM_ST_CONVERGED;
            break;
        }
        if (!(est < INFINITY)) {
            st = FADESIM_ST_NONFINITE;
            break;
        }
        if (refine_pending && est <= A.switch_tol) {
            // ---- one iterative-refinement step: rhs <- exact FP64 residual ----
            conv_test<G>(A, wsb, gp, b, t, a, rwire, hcoef, false, true, S1, Ifloor, colI, colD,
                         wmax);
            if (A.field) {
#pragma unroll
                for (int ii = 0; ii < TR; ++ii) {
                    strow<TC>(A.field + 2 * base + (size_t)(i0 + ii) * N + j0, yp[ii]);
                    strow<TC>(A.field + 2 * base + (size_t)M * N + (size_t)(i0 + ii) * N + j0, yq[ii]);
                }
            }
#pragma unroll
            for (int ii = 0; ii < TR; ++ii)
#pragma unroll
                for (int jj = 0; jj < TC; ++jj)
                    yp[ii][jj] = yq[ii][jj] = 0.f;
            if (t < N)
                S1[t] = colI[t];
            publish_all<G>(yp, yq, halo, t);
            __syncthreads();
            refine_pending = false;
            refined = true;
        } else if (!refine_pending && est <= stop_tol) {
            st = FADESIM_ST_CONVERGED;
            break;
        }
        if (sweep >= A.max_sweeps) {
            st = FADESIM_ST_MAXIT;
            break;
        }
        const double tgt = refine_pending ? A.switch_tol : stop_tol;
        const float k = (est > tgt) ? 0.8f * __log2f((float)(tgt / est)) / lrho : 16.0f;
        const int kk = (int)fminf(fmaxf(k, 16.0f), 256.0f);
        next_check = sweep + kk;
        if (next_check > A.max_sweeps)
            next_check = A.max_sweeps;
    }

    // ---- outputs ---------------------------------------------------------------
    if (t < N) {
        double I = colI[t];
        if (dwr && st == FADESIM_ST_CONVERGED)
            I -= colD[t];
        A.iout[(size_t)b * N + t] = I;
    }
    if (t == 0) {
        if (A.sweeps)
            A.sweeps[b] = sweep;
        if (A.status)
            A.status[b] = st;
        if (A.est)
            A.est[b] = fixed ? NAN : est;
    }
    if (A.field) {
#pragma unroll
        for (int ii = 0; ii < TR; ++ii) {
            float *P = A.field + 2 * base + (size_t)(i0 + ii) * N + j0;
            float *Q = P + (size_t)M * N;
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
}

template <class G>
__global__ void __launch_bounds__(G::NT) rbsor_kernel(RbArgs A)
{
    __shared__ float halo[G::HALO];
    __shared__ double S1[G::N];
    __shared__ double Ifloor[G::N];
    __shared__ double colI[G::N];
    __shared__ double colD[G::N];
    __shared__ float wmax[G::NW];
    __shared__ int s_inst;
    const RbShared sh{halo, S1, Ifloor, colI, colD, wmax};
    float *wsb = A.ws + (size_t)blockIdx.x * kWsPlanes * G::M * G::N;
    for (;;) {
        if (threadIdx.x == 0)
            s_inst = atomicAdd(A.queue, 1);
        __syncthreads();
        // wave-uniform: keep the instance index and everything derived from it in SGPRs
        const int b = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (b >= A.batch)
            break;
        solve_instance<G>(A, b, wsb, sh);
    }
}

thread_local hipEvent_t t_ev[2] = {nullptr, nullptr};
thread_local bool t_ev_valid = false;

// Resident workgroups of rbsor_kernel<G> on the current device (persistent grid).
template <class G>
int resident_slots()
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev >= 64)
        return 256;
    if (!cached[dev]) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rbsor_kernel<G>, G::NT, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        cached[dev] = per_cu * cus;
    }
    return cached[dev];
}

template <class G>
int launch_rb(RbArgs args, int M, int N, hipStream_t s)
{
    (void)M;
    (void)N;
    int grid = resident_slots<G>();
    if (grid > args.batch)
        grid = args.batch;
    const size_t ws_bytes = (size_t)grid * kWsPlanes * G::M * G::N * sizeof(float) + 256;
    char *ws = (char *)workspace_aux(ws_bytes);
    if (!ws)
        return FADESIM_E_NOMEM;
    args.ws = (float *)ws;
    args.queue = (int *)(ws + (size_t)grid * kWsPlanes * G::M * G::N * sizeof(float));
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, s, args.queue);
    FADESIM_HIP_CHECK(hipGetLastError());
    const bool timing = (args.flags & FADESIM_F_TIMING) != 0;
    if (timing) {
        if (!t_ev[0]) {
            FADESIM_HIP_CHECK(hipEventCreate(&t_ev[0]));
            FADESIM_HIP_CHECK(hipEventCreate(&t_ev[1]));
        }
        FADESIM_HIP_CHECK(hipEventRecord(t_ev[0], s));
    }
    hipLaunchKernelGGL((rbsor_kernel<G>), dim3(grid), dim3(G::NT), 0, s, args);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (timing) {
        FADESIM_HIP_CHECK(hipEventRecord(t_ev[1], s));
        t_ev_valid = true;
    }
    return FADESIM_OK;
}

} // namespace
} // namespace fadesim



using namespace fadesim;

extern "C" double fadesim_last_kernel_ms(void)
{
    if (!t_ev_valid)
        return -1.0;
    if (hipEventSynchronize(t_ev[1]) != hipSuccess)
        return -2.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t_ev[0], t_ev[1]) != hipSuccess)
        return -3.0;
    return (double)ms;
}

extern "C" double fadesim_default_omega(int rows, int cols)
{
    const int L = rows > cols ? rows : cols;
    const double s = std::sin(M_PI / (2.0 * L + 1.0));
    return 2.0 / (1.0 + 0.83 * s);
}

extern "C" int fadesim_solve_batch(const fadesim_batch_desc *d, void *stream)
{
    if (!d)
        return set_error(FADESIM_E_INVALID, "null descriptor");
    const int B = d->batch, M = d->rows, N = d->cols;
    if (B < 0)
        return set_error(FADESIM_E_INVALID, "batch must be >= 0");
    if (B == 0)
        return FADESIM_OK;
    if (!d->d_g || !d->d_r_wire || !d->d_vin || !d->d_iout)
        return set_error(FADESIM_E_INVALID, "d_g, d_r_wire, d_vin and d_iout are required");
    double w = d->omega > 0 ? d->omega : fadesim_default_omega(M, N);
    if (!(w > 0.0 && w < 2.0))
        return set_error(FADESIM_E_INVALID, "omega must lie in (0, 2)");
    if (d->max_sweeps < 0)
        return set_error(FADESIM_E_INVALID, "max_sweeps must be >= 0");
    const bool fixed = (d->flags & FADESIM_F_FIXED_SWEEPS) != 0;
    if (!fixed && !(d->tol > 0.0))
        return set_error(FADESIM_E_INVALID, "tol must be > 0");
    const double safety = d->safety > 0 ? d->safety : 2.0;
    if (safety < 1.0)
        return set_error(FADESIM_E_INVALID, "safety must be >= 1");
    hipStream_t s = (hipStream_t)stream;

    const double *vout = d->d_vout;
    if (!vout) {
        double *z = (double *)workspace((size_t)B * N * sizeof(double));
        if (!z)
            return FADESIM_E_NOMEM;
        hipLaunchKernelGGL(fill_zero_f64, dim3(256), dim3(256), 0, s, z, (size_t)B * N);
        FADESIM_HIP_CHECK(hipGetLastError());
        vout = z;
    }

    RbArgs args;
    args.batch = B;
    args.g = d->d_g;
    args.gidx = d->d_gidx;
    args.queue = nullptr;
    args.rw = d->d_r_wire;
    args.vin = d->d_vin;
    args.vout = vout;
    args.warm = d->d_warm;
    args.ws = nullptr;
    args.dwr_scale = 16.0;
    args.omega = (float)w;
    args.tol_eff = d->tol / safety;
    args.switch_tol = d->switch_tol > 0 ? d->switch_tol : 2e-5;
    args.max_sweeps = d->max_sweeps;
    args.flags = d->flags;
    args.iout = d->d_iout;
    args.sweeps = d->d_sweeps;
    args.status = d->d_status;
    args.est = d->d_est;
    args.field = d->d_field;

#define FADESIM_RB_CASE(TRv, TCv, TYv, TXv)                                 \
    if (M == TRv * TYv && N == TCv * TXv)                                   \
        return launch_rb<Geo<TRv, TCv, TYv, TXv>>(args, M, N, s);
    FADESIM_RB_CASE(4, 8, 32, 16) // 128 x 128, 512 threads
    FADESIM_RB_CASE(4, 8, 16, 8)  // 64 x 64, 128 threads
    FADESIM_RB_CASE(2, 4, 16, 8)  // 32 x 32, 128 threads
    FADESIM_RB_CASE(4, 8, 32, 8)  // 128 x 64
    FADESIM_RB_CASE(4, 8, 16, 16) // 64 x 128
#undef FADESIM_RB_CASE
    return set_error(FADESIM_E_UNSUPPORTED,
                     "rows x cols = " + std::to_string(M) + " x " + std::to_string(N) +
                         " not supported (batched kernel: 32x32, 64x64, 128x128, 128x64, 64x128)");
}
