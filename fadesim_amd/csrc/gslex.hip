// The reference's lexicographic Gauss-Seidel/SOR, reproduced bit for bit on
// the GPU: the solver behind the `sim` CLI's nodebasedGSMethod command.
//
// The sweep of NodebasedGSMethod::nodeBaseGS
// (crossbarHspice/V2/concrete_command.h:1218-1273) and array2D::nodeBasedGS
// (crossbarHspice/2dArray.cc:1554-1609) visits rows i = 0..n-1, columns
// j = 0..m-1, and updates u_ij then d_ij from the newest values:
//   u_ij needs u_i,j-1 (new), u_i,j+1 (old), d_ij (old);
//   d_ij needs d_i-1,j (new), d_i+1,j (old), u_ij (new).
// Every new value node (i,j) reads lies on anti-diagonal i+j-1, every old one
// on diagonal i+j or i+j+1.  So one sweep is n+m-1 wavefront steps: step t
// updates the nodes of diagonal t, each from the same operands, in the same
// expression order, as the CPU loop -- the same bits.
//
// Mapping: one 64-lane wavefront per crossbar, lane l owns rows l, l+64, ...
// (R rows per lane).  Row i visits its columns in order, so each row keeps
// its own sliding window: u_i,j-1 (new) in a register, the old values of
// columns j and j+1 and the coefficients of column j in a ring of P register
// slots filled P steps ahead from L2.  The two values that cross rows come
// from the neighbouring lane: d_i-1,j (new, computed by row i-1 one step
// earlier) and d_i+1,j (old, the look-ahead of row i+1).  No barrier, no LDS.
//
// Fields and coefficients live in the workspace in diagonal-major order,
// node (i,j) at q = (i+j+1) n + i for j in [-1, m-1], so the lanes of a step
// touch consecutive addresses.  Coefficient planes (double2 = value,
// reciprocal): c0 = RU[i][j+1], c1 = RD[j][i], c2 = RD[j][i+1], c3 = the
// wordline node's conductance sum, c4 the bitline node's, cg = (g_ij, the
// end drive of the bitline at row 0 or n-1).  At column -1, c0 holds
// RU[i][0], the segment that row i's first node reads.
//
// Division.  Every quotient of the reference is a division by a value that
// does not change over the sweeps (a segment resistance or a conductance
// sum), so its reciprocal y = RN(1/b) is computed once, and a/b as
//   q = RN(a y), r = a - b q (exact, one fma), RN(q + r y)
// which is RN(a/b) when y = RN(1/b) and q is within an ulp of a/b (Markstein's
// theorem for round-to-nearest, no underflow; a = 0 keeps its sign).  The
// setup checks that every divisor and reciprocal is a finite normal number;
// otherwise the kernel divides with the IEEE division instead.
//
// General boundary (array2D::nodeBasedGeneralGS, crossbarHspice/2dArray.cc:
// 1209-1374): the same wavefront with that solver's expression forms (end
// terms drive / 0 over the end segment, sums g + a + b) -- GEN = true.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "common.h"
#include "fadesim.h"
#include "gslex.h"

namespace fadesim {
namespace {

constexpr int kLanes = 64;

struct LexArgs {
    int n, m;
    int iter_number;
    double w, omw;
    int enable_break;
    double tol;
    const double *g, *ru, *rd;
    int ru_ld, rd_ld;
    const double *vin, *vout;
    const int *on_top, *on_left, *on_bottom, *on_right;
    const double *vtop, *vright;
    const double *init_u, *init_d;
    double2 *F, *c0, *c1, *c2, *c3, *c4, *cg;
    int *slow; // 1 when some divisor needs the IEEE division
    double *vup, *vdown, *iout;
    int *record, *sweeps;
};

__device__ __forceinline__ size_t qidx(int n, int i, int j) { return (size_t)(i + j + 1) * n + i; }

__device__ __forceinline__ bool nice(double b)
{
    const double a = fabs(b);
    const double y = fabs(1.0 / b);
    return a >= DBL_MIN && a <= DBL_MAX && y >= DBL_MIN && y <= DBL_MAX;
}

// Coefficient planes of every node (i,j), j in [-1, m-1]; initial fields.
template <bool GEN>
__global__ void lex_setup(LexArgs A)
{
    const int n = A.n, m = A.m;
    const size_t total = (size_t)(n + m) * n;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n);
        const int j = (int)(q / n) - 1 - i;
        if (j < -1 || j >= m)
            continue;
        const double *RU = A.ru + (size_t)i * A.ru_ld;
        if (j == -1) {
            const double r0 = RU[0];
            A.c0[q] = make_double2(r0, 1.0 / r0);
            if (!nice(r0)) // row i's first node divides by RU[i][0] in both forms
                atomicOr(A.slow, 1);
            continue;
        }
        const double *RDj = A.rd + (size_t)j * A.rd_ld;
        const double gij = A.g[(size_t)i * m + j];
        const double ra = RU[j], rb = RU[j + 1], da = RDj[i], db = RDj[i + 1];
        double su, sd, edge = 0.0;
        bool ok = true;
        if (GEN) {
            const bool onL = A.on_left[i] != 0, onR = A.on_right[i] != 0;
            const bool onT = A.on_top[j] != 0, onB = A.on_bottom[j] != 0;
            su = gij + ((j != 0 || onL) ? 1.0 / ra : 0.0) + ((j != m - 1 || onR) ? 1.0 / rb : 0.0);
            sd = gij + ((i != 0 || onT) ? 1.0 / da : 0.0) + ((i != n - 1 || onB) ? 1.0 / db : 0.0);
            if (i == 0)
                edge = onT ? A.vtop[j] : 0.0;
            else if (i == n - 1)
                edge = onB ? A.vout[j] : 0.0;
            ok = nice(rb) && nice(da) && nice(db);
        } else {
            su = (j == m - 1) ? 1.0 / ra + gij : 1.0 / ra + 1.0 / rb + gij;
            sd = (i == 0) ? 1.0 / db + gij : 1.0 / da + 1.0 / db + gij;
            if (i == n - 1)
                edge = A.vout[j];
            ok = (j == m - 1 || nice(rb)) && (i == 0 || nice(da)) && nice(db);
        }
        ok = ok && nice(su) && nice(sd);
        if (!ok)
            atomicOr(A.slow, 1);
        A.c0[q] = make_double2(rb, 1.0 / rb);
        A.c1[q] = make_double2(da, 1.0 / da);
        A.c2[q] = make_double2(db, 1.0 / db);
        A.c3[q] = make_double2(su, 1.0 / su);
        A.c4[q] = make_double2(sd, 1.0 / sd);
        A.cg[q] = make_double2(gij, edge);
        double u0, d0;
        if (A.init_u) {
            u0 = A.init_u[(size_t)i * m + j];
            d0 = A.init_d[(size_t)i * m + j];
        } else if (GEN) {
            u0 = A.on_left[i] != 0 ? A.vin[i] : 0.0;
            d0 = 0.0;
        } else {
            u0 = A.vin[i];
            d0 = A.vout[j];
        }
        A.F[q] = make_double2(u0, d0);
    }
}

template <bool FAST>
__device__ __forceinline__ double qdiv(double a, double2 b)
{
    if (!FAST)
        return a / b.x;
    const double q = a * b.y;
    const double r = __builtin_fma(-q, b.x, a);
    const double q1 = __builtin_fma(r, b.y, q);
    return a == 0.0 ? q : q1;
}

struct Slot {
    double2 c0, c1, c2, c3, c4, cg;
    double2 f; // old (u, d) of column j+1
};

struct RowState {
    double uL;      // new u_i,j-1
    double uO, dO;  // old u_ij, d_ij
    double2 ra;     // RU[i][j] and its reciprocal
    double dUp;     // new d_i-1,j
    double lnum;    // left drive of the row (numerator of its first wordline segment)
    double rnum;    // right drive (general boundary)
};

__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __shfl(v, src, kLanes);
}

// Fills a slot for row i at column c: the coefficients of column c (c0 also
// at c = -1) and the old fields of column c+1; out-of-range parts are skipped.
__device__ __forceinline__ void load_slot(const LexArgs &A, Slot &s, int i, int c)
{
    const int n = A.n, m = A.m;
    if (i < n && c >= -1 && c < m) {
        const size_t q = qidx(n, i, c);
        s.c0 = A.c0[q];
        if (c >= 0) {
            s.c1 = A.c1[q];
            s.c2 = A.c2[q];
            s.c3 = A.c3[q];
            s.c4 = A.c4[q];
            s.cg = A.cg[q];
        }
        if (c + 1 < m)
            s.f = A.F[qidx(n, i, c + 1)];
    }
}

template <int R, int P, bool GEN, bool FAST>
__device__ void lex_sweeps(const LexArgs &A, RowState (&rs)[R])
{
    const int n = A.n, m = A.m, lane = threadIdx.x;
    const double w = A.w, omw = A.omw;
    const int T = n + m; // steps t = -1 .. n+m-2
    Slot sl[R][P] = {};
    int record = -1, it = 0;
    for (; it < A.iter_number; ++it) {
        // prologue: slots of steps -1 .. P-2
#pragma unroll
        for (int k = 0; k < P; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + kLanes * r;
                load_slot(A, sl[r][k], i, k - 1 - i);
            }
#pragma unroll
        for (int r = 0; r < R; ++r)
            rs[r].dUp = 0.0;
        double mx = 0.0;
        for (int ub = 0; ub < T; ub += P) {
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const int t = ub + k - 1;
                // d_i+1,j (old): the look-ahead of row i+1
                double below[R];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    below[r] = shfl_d(sl[r][k].f.y, (lane + 1) & (kLanes - 1));
                double dn[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    Slot &S = sl[r][k];
                    RowState &X = rs[r];
                    const int i = lane + kLanes * r;
                    const int j = t - i;
                    dn[r] = 0.0;
                    if (i < n && j >= 0 && j < m) {
                        const double g = S.cg.x;
                        // wordline node
                        const double l = (j == 0) ? X.lnum : X.uL;
                        const double a1 = qdiv<FAST>(l, X.ra);
                        double s;
                        if (GEN) {
                            const double rr = (j == m - 1) ? X.rnum : S.f.x;
                            s = (a1 + qdiv<FAST>(rr, S.c0)) + X.dO * g;
                        } else {
                            s = (j == m - 1) ? a1 + g * X.dO : (a1 + qdiv<FAST>(S.f.x, S.c0)) + g * X.dO;
                        }
                        const double un = qdiv<FAST>(w * s, S.c3) + omw * X.uO;
                        // bitline node
                        double dbel;
                        if (lane == kLanes - 1)
                            dbel = (r + 1 < R) ? below[r + 1 < R ? r + 1 : r] : 0.0;
                        else
                            dbel = below[r];
                        const double rr = (i == n - 1) ? S.cg.y : dbel;
                        const double b2 = qdiv<FAST>(rr, S.c2);
                        double s2;
                        if (GEN) {
                            const double lt = (i == 0) ? S.cg.y : X.dUp;
                            s2 = (qdiv<FAST>(lt, S.c1) + b2) + un * g;
                        } else {
                            s2 = (i == 0) ? b2 + g * un : (qdiv<FAST>(X.dUp, S.c1) + b2) + g * un;
                        }
                        const double d1 = qdiv<FAST>(w * s2, S.c4) + omw * X.dO;
                        mx = fmax(mx, fabs(d1 - X.dO));
                        mx = fmax(mx, fabs(un - X.uO));
                        A.F[qidx(n, i, j)] = make_double2(un, d1);
                        X.uL = un;
                        dn[r] = d1;
                    }
                    // slide the window: column j+1 becomes current; then the
                    // slot is refilled for step t + P (the later rows of this
                    // step hide the load)
                    X.uO = S.f.x;
                    X.dO = S.f.y;
                    X.ra = S.c0;
                    load_slot(A, S, i, t + P - i);
                }
                // d_i-1,j+1 (new) for the next step: row i-1's result of this step
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double up = shfl_d(dn[r], (lane + kLanes - 1) & (kLanes - 1));
                    if (lane != 0)
                        rs[r].dUp = up;
                    else if (r + 1 < R)
                        rs[r + 1 < R ? r + 1 : r].dUp = up; // lane 63 of slot r feeds lane 0 of slot r+1
                }
            }
        }
        if (A.enable_break) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                mx = fmax(mx, __shfl_xor(mx, off, kLanes));
            if (mx < A.tol) {
                record = it;
                ++it;
                break;
            }
        }
    }
    if (lane == 0) {
        *A.record = record;
        *A.sweeps = it;
    }
}

template <int R, int P, bool GEN>
__global__ void __launch_bounds__(kLanes) lex_kernel(LexArgs A)
{
    const int n = A.n, m = A.m, lane = threadIdx.x;
    RowState rs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + kLanes * r;
        rs[r].uL = 0.0;
        rs[r].dUp = 0.0;
        rs[r].lnum = 0.0;
        rs[r].rnum = 0.0;
        if (i < n) {
            if (GEN) {
                rs[r].lnum = A.on_left[i] != 0 ? A.vin[i] : 0.0;
                rs[r].rnum = A.on_right[i] != 0 ? A.vright[i] : 0.0;
            } else {
                rs[r].lnum = A.vin[i];
            }
        }
    }
    if (*A.slow)
        lex_sweeps<R, P, GEN, false>(A, rs);
    else
        lex_sweeps<R, P, GEN, true>(A, rs);
}

// Row-major fields and the readout: I_j = sum_i (u_ij - d_ij) g_ij over
// ascending rows (concrete_command.h:1344-1348), or the sink wire current of a
// driven bottom (2dArray.cc:1351-1356) for the general boundary.
template <bool GEN>
__global__ void lex_readout(LexArgs A)
{
    const int n = A.n, m = A.m;
    const size_t nm = (size_t)n * m;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nm; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / m), j = (int)(k % m);
        const double2 f = A.F[qidx(n, i, j)];
        A.vup[k] = f.x;
        A.vdown[k] = f.y;
    }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        if (GEN) {
            A.iout[j] = A.on_bottom[j] != 0
                            ? (A.F[qidx(n, n - 1, j)].y - A.vout[j]) / A.rd[(size_t)j * A.rd_ld + n]
                            : 0.0;
            continue;
        }
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double2 f = A.F[qidx(n, i, j)];
            s += (f.x - f.y) * A.g[(size_t)i * m + j];
        }
        A.iout[j] = s;
    }
}

template <int R, int P>
void launch_rp(const LexArgs &a, bool gen, hipStream_t s)
{
    if (gen)
        hipLaunchKernelGGL((lex_kernel<R, P, true>), dim3(1), dim3(kLanes), 0, s, a);
    else
        hipLaunchKernelGGL((lex_kernel<R, P, false>), dim3(1), dim3(kLanes), 0, s, a);
}

} // namespace

size_t gslex_workspace_bytes(int n, int m)
{
    return (size_t)(n + m) * n * 7 * sizeof(double2) + 256;
}

int gslex_solve(const fadesim_gs64_desc *d, void *ws, hipStream_t s)
{
    const int n = d->rows, m = d->cols;
    if (n > kLanes * 8)
        return set_error(FADESIM_E_UNSUPPORTED,
                         "the lexicographic order handles up to 512 rows (use FADESIM_GS64_REDBLACK)");
    LexArgs a;
    a.n = n;
    a.m = m;
    a.iter_number = d->iter_number;
    a.w = d->omega;
    a.omw = 1 - d->omega;
    a.enable_break = d->enable_break;
    a.tol = d->tol;
    a.g = d->d_g;
    a.ru = d->d_ru;
    a.rd = d->d_rd;
    a.ru_ld = d->ru_ld;
    a.rd_ld = d->rd_ld;
    a.vin = d->d_vin;
    a.vout = d->d_vout;
    a.on_top = d->d_on_top;
    a.on_left = d->d_on_left;
    a.on_bottom = d->d_on_bottom;
    a.on_right = d->d_on_right;
    a.vtop = d->d_v_top;
    a.vright = d->d_v_right;
    a.init_u = d->d_init_up;
    a.init_d = d->d_init_down;
    const size_t Q = (size_t)(n + m) * n;
    double2 *p = (double2 *)ws;
    a.F = p;
    a.c0 = p + Q;
    a.c1 = p + 2 * Q;
    a.c2 = p + 3 * Q;
    a.c3 = p + 4 * Q;
    a.c4 = p + 5 * Q;
    a.cg = p + 6 * Q;
    a.slow = (int *)(p + 7 * Q);
    a.vup = d->d_vup;
    a.vdown = d->d_vdown;
    a.iout = d->d_iout;
    a.record = d->d_record;
    a.sweeps = d->d_sweeps;
    const bool gen = d->general != 0;
    FADESIM_HIP_CHECK(hipMemsetAsync(a.slow, 0, sizeof(int), s));
    const unsigned blocks = (unsigned)std::min<size_t>((Q + 255) / 256, 2048);
    if (gen)
        hipLaunchKernelGGL(lex_setup<true>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(lex_setup<false>, dim3(blocks), dim3(256), 0, s, a);
    FADESIM_HIP_CHECK(hipGetLastError());
    const int R = (n + kLanes - 1) / kLanes;
    if (R == 1)
        launch_rp<1, 4>(a, gen, s);
    else if (R == 2)
        launch_rp<2, 3>(a, gen, s);
    else if (R <= 4)
        launch_rp<4, 1>(a, gen, s);
    else
        launch_rp<8, 1>(a, gen, s);
    FADESIM_HIP_CHECK(hipGetLastError());
    const unsigned rb = (unsigned)std::min<size_t>(((size_t)n * m + 255) / 256, 1024);
    if (gen)
        hipLaunchKernelGGL(lex_readout<true>, dim3(rb), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(lex_readout<false>, dim3(rb), dim3(256), 0, s, a);
    FADESIM_HIP_CHECK(hipGetLastError());
    return FADESIM_OK;
}

} // namespace fadesim
