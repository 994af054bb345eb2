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
constexpr int kOff = 2; // diagonal offset of the planes (prologue steps read below diagonal 0)

// Workspace layout: 7 planes (F, c0, c1, c2, c3, c4, cg) of nd diagonals x ld
// rows, ld = 64 R (rows padded to the wavefront), node (i,c) of column
// c in [-1, m-1] at (i + c + 1 + kOff) ld + i; every other entry is zero.
struct LexLayout {
    int ld, nd;
    size_t plane; // elements per plane
};

__host__ __device__ inline LexLayout lex_layout(int n, int m, int R, int P, int U)
{
    LexLayout L;
    L.ld = kLanes * R;
    L.nd = n + m + U + 2 * P + kOff + 4;
    L.plane = (size_t)L.ld * L.nd;
    return L;
}

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
    LexLayout L;
    double2 *W;  // 7 planes
    int *slow;   // 1 when some divisor needs the IEEE division
    double *vup, *vdown, *iout;
    int *record, *sweeps;
};

enum { kF = 0, kC0, kC1, kC2, kC3, kC4, kCG };

__device__ __forceinline__ size_t node_at(const LexLayout &L, int i, int c)
{
    return (size_t)(i + c + 1 + kOff) * L.ld + i;
}

__device__ __forceinline__ bool nice(double b)
{
    const double a = fabs(b);
    const double y = fabs(1.0 / b);
    return a >= DBL_MIN && a <= DBL_MAX && y >= DBL_MIN && y <= DBL_MAX;
}

// Coefficient planes of every node (i,c), c in [-1, m-1], initial fields, and
// zeros elsewhere.
template <bool GEN>
__global__ void lex_setup(LexArgs A)
{
    const int n = A.n, m = A.m;
    const LexLayout L = A.L;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < L.plane; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % L.ld);
        const int c = (int)(e / L.ld) - kOff - 1 - i;
        double2 v[7];
        for (int p = 0; p < 7; ++p)
            v[p] = make_double2(0.0, 0.0);
        if (i < n && c == -1) {
            const double r0 = A.ru[(size_t)i * A.ru_ld];
            v[kC0] = make_double2(r0, 1.0 / r0);
            if (!nice(r0)) // row i's first node divides by RU[i][0] in both forms
                atomicOr(A.slow, 1);
        } else if (i < n && c >= 0 && c < m) {
            const int j = c;
            const double *RU = A.ru + (size_t)i * A.ru_ld;
            const double *RDj = A.rd + (size_t)j * A.rd_ld;
            const double gij = A.g[(size_t)i * m + j];
            const double ra = RU[j], rb = RU[j + 1], da = RDj[i], db = RDj[i + 1];
            double su, sd, edge = 0.0;
            bool ok;
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
            if (!(ok && nice(su) && nice(sd)))
                atomicOr(A.slow, 1);
            v[kC0] = make_double2(rb, 1.0 / rb);
            v[kC1] = make_double2(da, 1.0 / da);
            v[kC2] = make_double2(db, 1.0 / db);
            v[kC3] = make_double2(su, 1.0 / su);
            v[kC4] = make_double2(sd, 1.0 / sd);
            v[kCG] = make_double2(gij, edge);
            if (A.init_u)
                v[kF] = make_double2(A.init_u[(size_t)i * m + j], A.init_d[(size_t)i * m + j]);
            else if (GEN)
                v[kF] = make_double2(A.on_left[i] != 0 ? A.vin[i] : 0.0, 0.0);
            else
                v[kF] = make_double2(A.vin[i], A.vout[j]);
        }
        for (int p = 0; p < 7; ++p)
            A.W[p * L.plane + e] = v[p];
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

// Whole-wavefront rotates by one lane (DPP wave_rol:1 / wave_ror:1, GFX9):
// from_next(v) in lane l is v of lane l+1 (lane 63: lane 0), from_prev(v) is v
// of lane l-1 (lane 0: lane 63).
template <int CTRL>
__device__ __forceinline__ double rot_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_next(double v) { return rot_d<0x134>(v); }
__device__ __forceinline__ double from_prev(double v) { return rot_d<0x13C>(v); }

// The slot of a row for the step whose coefficient diagonal starts at element
// `diag` (uniform): coefficients of the row's column there, old fields of the
// next column.  Out-of-range entries are zeros of the padded planes.
__device__ __forceinline__ void load_slot(const double2 *__restrict__ W, size_t plane, int ld, size_t diag, int i,
                                          Slot &s)
{
    const size_t q = diag + i;
    s.c0 = W[kC0 * plane + q];
    s.c1 = W[kC1 * plane + q];
    s.c2 = W[kC2 * plane + q];
    s.c3 = W[kC3 * plane + q];
    s.c4 = W[kC4 * plane + q];
    s.cg = W[kCG * plane + q];
    s.f = W[kF * plane + q + ld];
}

// Runs only when the setup's division check agrees with FAST (both variants
// are launched; one returns at once).  U steps per loop iteration (a multiple
// of P): the compiler drains every load in flight at the loop's back edge, so
// the ring refills only pay off inside one iteration.
template <int R, int P, int U, bool GEN, bool FAST>
__global__ void __launch_bounds__(kLanes) lex_kernel(LexArgs A)
{
    if ((*A.slow != 0) == FAST)
        return;
    const int n = A.n, m = A.m, lane = threadIdx.x;
    const double w = A.w, omw = A.omw;
    const int T = n + m; // steps t = -1 .. n+m-2
    const int ld = A.L.ld;
    const size_t plane = A.L.plane;
    double2 *__restrict__ W = A.W;
    RowState rs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + kLanes * r;
        rs[r].uL = 0.0;
        rs[r].uO = rs[r].dO = 0.0;
        rs[r].ra = make_double2(0.0, 0.0);
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
    Slot sl[R][P];
    int record = -1, it = 0;
    for (; it < A.iter_number; ++it) {
        // prologue: the slots of steps -1 .. P-2
#pragma unroll
        for (int k = 0; k < P; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r)
                load_slot(W, plane, ld, (size_t)(k + kOff) * ld, lane + kLanes * r, sl[r][k]);
#pragma unroll
        for (int r = 0; r < R; ++r)
            rs[r].dUp = 0.0;
        double mx = 0.0;
        for (int ub = 0; ub < T; ub += U) {
#pragma unroll
            for (int kk = 0; kk < U; ++kk) {
                const int k = kk % P;
                const int t = ub + kk - 1;
                // d_i+1,j (old): the look-ahead of row i+1
                double below[R];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    below[r] = from_next(sl[r][k].f.y);
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
                        const double dbel = (lane == kLanes - 1) ? ((r + 1 < R) ? below[r + 1 < R ? r + 1 : r] : 0.0)
                                                                 : below[r];
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
                        W[kF * plane + (size_t)(t + 1 + kOff) * ld + i] = make_double2(un, d1);
                        X.uL = un;
                        dn[r] = d1;
                    }
                    // slide the window: column j+1 becomes current; then the
                    // slot is refilled for step t + P (the later rows of this
                    // step hide the load)
                    X.uO = S.f.x;
                    X.dO = S.f.y;
                    X.ra = S.c0;
                    load_slot(W, plane, ld, (size_t)(t + P + 1 + kOff) * ld, i, S);
                }
                // d_i-1,j+1 (new) for the next step: row i-1's result of this step
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double up = from_prev(dn[r]);
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

// Row-major fields and the readout: I_j = sum_i (u_ij - d_ij) g_ij over
// ascending rows (concrete_command.h:1344-1348), or the sink wire current of a
// driven bottom (2dArray.cc:1351-1356) for the general boundary.
template <bool GEN>
__global__ void lex_readout(LexArgs A)
{
    const int n = A.n, m = A.m;
    const size_t nm = (size_t)n * m;
    const double2 *F = A.W + kF * A.L.plane;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nm; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / m), j = (int)(k % m);
        const double2 f = F[node_at(A.L, i, j)];
        A.vup[k] = f.x;
        A.vdown[k] = f.y;
    }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        if (GEN) {
            A.iout[j] = A.on_bottom[j] != 0 ? (F[node_at(A.L, n - 1, j)].y - A.vout[j]) / A.rd[(size_t)j * A.rd_ld + n]
                                            : 0.0;
            continue;
        }
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double2 f = F[node_at(A.L, i, j)];
            s += (f.x - f.y) * A.g[(size_t)i * m + j];
        }
        A.iout[j] = s;
    }
}

template <int R, int P, int U, bool GEN>
void launch_variants(const LexArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((lex_kernel<R, P, U, GEN, true>), dim3(1), dim3(kLanes), 0, s, a);
    hipLaunchKernelGGL((lex_kernel<R, P, U, GEN, false>), dim3(1), dim3(kLanes), 0, s, a);
}

template <int R, int P, int U>
void launch_rpu(const LexArgs &a, bool gen, hipStream_t s)
{
    if (gen)
        launch_variants<R, P, U, true>(a, s);
    else
        launch_variants<R, P, U, false>(a, s);
}

// Rows per lane R, ring depth P and steps per loop iteration U by row count.
void shape(int n, int &R, int &P, int &U)
{
    R = (n + kLanes - 1) / kLanes;
    if (R <= 1) {
        R = 1;
        P = 3;
        U = 12;
    } else if (R == 2) {
        P = 2;
        U = 12;
    } else if (R <= 4) {
        R = 4;
        P = 1;
        U = 8;
    } else {
        R = 8;
        P = 1;
        U = 2;
    }
}

} // namespace

size_t gslex_workspace_bytes(int n, int m)
{
    int R, P, U;
    shape(n, R, P, U);
    return lex_layout(n, m, R, P, U).plane * 7 * sizeof(double2) + 256;
}

int gslex_solve(const fadesim_gs64_desc *d, void *ws, hipStream_t s)
{
    const int n = d->rows, m = d->cols;
    if (n > kLanes * 8)
        return set_error(FADESIM_E_UNSUPPORTED,
                         "the lexicographic order handles up to 512 rows (FADESIM_GS64_REDBLACK handles more)");
    int R, P, U;
    shape(n, R, P, U);
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
    a.L = lex_layout(n, m, R, P, U);
    a.W = (double2 *)ws;
    a.slow = (int *)(a.W + 7 * a.L.plane);
    a.vup = d->d_vup;
    a.vdown = d->d_vdown;
    a.iout = d->d_iout;
    a.record = d->d_record;
    a.sweeps = d->d_sweeps;
    const bool gen = d->general != 0;
    FADESIM_HIP_CHECK(hipMemsetAsync(a.slow, 0, sizeof(int), s));
    const unsigned blocks = (unsigned)std::min<size_t>((a.L.plane + 255) / 256, 2048);
    if (gen)
        hipLaunchKernelGGL(lex_setup<true>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(lex_setup<false>, dim3(blocks), dim3(256), 0, s, a);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (R == 1)
        launch_rpu<1, 3, 12>(a, gen, s);
    else if (R == 2)
        launch_rpu<2, 2, 12>(a, gen, s);
    else if (R == 4)
        launch_rpu<4, 1, 8>(a, gen, s);
    else
        launch_rpu<8, 1, 2>(a, gen, s);
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
