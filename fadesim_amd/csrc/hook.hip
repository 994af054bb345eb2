// The PimTorch ir_drop_gs hook with the reference's own iteration, for
// callers that need its exact (often unconverged) results: line
// Gauss-Seidel/SOR, all wordlines left->right then all bitlines bottom->top
// each sweep (pimtorch/.../SplitArrCuda/split_arr_matmul.cu:27-119), readout
// sum_i (u - d) G (:121-140), sinh-nonlinear cells by the lagged Picard
// update (:142-196, 241-262).  Separate source / sink load conductances are
// ir_gs_ou_local (:690-815); OU mode, every ou_rows x ou_cols block of a tile
// solved as its own crossbar with position-dependent loads, is ir_gs_ou
// (:1086-1327).
//
// One workgroup per crossbar (b, x, y).  The fields are updated in place:
// walking a wordline left->right, the left neighbour is already new and the
// right one still old -- exactly the reference's v_up / v_up_prev pairing --
// and likewise bottom->top on a bitline.  Fields and the nonlinear
// conductances live in the workspace (L2); the wordline phase keeps its
// running left value and the bitline phase its running lower value in
// registers.  Arithmetic follows the reference expression by expression
// (-ffp-contract=off); only sinh differs in the last bits between devices.
//
// Nonlinear: sweep k uses G computed from the state two sweeps back (the
// reference computes it from the buffer it is about to overwrite), so each
// sweep first computes the next sweep's G from the current state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "common.h"
#include "fadesim.h"

namespace fadesim {
namespace {

constexpr int kHookThreads = 256;

template <typename T>
struct HookArgs {
    const T *input;
    const T *G;
    int bsize, row, col, numX, numY;
    int ou_rows, ou_cols, nOuX, nOuY; // OU mode: each ou_rows x ou_cols block is its own crossbar
    T gl, gd;                         // source / sink load conductances (whole-tile mode)
    int sweeps_this_launch;
    int first_sweep; // 0 on the first launch: initialise the fields
    T gw, be, omb, thr;
    int nonlinear;
    double nl[6];
    T *u, *d;         // [grid][row][col]
    T *Gc, *Gn;       // nonlinear conductances of this and the next sweep
    unsigned *moved;  // break mode: set when any node moved by more than thr
    int want_moved;
};

template <typename T> __device__ __forceinline__ T dev_sinh(T x);
template <> __device__ __forceinline__ float dev_sinh<float>(float x) { return sinhf(x); }
template <> __device__ __forceinline__ double dev_sinh<double>(double x) { return sinh(x); }
template <typename T> __device__ __forceinline__ T dev_abs(T x) { return x < (T)0 ? -x : x; }

// Nonlinear cell conductance of split_arr_matmul.cu:142-196 for voltage V.
template <typename T>
__device__ __forceinline__ T nonlinear_g(T V, T z, const double *nl)
{
    if (V < (T)1e-5 && V > -(T)1e-5)
        V = V > (T)0 ? (T)1e-5 : -(T)1e-5;
    V = dev_abs(V);
    const T Ion = (T)nl[3] * dev_sinh<T>(V * (T)nl[2]);
    const T Ioff = (T)nl[5] * dev_sinh<T>(V * (T)nl[4]);
    const T R = ((T)1 - z) * V / Ion + z * V / Ioff;
    return (T)1 / R;
}

// One solved crossbar: tile (b, x, y) or, in OU mode, OU (ou_x, ou_y) of it.
template <typename T>
struct HookUnit {
    int b, x, y, oux, ouy, r0, c0, rows, cols;
    T gl, gd;
};

template <typename T>
__device__ __forceinline__ HookUnit<T> hook_unit(const HookArgs<T> &A, int xb)
{
    HookUnit<T> U;
    U.ouy = xb % A.nOuY;
    U.oux = (xb / A.nOuY) % A.nOuX;
    const int t2 = xb / (A.nOuY * A.nOuX);
    U.y = t2 % A.numY;
    U.x = (t2 / A.numY) % A.numX;
    U.b = t2 / (A.numY * A.numX);
    if (A.ou_rows > 0) {
        // the OU sees the source through c0 + 1 wire segments and the sink
        // through row - (last row) segments (split_arr_matmul.cu:1109, 1166)
        U.rows = A.ou_rows;
        U.cols = A.ou_cols;
        U.r0 = U.oux * A.ou_rows;
        U.c0 = U.ouy * A.ou_cols;
        U.gl = A.gw / (T)(U.c0 + 1);
        U.gd = A.gw / (T)(A.row - (U.r0 + A.ou_rows - 1));
    } else {
        U.rows = A.row;
        U.cols = A.col;
        U.r0 = U.c0 = 0;
        U.gl = A.gl;
        U.gd = A.gd;
    }
    return U;
}

template <typename T>
__global__ void __launch_bounds__(kHookThreads) hook_kernel(HookArgs<T> A)
{
    const int xb = blockIdx.x; // (b, x, y[, ou_x, ou_y]) flattened
    const HookUnit<T> U = hook_unit(A, xb);
    const int b = U.b, x = U.x, y = U.y;
    const int row = U.rows, col = U.cols, t = threadIdx.x;
    const int nth = blockDim.x; // a multiple of 64 up to kHookThreads, sized to the unit's lines
    const size_t cells = (size_t)row * col;
    T *u = A.u + (size_t)xb * cells;
    T *d = A.d + (size_t)xb * cells;
    const T *in = A.input + ((size_t)b * A.numX + x) * A.row + U.r0;
    const T gl = U.gl, gd = U.gd;
    // linear G of cell (i, c): G[x][r0 + i][y*col + c0 + c]
    auto Glin = [&](int i, int c) -> T {
        return A.G[((size_t)x * A.row + U.r0 + i) * ((size_t)A.numY * A.col) + (size_t)y * A.col + U.c0 + c];
    };
    T *Gc = A.nonlinear ? A.Gc + (size_t)xb * cells : nullptr;
    T *Gn = A.nonlinear ? A.Gn + (size_t)xb * cells : nullptr;
    const T gw = A.gw, be = A.be, omb = A.omb;

    if (A.first_sweep) {
        for (size_t k = t; k < cells; k += nth) {
            u[k] = in[k / col];
            d[k] = (T)0;
        }
        __syncthreads();
        if (A.nonlinear) {
            // G of sweeps 0 and 1 both come from the initial state
            for (size_t k = t; k < cells; k += nth) {
                const int i = (int)(k / col), c = (int)(k % col);
                const T gl = Glin(i, c);
                T z = ((T)1 / gl - (T)A.nl[0]) / (T)(A.nl[1] - A.nl[0]);
                if (z < (T)0)
                    z = (T)0;
                Gc[k] = Gn[k] = nonlinear_g<T>(u[k] - d[k], z, A.nl);
            }
            __syncthreads();
        }
    }
    bool moved = false;
    for (int s = 0; s < A.sweeps_this_launch; ++s) {
        if (A.nonlinear && !(A.first_sweep && s == 0)) {
            // Gc <- Gn (computed last sweep); Gn <- G of the current state
            for (size_t k = t; k < cells; k += nth) {
                const int i = (int)(k / col), c = (int)(k % col);
                const T gl = Glin(i, c);
                T z = ((T)1 / gl - (T)A.nl[0]) / (T)(A.nl[1] - A.nl[0]);
                if (z < (T)0)
                    z = (T)0;
                Gc[k] = Gn[k];
                Gn[k] = nonlinear_g<T>(u[k] - d[k], z, A.nl);
            }
            __syncthreads();
        }
        auto Gof = [&](int i, int c) -> T { return A.nonlinear ? Gc[(size_t)i * col + c] : Glin(i, c); };
        // wordlines, left -> right (split_arr_matmul.cu:75-119)
        for (int i = t; i < row; i += nth) {
            T *ur = u + (size_t)i * col;
            const T *dr = d + (size_t)i * col;
            T left = in[i];
            T old = ur[0];
            T G0 = Gof(i, 0);
            T sumg = gw + gl + G0;
            T nv = be * (gl * left + gw * ur[1] + G0 * dr[0]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > A.thr;
            ur[0] = nv;
            left = nv;
            int j;
            for (j = 1; j < col - 1; ++j) {
                old = ur[j];
                const T Gj = Gof(i, j);
                sumg = gw * (T)2 + Gj;
                nv = be * (gw * (left + ur[j + 1]) + Gj * dr[j]) / sumg + omb * old;
                moved |= dev_abs(nv - old) > A.thr;
                ur[j] = nv;
                left = nv;
            }
            old = ur[j];
            const T Gj = Gof(i, j);
            sumg = gw + Gj;
            nv = be * (gw * left + Gj * dr[j]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > A.thr;
            ur[j] = nv;
        }
        __syncthreads();
        // bitlines, bottom -> top (split_arr_matmul.cu:27-73)
        for (int c = t; c < col; c += nth) {
            int i = row - 1;
            T old = d[(size_t)i * col + c];
            T Gi = Gof(i, c);
            T sumg = gw + gd + Gi;
            T nv = be * (gw * d[(size_t)(i - 1) * col + c] + gd * (T)0 + Gi * u[(size_t)i * col + c]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > A.thr;
            d[(size_t)i * col + c] = nv;
            T down = nv;
            for (i = row - 2; i > 0; --i) {
                old = d[(size_t)i * col + c];
                Gi = Gof(i, c);
                sumg = gw * (T)2 + Gi;
                nv = be * (gw * (d[(size_t)(i - 1) * col + c] + down) + Gi * u[(size_t)i * col + c]) / sumg + omb * old;
                moved |= dev_abs(nv - old) > A.thr;
                d[(size_t)i * col + c] = nv;
                down = nv;
            }
            old = d[c];
            Gi = Gof(0, c);
            sumg = gw + Gi;
            nv = be * (gw * down + Gi * u[c]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > A.thr;
            d[c] = nv;
        }
        __syncthreads();
    }
    if (A.want_moved && moved)
        atomicOr(A.moved, 1u);
}

// Whole-tile units of at most kLdsMax x kLdsMax cells: the same
// sweeps, operation by operation, with the wordline field in LDS for every
// sweep of the launch and, in FP32, the bitline field too (FP64: 132 KB for
// one field, so the bitline field stays in the workspace).  LDS rows have
// stride cols + 1, so the wordline walk -- one lane per row -- and the
// bitline walk -- one lane per column -- are free of bank conflicts.
// Everything a step reads from global memory -- the conductance (in the
// bitline walk from the tile, lanes on consecutive columns; in the wordline
// walk from a transposed copy made on the first launch, lanes on
// consecutive rows) and, in FP64, the bitline field's old values -- is
// loaded a block of kLdsAhead steps ahead.  The fields go to the workspace at
// the end of the launch (and come back at the start of the next one in break
// mode).
constexpr int kLdsMax = 128;

// NT threads: 128 (a lane per row or column of the walks) for linear cells;
// 256 for sinh cells, whose per-sweep conductance update over every cell
// then runs on four waves
template <typename T, int NT, bool NL>
__global__ void __launch_bounds__(NT) hook_lds_kernel(HookArgs<T> A, T *__restrict__ GTall, T *__restrict__ NTall,
                                                      T *__restrict__ Zall)
{
    constexpr int kLdsThreads = NT;
    // steps fetched ahead (measured: 16 in FP32; 8 in FP64, whose walk also
    // fetches the bitline field and holds twice the registers per value)
    constexpr int kLdsAhead = sizeof(T) == 4 ? 16 : 8;
    constexpr bool DL = sizeof(T) == 4; // the bitline field in LDS too
    __shared__ T su[kLdsMax * (kLdsMax + 1)];
    __shared__ T sd[DL ? kLdsMax * (kLdsMax + 1) : 1];
    const int xb = blockIdx.x;
    const HookUnit<T> U = hook_unit(A, xb);
    const int x = U.x, y = U.y;
    const int row = U.rows, col = U.cols, ld = col + 1, t = threadIdx.x;
    const size_t cells = (size_t)row * col;
    T *u = A.u + (size_t)xb * cells;
    T *d = A.d + (size_t)xb * cells;
    const T *in = A.input + ((size_t)U.b * A.numX + x) * A.row;
    // G(i, c) = Gt[i * gstride + c]; its transposed copy gT[c * row + i]
    const size_t gstride = (size_t)A.numY * A.col;
    const T *__restrict__ Gt = A.G + (size_t)x * A.row * gstride + (size_t)y * A.col;
    T *__restrict__ gT = GTall + (size_t)xb * cells;
    const T gl = U.gl, gd = U.gd, gw = A.gw, be = A.be, omb = A.omb, thr = A.thr;
    // the bitline field: LDS (FP32) or the workspace (FP64)
    auto dget = [&](int i, int c) -> T { return DL ? sd[i * ld + c] : d[(size_t)i * col + c]; };
    auto dput = [&](int i, int c, T v) {
        if constexpr (DL)
            sd[i * ld + c] = v;
        else
            d[(size_t)i * col + c] = v;
    };

    if (A.first_sweep) {
        for (size_t k = t; k < cells; k += kLdsThreads) {
            const int i = (int)(k / col), c = (int)(k % col);
            su[i * ld + c] = in[i];
            dput(i, c, (T)0);
        }
        for (size_t k = t; k < cells; k += kLdsThreads) {
            const int c = (int)(k / row), i = (int)(k % row);
            gT[k] = Gt[(size_t)i * gstride + c];
        }
    } else {
        for (size_t k = t; k < cells; k += kLdsThreads) {
            const int i = (int)(k / col), c = (int)(k % col);
            su[i * ld + c] = u[k];
            if constexpr (DL)
                sd[i * ld + c] = d[k];
        }
    }
    __syncthreads(); // the fields, and this workgroup's transposed conductance

    // Sinh cells (one launch for all sweeps): F_s, the conductance of the
    // state at the start of sweep s, goes to buffer s & 1 (natural layout:
    // A.Gc, A.Gn; transposed: NTall's two planes); sweep s walks with
    // F_{max(s-1, 0)} -- the reference's two-sweeps-back conductance
    // (split_arr_matmul.cu:241-262): its kernel reads the buffer it is about
    // to overwrite.  z of each cell (its linear conductance's place between
    // Ron and Roff) is computed once, by the reference's expression.
    constexpr bool nl = NL;
    T *Fn[2] = {nullptr, nullptr}, *FT[2] = {nullptr, nullptr}, *zc = nullptr;
    if constexpr (nl) {
        Fn[0] = A.Gc + (size_t)xb * cells;
        Fn[1] = A.Gn + (size_t)xb * cells;
        FT[0] = NTall + (size_t)xb * cells;
        FT[1] = NTall + ((size_t)gridDim.x + xb) * cells;
        zc = Zall + (size_t)xb * cells;
        for (size_t k = t; k < cells; k += kLdsThreads) {
            const int i = (int)(k / col), c = (int)(k % col);
            const T g = Gt[(size_t)i * gstride + c];
            T z = ((T)1 / g - (T)A.nl[0]) / (T)(A.nl[1] - A.nl[0]);
            if (z < (T)0)
                z = (T)0;
            zc[k] = z;
        }
    }
    // F of the present state into buffer b, both layouts
    auto nonlinear_update = [&](int b) {
        for (size_t k = t; k < cells; k += kLdsThreads) {
            const int i = (int)(k / col), c = (int)(k % col);
            const T v = nonlinear_g<T>(su[i * ld + c] - dget(i, c), zc[k], A.nl);
            Fn[b][k] = v;
            FT[b][(size_t)c * row + i] = v;
        }
    };

    bool moved = false;
    for (int s = 0; s < A.sweeps_this_launch; ++s) {
        // the walks' conductance: linear, or F_{max(s-1, 0)} of sinh cells
        const T *__restrict__ gnat = Gt;
        const T *__restrict__ gtr = gT;
        size_t gst = gstride;
        if constexpr (nl) {
            nonlinear_update(s & 1);
            __syncthreads();
            const int b = s > 0 ? (s - 1) & 1 : 0;
            gnat = Fn[b];
            gtr = FT[b];
            gst = (size_t)col;
        }
        // wordlines, left -> right (split_arr_matmul.cu:75-119)
        if (t < row) {
            const int i = t;
            T *sur = su + i * ld;
            const T *gti = gtr + i; // G(i, j) = gti[j * row]
            T left = in[i];
            T old = sur[0];
            const T G0 = gti[0];
            T sumg = gw + gl + G0;
            T nv = be * (gl * left + gw * sur[1] + G0 * dget(i, 0)) / sumg + omb * old;
            moved |= dev_abs(nv - old) > thr;
            sur[0] = nv;
            left = nv;
            // a block's loads past the walk's end reread node col - 1 and go
            // unused: unconditional, so the compiler waits for each block's
            // own loads only, not for the block fetched ahead
            T gc[kLdsAhead], gn[kLdsAhead], dc[DL ? 1 : kLdsAhead], dn[DL ? 1 : kLdsAhead];
#pragma unroll
            for (int k = 0; k < kLdsAhead; ++k) {
                const int j = min(1 + k, col - 1);
                gc[k] = gti[(size_t)j * row];
                if constexpr (!DL)
                    dc[k] = dget(i, j);
            }
            for (int j0 = 1; j0 < col - 1; j0 += kLdsAhead) {
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    const int j = min(j0 + kLdsAhead + k, col - 1);
                    gn[k] = gti[(size_t)j * row];
                    if constexpr (!DL)
                        dn[k] = dget(i, j);
                }
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    const int j = j0 + k;
                    if (j < col - 1) {
                        old = sur[j];
                        const T Gj = gc[k];
                        const T dj = DL ? dget(i, j) : dc[DL ? 0 : k];
                        sumg = gw * (T)2 + Gj;
                        nv = be * (gw * (left + sur[j + 1]) + Gj * dj) / sumg + omb * old;
                        moved |= dev_abs(nv - old) > thr;
                        sur[j] = nv;
                        left = nv;
                    }
                }
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    gc[k] = gn[k];
                    if constexpr (!DL)
                        dc[k] = dn[k];
                }
            }
            const int j = col - 1;
            old = sur[j];
            const T Gj = gti[(size_t)j * row];
            sumg = gw + Gj;
            nv = be * (gw * left + Gj * dget(i, j)) / sumg + omb * old;
            moved |= dev_abs(nv - old) > thr;
            sur[j] = nv;
        }
        __syncthreads();
        // bitlines, bottom -> top (split_arr_matmul.cu:27-73)
        if (t < col) {
            const int c = t;
            const T *gcol = gnat + c; // G(i, c) = gcol[i * gst]
            int i = row - 1;
            T old = dget(i, c);
            T up = dget(i - 1, c);
            T Gi = gcol[(size_t)i * gst];
            T sumg = gw + gd + Gi;
            T nv = be * (gw * up + gd * (T)0 + Gi * su[i * ld + c]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > thr;
            dput(i, c, nv);
            T down = nv;
            // old value of the next node up: the up value just read
            T gc[kLdsAhead], gn[kLdsAhead], uc[DL ? 1 : kLdsAhead], un[DL ? 1 : kLdsAhead];
#pragma unroll
            for (int k = 0; k < kLdsAhead; ++k) {
                const int r = max(row - 2 - k, 0);
                gc[k] = gcol[(size_t)r * gst];
                if constexpr (!DL)
                    uc[k] = dget(max(r - 1, 0), c); // the old value above node r
            }
            for (int i0 = row - 2; i0 > 0; i0 -= kLdsAhead) {
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    const int r = max(i0 - kLdsAhead - k, 0);
                    gn[k] = gcol[(size_t)r * gst];
                    if constexpr (!DL)
                        un[k] = dget(max(r - 1, 0), c);
                }
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    i = i0 - k;
                    if (i > 0) {
                        old = up; // d(i, c) before this sweep wrote it
                        up = DL ? dget(i - 1, c) : uc[DL ? 0 : k];
                        Gi = gc[k];
                        sumg = gw * (T)2 + Gi;
                        nv = be * (gw * (up + down) + Gi * su[i * ld + c]) / sumg + omb * old;
                        moved |= dev_abs(nv - old) > thr;
                        dput(i, c, nv);
                        down = nv;
                    }
                }
#pragma unroll
                for (int k = 0; k < kLdsAhead; ++k) {
                    gc[k] = gn[k];
                    if constexpr (!DL)
                        uc[k] = un[k];
                }
            }
            old = up; // d(0, c) before this sweep
            Gi = gcol[0];
            sumg = gw + Gi;
            nv = be * (gw * down + Gi * su[c]) / sumg + omb * old;
            moved |= dev_abs(nv - old) > thr;
            dput(0, c, nv);
        }
        __syncthreads();
    }
    for (size_t k = t; k < cells; k += kLdsThreads) {
        const int i = (int)(k / col), c = (int)(k % col);
        u[k] = su[i * ld + c];
        if constexpr (DL)
            d[k] = sd[i * ld + c];
    }
    if constexpr (nl) {
        // the readout takes the conductance of the last sweep from A.Gc (a
        // call without sweeps reads out with the linear conductance)
        const int S = A.sweeps_this_launch;
        if (S >= 2 && ((S - 2) & 1)) {
            for (size_t k = t; k < cells; k += kLdsThreads)
                Fn[0][k] = Fn[1][k];
        }
    }
    if (A.want_moved && moved)
        atomicOr(A.moved, 1u);
}

// Readout: out[b][x][y*col + c] = sum_i (u - d) G, or the voltage differences;
// OU mode: out[b][x][ou_x][y*col + c0 + c], the current of each OU.
template <typename T>
__global__ void hook_readout(HookArgs<T> A, T *out, int ret_v)
{
    const int xb = blockIdx.x;
    const HookUnit<T> U = hook_unit(A, xb);
    const int b = U.b, x = U.x, y = U.y;
    const int row = U.rows, col = U.cols;
    const size_t cells = (size_t)row * col;
    const T *u = A.u + (size_t)xb * cells;
    const T *d = A.d + (size_t)xb * cells;
    if (ret_v) {
        for (size_t k = threadIdx.x; k < cells; k += blockDim.x)
            out[(size_t)xb * cells + k] = u[k] - d[k];
        return;
    }
    for (int c = threadIdx.x; c < col; c += blockDim.x) {
        T s = (T)0;
        for (int i = 0; i < row; ++i) {
            const size_t k = (size_t)i * col + c;
            const T G = A.nonlinear ? A.Gc[(size_t)xb * cells + k]
                                    : A.G[((size_t)x * A.row + U.r0 + i) * ((size_t)A.numY * A.col) +
                                          (size_t)y * A.col + U.c0 + c];
            s += (u[k] - d[k]) * G;
        }
        out[((((size_t)b * A.numX + x) * A.nOuX + U.oux) * A.numY + y) * A.col + U.c0 + c] = s;
    }
}

__global__ void set_u32(unsigned *p, unsigned v) { *p = v; }
__global__ void set_i32(int *p, int v) { *p = v; }

template <typename T>
int run_hook(const fadesim_hook_desc *D, hipStream_t s)
{
    const bool ou = D->ou_rows > 0;
    const int nOuX = ou ? D->row_size / D->ou_rows : 1, nOuY = ou ? D->col_size / D->ou_cols : 1;
    const int grid = D->batch_size * D->numX * D->numY * nOuX * nOuY;
    const size_t cells = ou ? (size_t)D->ou_rows * D->ou_cols : (size_t)D->row_size * D->col_size;
    const size_t nfield = (size_t)grid * cells;
    // whole tiles up to kLdsMax x kLdsMax: hook_lds_kernel -- one more plane
    // (the transposed conductance), and for sinh cells three more (z and the
    // two transposed conductance buffers); sinh cells in one launch only
    const bool lds = !ou && D->row_size <= kLdsMax && D->col_size <= kLdsMax && !(D->nonlinear && D->enable_break);
    const size_t planes = (D->nonlinear ? 4 : 2) + (lds ? (D->nonlinear ? 4 : 1) : 0);
    char *ws = (char *)workspace(planes * nfield * sizeof(T) + 256);
    if (!ws)
        return FADESIM_E_NOMEM;
    HookArgs<T> A;
    A.input = (const T *)D->d_input;
    A.G = (const T *)D->d_G;
    A.bsize = D->batch_size;
    A.row = D->row_size;
    A.col = D->col_size;
    A.numX = D->numX;
    A.numY = D->numY;
    A.ou_rows = ou ? D->ou_rows : 0;
    A.ou_cols = ou ? D->ou_cols : 0;
    A.nOuX = nOuX;
    A.nOuY = nOuY;
    A.gw = (T)D->g_wire;
    A.gl = (T)(D->g_load_left > 0 ? D->g_load_left : D->g_wire);
    A.gd = (T)(D->g_load_down > 0 ? D->g_load_down : D->g_wire);
    A.be = (T)D->beta;
    A.omb = (T)1 - A.be;
    A.thr = (T)D->break_threshold;
    A.nonlinear = D->nonlinear;
    for (int k = 0; k < 6; ++k)
        A.nl[k] = D->nonlinear_params[k];
    A.u = (T *)ws;
    A.d = A.u + nfield;
    A.Gc = D->nonlinear ? A.d + nfield : nullptr;
    A.Gn = D->nonlinear ? A.Gc + nfield : nullptr;
    A.moved = (unsigned *)(ws + planes * nfield * sizeof(T));
    T *GT = lds ? A.u + (D->nonlinear ? 4 : 2) * nfield : nullptr;
    T *NTp = lds && D->nonlinear ? GT + nfield : nullptr; // two planes
    T *Z = lds && D->nonlinear ? GT + 3 * nfield : nullptr;
    // the workspace kernel: OU blocks a thread per line of the block, in
    // whole waves (more workgroups per CU); whole tiles kHookThreads, whose
    // per-cell passes (sinh cells) gain from every thread
    const int unit_lines = std::max(D->ou_rows, D->ou_cols);
    const int hook_threads = ou ? std::min(kHookThreads, std::max(64, (unit_lines + 63) / 64 * 64)) : kHookThreads;
    // one launch of `sweeps` sweeps
    auto sweep_launch = [&]() {
        if (lds && A.nonlinear)
            hipLaunchKernelGGL((hook_lds_kernel<T, 256, true>), dim3(grid), dim3(256), 0, s, A, GT, NTp, Z);
        else if (lds)
            hipLaunchKernelGGL((hook_lds_kernel<T, 128, false>), dim3(grid), dim3(128), 0, s, A, GT, NTp, Z);
        else
            hipLaunchKernelGGL((hook_kernel<T>), dim3(grid), dim3(hook_threads), 0, s, A);
    };
    int done = 0;
    if (!D->enable_break) {
        A.sweeps_this_launch = D->iter_times;
        A.first_sweep = 1;
        A.want_moved = 0;
        sweep_launch();
        FADESIM_HIP_CHECK(hipGetLastError());
        done = D->iter_times;
    } else {
        // one sweep per launch and a host check of the "moved" flag, as the
        // reference's host loop does (split_arr_matmul.cu:295-306)
        A.sweeps_this_launch = 1;
        A.want_moved = 1;
        for (int it = 0; it < D->iter_times; ++it) {
            A.first_sweep = (it == 0);
            hipLaunchKernelGGL(set_u32, dim3(1), dim3(1), 0, s, A.moved, 0u);
            sweep_launch();
            FADESIM_HIP_CHECK(hipGetLastError());
            unsigned h = 0;
            FADESIM_HIP_CHECK(hipMemcpyAsync(&h, A.moved, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            FADESIM_HIP_CHECK(hipStreamSynchronize(s));
            ++done;
            if (!h)
                break;
        }
    }
    if (done == 0) {
        // no sweep: the fields still need their initial state for the
        // readout, which, as the reference's, uses the linear conductance
        // (sinh cells' conductance is first formed inside a sweep)
        A.nonlinear = 0;
        A.sweeps_this_launch = 0;
        A.first_sweep = 1;
        A.want_moved = 0;
        sweep_launch();
        FADESIM_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL((hook_readout<T>), dim3(grid), dim3(128), 0, s, A, (T *)D->d_out, D->ret_v);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (D->d_sweeps) {
        hipLaunchKernelGGL(set_i32, dim3(1), dim3(1), 0, s, D->d_sweeps, done);
        FADESIM_HIP_CHECK(hipGetLastError());
    }
    return FADESIM_OK;
}

} // namespace
} // namespace fadesim

using namespace fadesim;

extern "C" int fadesim_hook_ir_drop_gs(const fadesim_hook_desc *d, void *stream)
{
    if (!d)
        return set_error(FADESIM_E_INVALID, "null descriptor");
    if (d->row_size < 2)
        return set_error(FADESIM_E_INVALID, "row_size should be >=2");
    if (d->col_size < 2)
        return set_error(FADESIM_E_INVALID, "col_size should be >=2");
    if (d->batch_size < 1 || d->numX < 1 || d->numY < 1 || d->iter_times < 0)
        return set_error(FADESIM_E_INVALID, "batch_size, numX, numY must be >= 1 and iter_times >= 0");
    if (!d->d_input || !d->d_G || !d->d_out)
        return set_error(FADESIM_E_INVALID, "input, G and out are required");
    if (d->ou_rows > 0 || d->ou_cols > 0) {
        if (d->ou_rows < 2 || d->ou_cols < 2 || d->row_size % d->ou_rows || d->col_size % d->ou_cols)
            return set_error(FADESIM_E_INVALID, "OU sizes must be >= 2 and divide row_size and col_size");
        if (d->ret_v)
            return set_error(FADESIM_E_INVALID, "OU mode returns currents only");
    }
    if (d->dtype_bytes == 4)
        return run_hook<float>(d, (hipStream_t)stream);
    if (d->dtype_bytes == 8)
        return run_hook<double>(d, (hipStream_t)stream);
    return set_error(FADESIM_E_INVALID, "dtype_bytes must be 4 (float) or 8 (double)");
}
