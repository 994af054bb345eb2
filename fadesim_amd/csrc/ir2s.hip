// IR2S, the reference's wire-solve iteration, with its own semantics for the
// callers that depend on its (often unconverged) results: the PimTorch hook
// function ir2s (pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/
// split_arr_matmul.cu:336-545) and the V1 CLI command `fastsolve`
// (crossbarHspice/2dArray.cc:318-327, 1691-1900 over
// crossbarHspice/ir_drop_fast_solver.h:24-163).
//
// One iteration: the cell currents I = (u - d) G of the present state, then
// every wordline and every bitline solved exactly for them -- wordline i
// carries in segment j the currents of columns >= j, bitline j carries below
// row i the currents of rows <= i:
//   u_i0 = vin_i - r_src S_i0,   u_ij = u_i,j-1 - r S_ij        (S: suffix sums)
//   d_n-1,j = r_snk T_n-1,j + vout_j,  d_ij = d_i+1,j + r T_ij   (T: prefix sums)
// Both layers use the currents of the same state (Jacobi between layers).
//
// OU mode (ir2s_ou, split_arr_matmul.cu:816-1085): every ou_rows x ou_cols
// block of a tile is solved as its own crossbar, its source c0 + 1 wire
// segments away and its sink row - (last row) segments away; separate source
// and sink loads on whole tiles are ir2s_ou_local (:546-688).
//
// One workgroup per crossbar.  The currents of an iteration go to a workspace
// plane first (all cells in parallel); a thread per wordline then walks its
// row and a thread per bitline its column, each with the reference's
// accumulation order, so the FP arithmetic is the reference's operation for
// operation (-ffp-contract=off): the CUDA hook multiplies by r and carries
// each line's rounding residue of the running sum into the next iteration
// (its I_row_sum / I_col_sum are never reset), the CPU solver divides by
// g_wire = 1/r and resets them.  Sinh-nonlinear cells: iteration k uses G of
// the state two iterations back (the buffer the reference is about to
// overwrite, split_arr_matmul.cu:476-494), so each iteration computes the next
// one's G from the present state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "common.h"
#include "fadesim.h"

namespace fadesim {
namespace {

constexpr int kIr2sThreads = 256;

template <typename T>
struct Ir2sArgs {
    const T *vin, *G, *vout;
    int layout; // FADESIM_IR2S_HOOK or FADESIM_IR2S_PLAIN
    int bsize, row, col, numX, numY;
    int ou_rows, ou_cols, nOuX, nOuY; // OU mode (hook layout): blocks solved as crossbars
    int iters_this_launch;
    int first; // initialise the state (u = vin, d = 0)
    T r, rsrc, rsnk, gwire;
    int divide_by_g, carry_residue;
    int brk; // 0 none, FADESIM_IR2S_BREAK_BATCH, FADESIM_IR2S_BREAK_EACH
    T thr;
    int nonlinear;
    double nl[6];
    T *u, *d, *cur;      // [grid][row][col]
    T *wres, *bres;      // carried residues [grid][row], [grid][col]
    T *Gc, *Gn;          // nonlinear conductances of this and the next iteration
    unsigned *moved;     // batch break: any |dv| > thr in this launch
    int *count;          // per-crossbar break: iteration count (reference convention)
};

template <typename T> __device__ __forceinline__ T dsinh(T x);
template <> __device__ __forceinline__ float dsinh<float>(float x) { return sinhf(x); }
template <> __device__ __forceinline__ double dsinh<double>(double x) { return sinh(x); }
template <typename T> __device__ __forceinline__ T dabs(T x) { return x < (T)0 ? -x : x; }

// Nonlinear cell conductance for voltage V (split_arr_matmul.cu:142-196).
template <typename T>
__device__ __forceinline__ T cell_g(T V, T z, const double *nl)
{
    if (V < (T)1e-5 && V > -(T)1e-5)
        V = V > (T)0 ? (T)1e-5 : -(T)1e-5;
    V = dabs(V);
    const T R = ((T)1 - z) * V / ((T)nl[3] * dsinh<T>(V * (T)nl[2])) + z * V / ((T)nl[5] * dsinh<T>(V * (T)nl[4]));
    return (T)1 / R;
}

template <typename T>
struct Cross {
    const T *in;   // [rows] drives
    const T *vo;   // [cols] sink voltages or null
    size_t gbase;  // G of cell (i, c) at G[gbase + i * gld + c]
    size_t gld;
    int rows, cols;
    T rsrc, rsnk;  // source / sink segment resistance
    size_t obase;  // readout: out[obase + c]
};

template <typename T>
__device__ __forceinline__ Cross<T> cross_of(const Ir2sArgs<T> &A, int xb)
{
    Cross<T> X;
    X.rows = A.row;
    X.cols = A.col;
    X.rsrc = A.rsrc;
    X.rsnk = A.rsnk;
    if (A.layout == FADESIM_IR2S_HOOK) {
        const int ouy = xb % A.nOuY, oux = (xb / A.nOuY) % A.nOuX, t2 = xb / (A.nOuY * A.nOuX);
        const int b = t2 / (A.numX * A.numY), x = (t2 / A.numY) % A.numX, y = t2 % A.numY;
        int r0 = 0, c0 = 0;
        if (A.ou_rows > 0) {
            // the OU's source is c0 + 1 wire segments away, its sink
            // row - (last row) segments (split_arr_matmul.cu:846, 891)
            X.rows = A.ou_rows;
            X.cols = A.ou_cols;
            r0 = oux * A.ou_rows;
            c0 = ouy * A.ou_cols;
            X.rsrc = A.r * (T)(c0 + 1);
            X.rsnk = A.r * (T)(A.row - (r0 + A.ou_rows - 1));
        }
        X.in = A.vin + ((size_t)b * A.numX + x) * A.row + r0;
        X.gbase = ((size_t)x * A.row + r0) * ((size_t)A.numY * A.col) + (size_t)y * A.col + c0;
        X.gld = (size_t)A.numY * A.col;
        X.obase = ((((size_t)b * A.numX + x) * A.nOuX + oux) * A.numY + y) * A.col + c0;
    } else {
        X.in = A.vin + (size_t)xb * A.row;
        X.gbase = (size_t)xb * A.row * A.col;
        X.gld = A.col;
        X.obase = (size_t)xb * A.col;
    }
    X.vo = A.vout ? A.vout + (size_t)xb * A.col : nullptr;
    return X;
}

template <typename T>
__global__ void __launch_bounds__(kIr2sThreads) ir2s_kernel(Ir2sArgs<T> A)
{
    const int xb = blockIdx.x;
    const Cross<T> X = cross_of(A, xb);
    const int row = X.rows, col = X.cols, t = threadIdx.x;
    const int nth = blockDim.x; // a multiple of 64 up to kIr2sThreads, sized to the crossbar's lines
    const size_t cells = (size_t)row * col;
    T *u = A.u + (size_t)xb * cells, *d = A.d + (size_t)xb * cells, *cur = A.cur + (size_t)xb * cells;
    T *wres = A.wres + (size_t)xb * row, *bres = A.bres + (size_t)xb * col;
    T *Gc = A.nonlinear ? A.Gc + (size_t)xb * cells : nullptr;
    T *Gn = A.nonlinear ? A.Gn + (size_t)xb * cells : nullptr;
    auto Glin = [&](size_t k) -> T { return A.G[X.gbase + (k / col) * X.gld + (k % col)]; };
    auto zof = [&](size_t k) -> T {
        T z = ((T)1 / Glin(k) - (T)A.nl[0]) / (T)(A.nl[1] - A.nl[0]);
        return z < (T)0 ? (T)0 : z;
    };
    __shared__ int s_stop;
    __shared__ unsigned s_moved;

    if (A.first) {
        for (size_t k = t; k < cells; k += nth) {
            u[k] = X.in[k / col];
            d[k] = (T)0;
        }
        for (int i = t; i < row; i += nth)
            wres[i] = (T)0;
        for (int c = t; c < col; c += nth)
            bres[c] = (T)0;
        __syncthreads();
        if (A.nonlinear) {
            for (size_t k = t; k < cells; k += nth)
                Gc[k] = Gn[k] = cell_g<T>(u[k] - d[k], zof(k), A.nl);
            __syncthreads();
        }
        if (t == 0 && A.count)
            A.count[xb] = 0;
    }
    if (t == 0) {
        s_stop = 0;
        s_moved = 0u;
    }
    __syncthreads();
    for (int it = 0; it < A.iters_this_launch; ++it) {
        if (A.nonlinear && !(A.first && it == 0)) {
            for (size_t k = t; k < cells; k += nth) {
                Gc[k] = Gn[k];
                Gn[k] = cell_g<T>(u[k] - d[k], zof(k), A.nl);
            }
            __syncthreads();
        }
        for (size_t k = t; k < cells; k += nth)
            cur[k] = (u[k] - d[k]) * (A.nonlinear ? Gc[k] : Glin(k));
        __syncthreads();
        T dmax = (T)0;
        // wordlines: segment j carries the currents of columns >= j
        for (int i = t; i < row; i += nth) {
            const T *ci = cur + (size_t)i * col;
            T *ui = u + (size_t)i * col;
            T s = A.carry_residue ? wres[i] : (T)0;
            for (int j = 0; j < col; ++j)
                s += ci[j];
            T v = X.in[i] - s * X.rsrc;
            dmax = fmax(dmax, dabs(v - ui[0]));
            ui[0] = v;
            s -= ci[0];
            for (int j = 1; j < col; ++j) {
                v = ui[j - 1] - (A.divide_by_g ? s / A.gwire : s * A.r);
                dmax = fmax(dmax, dabs(v - ui[j]));
                ui[j] = v;
                s -= ci[j];
            }
            if (A.carry_residue)
                wres[i] = s;
        }
        // bitlines: the segment below row i carries the currents of rows <= i
        for (int c = t; c < col; c += nth) {
            T s = A.carry_residue ? bres[c] : (T)0;
            for (int i = 0; i < row; ++i)
                s += cur[(size_t)i * col + c];
            size_t k = (size_t)(row - 1) * col + c;
            T v = s * X.rsnk + (X.vo ? X.vo[c] : (T)0);
            dmax = fmax(dmax, dabs(v - d[k]));
            d[k] = v;
            s -= cur[k];
            for (int i = row - 2; i >= 0; --i) {
                k = (size_t)i * col + c;
                v = d[k + col] + (A.divide_by_g ? s / A.gwire : s * A.r);
                dmax = fmax(dmax, dabs(v - d[k]));
                d[k] = v;
                s -= cur[k];
            }
            if (A.carry_residue)
                bres[c] = s;
        }
        if (A.brk == FADESIM_IR2S_BREAK_BATCH) {
            if (dmax > A.thr)
                s_moved = 1u; // benign race: every writer stores 1
        } else if (A.brk == FADESIM_IR2S_BREAK_EACH) {
            if (!(dmax < A.thr))
                s_moved = 1u;
        }
        __syncthreads();
        if (A.brk == FADESIM_IR2S_BREAK_EACH) {
            const bool stop = s_moved == 0u;
            if (t == 0 && A.count)
                A.count[xb] += 1;
            __syncthreads();
            if (t == 0)
                s_moved = 0u;
            __syncthreads();
            if (stop) {
                if (t == 0)
                    s_stop = 1;
                break;
            }
        }
    }
    __syncthreads();
    if (A.brk == FADESIM_IR2S_BREAK_BATCH && t == 0 && s_moved)
        atomicOr(A.moved, 1u);
    if (A.brk == FADESIM_IR2S_BREAK_EACH && t == 0 && A.count && !s_stop)
        A.count[xb] += 1; // the reference's count when the threshold was never met: times + 1
}

// Readout: currents sum_i (u - d) G per column, or the voltage differences;
// optionally the node voltages.
template <typename T>
__global__ void ir2s_readout(Ir2sArgs<T> A, T *out, int ret_v, T *vup, T *vdown)
{
    const int xb = blockIdx.x;
    const Cross<T> X = cross_of(A, xb);
    const int row = X.rows, col = X.cols;
    const size_t cells = (size_t)row * col;
    const T *u = A.u + (size_t)xb * cells, *d = A.d + (size_t)xb * cells;
    for (size_t k = threadIdx.x; k < cells; k += blockDim.x) {
        if (ret_v)
            out[(size_t)xb * cells + k] = u[k] - d[k];
        if (vup)
            vup[(size_t)xb * cells + k] = u[k];
        if (vdown)
            vdown[(size_t)xb * cells + k] = d[k];
    }
    if (ret_v)
        return;
    for (int c = threadIdx.x; c < col; c += blockDim.x) {
        T s = (T)0;
        for (int i = 0; i < row; ++i) {
            const size_t k = (size_t)i * col + c;
            const T G = A.nonlinear ? A.Gc[(size_t)xb * cells + k] : A.G[X.gbase + (size_t)i * X.gld + c];
            s += (u[k] - d[k]) * G;
        }
        out[X.obase + c] = s;
    }
}

__global__ void ir2s_set_u32(unsigned *p, unsigned v) { *p = v; }
__global__ void ir2s_set_i32(int *p, int v) { *p = v; }

template <typename T>
int run_ir2s(const fadesim_ir2s_desc *D, hipStream_t s)
{
    const bool ou = D->ou_rows > 0;
    const int nOuX = ou ? D->row_size / D->ou_rows : 1, nOuY = ou ? D->col_size / D->ou_cols : 1;
    const int grid = D->batch_size * D->numX * D->numY * nOuX * nOuY;
    // OU blocks: a thread per line of the block, in whole waves (more
    // workgroups per CU); whole tiles: kIr2sThreads, whose per-cell passes
    // gain from every thread (measured)
    const int ulines = std::max(D->ou_rows, D->ou_cols);
    const int nthreads = ou ? std::min(kIr2sThreads, std::max(64, (ulines + 63) / 64 * 64)) : kIr2sThreads;
    const int urows = ou ? D->ou_rows : D->row_size, ucols = ou ? D->ou_cols : D->col_size;
    const size_t cells = (size_t)urows * ucols;
    const size_t nfield = (size_t)grid * cells;
    const size_t planes = D->nonlinear ? 5 : 3;
    const size_t lines = (size_t)grid * (urows + ucols);
    const size_t bytes = (planes * nfield + lines) * sizeof(T) + (size_t)grid * sizeof(int) + 256;
    char *ws = (char *)workspace(bytes);
    if (!ws)
        return FADESIM_E_NOMEM;
    Ir2sArgs<T> A;
    A.vin = (const T *)D->d_vin;
    A.G = (const T *)D->d_G;
    A.vout = (const T *)D->d_vout;
    A.layout = D->layout;
    A.bsize = D->batch_size;
    A.row = D->row_size;
    A.col = D->col_size;
    A.numX = D->numX;
    A.numY = D->numY;
    A.ou_rows = ou ? D->ou_rows : 0;
    A.ou_cols = ou ? D->ou_cols : 0;
    A.nOuX = nOuX;
    A.nOuY = nOuY;
    A.r = (T)D->r_wire;
    A.gwire = (T)(1.0 / D->r_wire);
    A.rsrc = (T)(D->r_source > 0 ? D->r_source : D->r_wire);
    A.rsnk = (T)(D->r_sink > 0 ? D->r_sink : D->r_wire);
    A.divide_by_g = D->divide_by_g;
    A.carry_residue = D->carry_residue;
    A.brk = D->enable_break;
    A.thr = (T)D->break_threshold;
    A.nonlinear = D->nonlinear;
    for (int k = 0; k < 6; ++k)
        A.nl[k] = D->nonlinear_params[k];
    A.u = (T *)ws;
    A.d = A.u + nfield;
    A.cur = A.d + nfield;
    A.Gc = D->nonlinear ? A.cur + nfield : nullptr;
    A.Gn = D->nonlinear ? A.Gc + nfield : nullptr;
    A.wres = A.u + planes * nfield;
    A.bres = A.wres + (size_t)grid * urows;
    A.count = (int *)(A.bres + (size_t)grid * ucols);
    A.moved = (unsigned *)(A.count + grid);
    int done = 0;
    if (D->enable_break != FADESIM_IR2S_BREAK_BATCH) {
        A.iters_this_launch = D->iter_times;
        A.first = 1;
        hipLaunchKernelGGL((ir2s_kernel<T>), dim3(grid), dim3(nthreads), 0, s, A);
        FADESIM_HIP_CHECK(hipGetLastError());
        done = D->iter_times;
    } else {
        // one iteration per launch and a host check of the batch-wide
        // "moved" flag, as the reference's host loop (split_arr_matmul.cu:517-527)
        A.iters_this_launch = 1;
        for (int it = 0; it < D->iter_times; ++it) {
            A.first = (it == 0);
            hipLaunchKernelGGL(ir2s_set_u32, dim3(1), dim3(1), 0, s, A.moved, 0u);
            hipLaunchKernelGGL((ir2s_kernel<T>), dim3(grid), dim3(nthreads), 0, s, A);
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
        A.iters_this_launch = 0;
        A.first = 1;
        hipLaunchKernelGGL((ir2s_kernel<T>), dim3(grid), dim3(nthreads), 0, s, A);
        FADESIM_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL((ir2s_readout<T>), dim3(grid), dim3(128), 0, s, A, (T *)D->d_out, D->ret_v, (T *)D->d_vup,
                       (T *)D->d_vdown);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (D->d_iters) {
        if (D->enable_break == FADESIM_IR2S_BREAK_EACH) {
            FADESIM_HIP_CHECK(hipMemcpyAsync(D->d_iters, A.count, (size_t)grid * sizeof(int),
                                             hipMemcpyDeviceToDevice, s));
        } else { // one count for the whole call
            hipLaunchKernelGGL(ir2s_set_i32, dim3(1), dim3(1), 0, s, D->d_iters, done);
            FADESIM_HIP_CHECK(hipGetLastError());
        }
    }
    return FADESIM_OK;
}

} // namespace
} // namespace fadesim

using namespace fadesim;

extern "C" int fadesim_ir2s(const fadesim_ir2s_desc *d, void *stream)
{
    if (!d)
        return set_error(FADESIM_E_INVALID, "null descriptor");
    if (d->row_size < 2)
        return set_error(FADESIM_E_INVALID, "row_size should be >=2");
    if (d->col_size < 2)
        return set_error(FADESIM_E_INVALID, "col_size should be >=2");
    if (d->batch_size < 1 || d->numX < 1 || d->numY < 1 || d->iter_times < 0)
        return set_error(FADESIM_E_INVALID, "batch_size, numX, numY must be >= 1 and iter_times >= 0");
    if (d->layout != FADESIM_IR2S_HOOK && d->layout != FADESIM_IR2S_PLAIN)
        return set_error(FADESIM_E_INVALID, "layout must be FADESIM_IR2S_HOOK or FADESIM_IR2S_PLAIN");
    if (d->layout == FADESIM_IR2S_PLAIN && (d->numX != 1 || d->numY != 1))
        return set_error(FADESIM_E_INVALID, "the plain layout takes numX = numY = 1");
    if (d->enable_break < 0 || d->enable_break > FADESIM_IR2S_BREAK_EACH)
        return set_error(FADESIM_E_INVALID, "enable_break must be 0, FADESIM_IR2S_BREAK_BATCH or _EACH");
    if (!(d->r_wire > 0.0))
        return set_error(FADESIM_E_INVALID, "r_wire must be > 0");
    if (d->ou_rows > 0 || d->ou_cols > 0) {
        if (d->layout != FADESIM_IR2S_HOOK || d->ou_rows < 2 || d->ou_cols < 2 || d->row_size % d->ou_rows ||
            d->col_size % d->ou_cols)
            return set_error(FADESIM_E_INVALID,
                             "OU mode: hook layout, OU sizes >= 2 dividing row_size and col_size");
        if (d->ret_v || d->d_vup || d->d_vdown)
            return set_error(FADESIM_E_INVALID, "OU mode returns currents only");
    }
    if (!d->d_vin || !d->d_G || !d->d_out)
        return set_error(FADESIM_E_INVALID, "vin, G and out are required");
    if (d->dtype_bytes == 4)
        return run_ir2s<float>(d, (hipStream_t)stream);
    if (d->dtype_bytes == 8)
        return run_ir2s<double>(d, (hipStream_t)stream);
    return set_error(FADESIM_E_INVALID, "dtype_bytes must be 4 (float) or 8 (double)");
}
