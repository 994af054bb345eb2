// fadesim_gs64_solve: one crossbar in FP64 with per-segment wire resistances
// and the reference stopping rule -- the solver behind the `sim` CLI's
// nodebasedGSMethod command.  The reference's lexicographic order is
// gslex.hip; this file holds the entry point and the red-black alternative:
//
// Same network and update as NodebasedGSMethod::nodeBaseGS
// (crossbarHspice/V2/concrete_command.h:1218-1273) in absolute node voltages,
// with wordline segments read from RU (array2D::nodeBasedGS,
// crossbarHspice/2dArray.cc:1562-1588 -- V2 reads them from RDLine, which is
// the same for square arrays of one wire resistance).  Red-black order:
// colour of u_ij is (i+j)&1, of d_ij (i+j+1)&1, colour 0 first.  Node update
// without fused multiply-adds (built with -ffp-contract=off):
//   u: num = (aL*L + aR*R) + g*d,  u' = wD*num + (1-w)*u,  wD = w/((aL+aR)+g)
//   d: num = (bU*U + bD*Dn) + g*u, d' = wD*num + (1-w)*d,  wD = w/((bU+bD)+g)
// Stop after the first sweep whose max |x' - x| < tol (concrete_command.h:1256-1272).
// The readout sums (u-d)*g over ascending rows per column, like the reference.
//
// General boundary (array2D::nodeBasedGeneralGS, crossbarHspice/2dArray.cc:
// 1209-1374): any edge driven (voltage source through its end segment) or
// open, per line; initial state u = v_left on driven rows (0 elsewhere),
// d = 0; readout the sink wire current (d_n-1,j - v_bottom_j) / RD[j][n] of
// driven bottoms, 0 elsewhere.  The same kernel with per-node end terms.
//
// One workgroup; fields and coefficients live in the workspace (L2 resident
// for the array sizes the CLI handles).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "common.h"
#include "fadesim.h"
#include "gslex.h"

namespace fadesim {
namespace {

constexpr int kGsThreads = 1024;

struct GsArgs {
    int n, m;
    const double *g, *ru, *rd;
    int ru_ld, rd_ld;
    const double *vin, *vout;
    int general;                              // general-boundary network (nodeBasedGeneralGS)
    const int *on_top, *on_left, *on_bottom, *on_right;
    const double *vtop, *vright;
    const double *init_u, *init_d;
    int iter_number;
    double w;
    int enable_break;
    double tol;
    double *u, *d, *coef; // workspace: u, d [n*m]; coef [6][n*m]
    double *vup, *vdown, *iout;
    int *record, *sweeps;
};

__global__ void __launch_bounds__(kGsThreads) gs64_kernel(GsArgs A)
{
    const int n = A.n, m = A.m, t = threadIdx.x;
    const size_t nm = (size_t)n * m;
    double *aL = A.coef, *aR = A.coef + nm, *bU = A.coef + 2 * nm, *bD = A.coef + 3 * nm;
    double *wDu = A.coef + 4 * nm, *wDd = A.coef + 5 * nm;
    double *u = A.u, *d = A.d;
    __shared__ double wred[kGsThreads / 64];

    const bool gen = A.general != 0;
    for (size_t k = t; k < nm; k += kGsThreads) {
        const int i = (int)(k / m), j = (int)(k % m);
        const bool onL = gen ? A.on_left[i] != 0 : true, onR = gen ? A.on_right[i] != 0 : false;
        const bool onT = gen ? A.on_top[j] != 0 : false, onB = gen ? A.on_bottom[j] != 0 : true;
        aL[k] = (j > 0 || onL) ? 1.0 / A.ru[(size_t)i * A.ru_ld + j] : 0.0;
        aR[k] = (j < m - 1 || onR) ? 1.0 / A.ru[(size_t)i * A.ru_ld + j + 1] : 0.0;
        bU[k] = (i > 0 || onT) ? 1.0 / A.rd[(size_t)j * A.rd_ld + i] : 0.0;
        bD[k] = (i < n - 1 || onB) ? 1.0 / A.rd[(size_t)j * A.rd_ld + i + 1] : 0.0;
        wDu[k] = A.w / ((aL[k] + aR[k]) + A.g[k]);
        wDd[k] = A.w / ((bU[k] + bD[k]) + A.g[k]);
        if (A.init_u) {
            u[k] = A.init_u[k];
            d[k] = A.init_d[k];
        } else {
            u[k] = gen ? (onL ? A.vin[i] : 0.0) : A.vin[i];
            d[k] = gen ? 0.0 : A.vout[j];
        }
    }
    __syncthreads();
    const double omw = 1.0 - A.w;
    int record = -1, it;
    for (it = 0; it < A.iter_number; ++it) {
        double mx = 0.0;
        for (int c = 0; c < 2; ++c) {
            for (size_t k = t; k < nm; k += kGsThreads) {
                const int i = (int)(k / m), j = (int)(k % m);
                if (((i + j) & 1) == c) {
                    const double L = j ? u[k - 1] : A.vin[i];
                    const double R = (j < m - 1) ? u[k + 1] : (gen ? A.vright[i] : 0.0);
                    const double num = (aL[k] * L + aR[k] * R) + A.g[k] * d[k];
                    const double nv = wDu[k] * num + omw * u[k];
                    mx = fmax(mx, fabs(nv - u[k]));
                    u[k] = nv;
                } else {
                    const double U = i ? d[k - m] : (gen ? A.vtop[j] : 0.0);
                    const double Dn = (i < n - 1) ? d[k + m] : A.vout[j];
                    const double num = (bU[k] * U + bD[k] * Dn) + A.g[k] * u[k];
                    const double nv = wDd[k] * num + omw * d[k];
                    mx = fmax(mx, fabs(nv - d[k]));
                    d[k] = nv;
                }
            }
            __syncthreads();
        }
        if (A.enable_break) {
            for (int off = 32; off > 0; off >>= 1)
                mx = fmax(mx, __shfl_xor(mx, off, 64));
            if ((t & 63) == 0)
                wred[t >> 6] = mx;
            __syncthreads();
            double all = 0.0;
            for (int w = 0; w < kGsThreads / 64; ++w)
                all = fmax(all, wred[w]);
            __syncthreads(); // wred is rewritten next sweep
            if (all < A.tol) {
                record = it;
                ++it;
                break;
            }
        }
    }
    for (size_t k = t; k < nm; k += kGsThreads) {
        A.vup[k] = u[k];
        A.vdown[k] = d[k];
    }
    for (int j = t; j < m; j += kGsThreads) {
        if (gen) { // the sink wire current of a driven bottom (2dArray.cc:1351-1356)
            const size_t k = (size_t)(n - 1) * m + j;
            A.iout[j] = A.on_bottom[j] ? (d[k] - A.vout[j]) / A.rd[(size_t)j * A.rd_ld + n] : 0.0;
            continue;
        }
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const size_t k = (size_t)i * m + j;
            s += (u[k] - d[k]) * A.g[k];
        }
        A.iout[j] = s;
    }
    if (t == 0) {
        *A.record = record;
        *A.sweeps = it;
    }
}

} // namespace
} // namespace fadesim

using namespace fadesim;

extern "C" int fadesim_gs64_solve(const fadesim_gs64_desc *d, void *stream)
{
    if (!d)
        return set_error(FADESIM_E_INVALID, "null descriptor");
    if (d->struct_size != sizeof(fadesim_gs64_desc))
        return set_error(FADESIM_E_INVALID, "fadesim_gs64_desc.struct_size is " + std::to_string(d->struct_size) +
                                                ", the library's descriptor has " +
                                                std::to_string(sizeof(fadesim_gs64_desc)) + " bytes");
    if (d->order != FADESIM_GS64_LEXICOGRAPHIC && d->order != FADESIM_GS64_REDBLACK)
        return set_error(FADESIM_E_INVALID, "order must be FADESIM_GS64_LEXICOGRAPHIC or FADESIM_GS64_REDBLACK");
    if ((d->d_init_up == nullptr) != (d->d_init_down == nullptr))
        return set_error(FADESIM_E_INVALID, "d_init_up and d_init_down are given together or not at all");
    if (!(d->rows >= 2 && d->cols >= 2))
        return set_error(FADESIM_E_INVALID, "m and n should greater than 2");
    if (!d->d_g || !d->d_ru || !d->d_rd || !d->d_vin || !d->d_vout || !d->d_vup || !d->d_vdown ||
        !d->d_iout || !d->d_record || !d->d_sweeps)
        return set_error(FADESIM_E_INVALID, "all device pointers are required");
    if (d->ru_ld < d->cols + 1 || d->rd_ld < d->rows + 1)
        return set_error(FADESIM_E_INVALID, "ru_ld must be >= cols+1 and rd_ld >= rows+1");
    if (d->iter_number < 0)
        return set_error(FADESIM_E_INVALID, "iter_number must be >= 0");
    if (d->general && (!d->d_on_top || !d->d_on_left || !d->d_on_bottom || !d->d_on_right || !d->d_v_top ||
                       !d->d_v_right))
        return set_error(FADESIM_E_INVALID, "the general boundary needs all edge flags and voltages");
    const size_t nm = (size_t)d->rows * d->cols;
    if (d->order == FADESIM_GS64_LEXICOGRAPHIC) {
        void *lws = workspace(gslex_workspace_bytes(d->rows, d->cols));
        if (!lws)
            return FADESIM_E_NOMEM;
        return gslex_solve(d, lws, (hipStream_t)stream);
    }
    double *ws = (double *)workspace(8 * nm * sizeof(double));
    if (!ws)
        return FADESIM_E_NOMEM;
    GsArgs a;
    a.n = d->rows;
    a.m = d->cols;
    a.g = d->d_g;
    a.ru = d->d_ru;
    a.rd = d->d_rd;
    a.ru_ld = d->ru_ld;
    a.rd_ld = d->rd_ld;
    a.vin = d->d_vin;
    a.vout = d->d_vout;
    a.general = d->general;
    a.on_top = d->d_on_top;
    a.on_left = d->d_on_left;
    a.on_bottom = d->d_on_bottom;
    a.on_right = d->d_on_right;
    a.vtop = d->d_v_top;
    a.vright = d->d_v_right;
    a.init_u = d->d_init_up;
    a.init_d = d->d_init_down;
    a.iter_number = d->iter_number;
    a.w = d->omega;
    a.enable_break = d->enable_break;
    a.tol = d->tol;
    a.u = ws;
    a.d = ws + nm;
    a.coef = ws + 2 * nm;
    a.vup = d->d_vup;
    a.vdown = d->d_vdown;
    a.iout = d->d_iout;
    a.record = d->d_record;
    a.sweeps = d->d_sweeps;
    hipLaunchKernelGGL(gs64_kernel, dim3(1), dim3(kGsThreads), 0, (hipStream_t)stream, a);
    FADESIM_HIP_CHECK(hipGetLastError());
    return FADESIM_OK;
}
