// Nonlinear (sinh) cell conductance of a crossbar state, for the Picard
// iteration of the batched solver.
//
// Cell model of pimtorch/.../SplitArrCuda/split_arr_matmul.cu:142-196 (and
// conductance_variation.py:29-40): for the voltage V across the cell,
//   |V| clamped to >= 1e-5,  I_on = b_on sinh(a_on |V|),  I_off = b_off sinh(a_off |V|)
//   R(V) = (1 - z) |V| / I_on + z |V| / I_off,   z = max(0, (1/g_lin - Ron) / (Roff - Ron))
// and the cell conductance is 1 / R(V).  Evaluated in FP64, stored FP32 (the
// batched solver's conductance type).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "common.h"
#include "fadesim.h"

namespace fadesim {
namespace {

struct CellArgs {
    int batch, rows, cols;
    const float *glin;
    const int *gidx;
    const double *vin, *vout;
    const float *field;
    double Ron, Roff, a_on, b_on, a_off, b_off;
    float *g;
};

__global__ void __launch_bounds__(256) cell_conductance_kernel(CellArgs A)
{
    const size_t MN = (size_t)A.rows * A.cols;
    const size_t total = (size_t)A.batch * MN;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
        const size_t b = k / MN, id = k % MN;
        const int i = (int)(id / A.cols), j = (int)(id % A.cols);
        const size_t slot = A.gidx ? (size_t)A.gidx[b] : b;
        const double gl = (double)A.glin[slot * MN + id];
        double V = A.vin[b * A.rows + i] - (A.vout ? A.vout[b * A.cols + j] : 0.0);
        if (A.field)
            V -= (double)A.field[b * 2 * MN + id] + (double)A.field[b * 2 * MN + MN + id];
        if (V < 1e-5 && V > -1e-5)
            V = V > 0.0 ? 1e-5 : -1e-5;
        V = fabs(V);
        double z = (1.0 / gl - A.Ron) / (A.Roff - A.Ron);
        if (z < 0.0)
            z = 0.0;
        const double R = (1.0 - z) * V / (A.b_on * sinh(A.a_on * V)) + z * V / (A.b_off * sinh(A.a_off * V));
        A.g[k] = (float)(1.0 / R);
    }
}

} // namespace
} // namespace fadesim

using namespace fadesim;

extern "C" int fadesim_cell_conductance(int batch, int rows, int cols, const float *d_glin, const int *d_gidx,
                                        const double *d_vin, const double *d_vout, const float *d_field,
                                        const double *params, float *d_g, void *stream)
{
    if (batch < 0 || rows < 1 || cols < 1 || !d_glin || !d_vin || !params || !d_g)
        return set_error(FADESIM_E_INVALID, "fadesim_cell_conductance: bad arguments");
    if (!(params[1] > params[0]) || !(params[3] > 0.0) || !(params[5] > 0.0))
        return set_error(FADESIM_E_INVALID, "fadesim_cell_conductance: need Roff > Ron and b_on, b_off > 0");
    if (batch == 0)
        return FADESIM_OK;
    CellArgs a{batch, rows, cols, d_glin, d_gidx, d_vin, d_vout, d_field,
               params[0], params[1], params[2], params[3], params[4], params[5], d_g};
    const size_t total = (size_t)batch * rows * cols;
    size_t grid = (total + 255) / 256;
    if (grid > 65536)
        grid = 65536;
    hipLaunchKernelGGL(cell_conductance_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, a);
    FADESIM_HIP_CHECK(hipGetLastError());
    return FADESIM_OK;
}
