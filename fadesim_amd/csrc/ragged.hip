// Mixed-size batched IR-drop solve in one persistent launch: crossbars of
// 64x64, 128x128, 256x256 and 512x512 cells in any order and mix.
//
// Same red-black SOR core as the uniform solver (rb_core.h).  A 64x64 or
// 128x128 crossbar is one workgroup's work item.  A 256x256 or 512x512 one
// holds 2-8x more state than a CU's registers and LDS, so it is split into
// strips of full rows, 64x256 (4 strips) or 32x512 (16 strips), each one
// workgroup of the same 16384-cell register tile as the 128x128 path; the
// strips of a crossbar exchange their boundary bitline rows after every
// half-sweep (StripCtx).
//
// Schedule: the grid is the resident workgroup count, cut into teams of 16
// consecutive workgroups.  Phase 1: each team takes "big tasks" from a queue
// -- one 512x512 crossbar (16 strips) or up to four 256x256 crossbars (4
// strips each) -- its rank-0 workgroup draws the task and broadcasts it to
// the team, so the strips of a crossbar always run together on co-resident
// workgroups.  Phase 2: when the big queue is empty, every workgroup takes
// single-workgroup items (128x128 first, then 64x64) from a second queue.
// The plan orders the queues longest first.
//
// All waits are bounded (rb::wait_geq): a team member that times out raises
// the abort flag, every waiter sees it, and the kernel drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "fadesim.h"
#include "rb_core.h"

namespace fadesim {
namespace {

using G64 = rb::Geo<2, 4, 32, 16>;  // 64 x 64, 512 threads
using G128 = rb::Geo<4, 8, 32, 16>; // 128 x 128
using G256 = rb::Geo<4, 8, 16, 32>; // 64 x 256 strip, 4 per crossbar
using G512 = rb::Geo<4, 8, 8, 64>;  // 32 x 512 strip, 16 per crossbar
constexpr int kThreads = 512;
constexpr int kTeam = 16;
constexpr int kNmax = 512;
constexpr int kSlotCells = 16384; // cells of one workgroup's tile (all classes)
constexpr int kTaskInts = 5;      // big task: class, up to four instances
constexpr int kHaloMax = std::max(std::max(G64::HALO, G128::HALO), std::max(G256::HALO, G512::HALO));
constexpr int kCbufMax = 2 * std::max(std::max(G64::TY * G64::N, G128::TY * G128::N),
                                      std::max(G256::TY * G256::N, G512::TY * G512::N));
static_assert(G64::NT == kThreads && G128::NT == kThreads && G256::NT == kThreads && G512::NT == kThreads,
              "one block size for every class");
static_assert(G128::M * G128::N == kSlotCells && G256::M * G256::N == kSlotCells &&
                  G512::M * G512::N == kSlotCells,
              "workspace slot");
static_assert(G256::M * 4 == 256 && G512::M * 16 == 512, "strip counts");

// per-team exchange buffers, for each of two consecutive tasks
constexpr size_t kRowsPerTeamTask = (size_t)2 * kTeam * 2 * rb::kRowsLd;   // floats
constexpr size_t kPartPerTeamTask = (size_t)2 * kTeam * 2 * rb::kRowsLd;   // doubles
constexpr size_t kSetupPerTeamTask = (size_t)kTeam * 4 * rb::kRowsLd;      // doubles

struct RaggedArgs {
    const int *cls;           // [K] 0: 64, 1: 128, 2: 256, 3: 512
    const float *g;
    const long long *g_off;   // [K] element offset of instance k (rows*cols cells)
    const double *rw;         // [K]
    const double *vin;
    const long long *vin_off; // [K]
    const double *vout;       // packed like iout, or null
    const long long *out_off; // [K]
    const float *warm;        // packed [2][rows][cols] per instance at 2*g_off, or null
    float *field;
    double *iout;
    int *sweeps, *status;
    double *est;
    const double *zeros;      // kNmax zeros (vout when not given)
    const int *big;           // [n_big][kTaskInts]
    int n_big;
    const int *small;         // [n_small] instance ids
    int n_small;
    int *ctr;                 // [0] big queue, [1] small queue, [2] abort
    int *team_ver;            // [teams]
    int *team_task;           // [teams][n_big + 1]: the task of each broadcast version
    int *flags;               // [n_big][kTeam][2]
    int *cnts;                // [n_big][4]
    float *rows;              // [teams][2][kRowsPerTeamTask]
    double *part;             // [teams][2][kPartPerTeamTask]
    double *setup;            // [teams][2][kSetupPerTeamTask]
    float *ws;                // [grid][7][kSlotCells]
    rb::Ctl ctl[4];
};

__device__ __forceinline__ rb::Inst make_inst(const RaggedArgs &A, int k, int n)
{
    rb::Inst I;
    const long long go = A.g_off[k];
    I.g = A.g + go;
    I.rwire = A.rw[k];
    I.vin = A.vin + A.vin_off[k];
    I.vout = A.vout ? A.vout + A.out_off[k] : A.zeros;
    I.warm = A.warm ? A.warm + 2 * go : nullptr;
    I.field = A.field ? A.field + 2 * go : nullptr;
    I.iout = A.iout + A.out_off[k];
    I.sweeps = A.sweeps ? A.sweeps + k : nullptr;
    I.status = A.status ? A.status + k : nullptr;
    I.est = A.est ? A.est + k : nullptr;
    I.rejected = nullptr;
    I.rows = n;
    I.coarse = nullptr;
    I.cscale = 0.f;
    I.cch = I.clevel = 0;
    return I;
}

__global__ void __launch_bounds__(kThreads) ragged_kernel(RaggedArgs A)
{
    __shared__ __align__(16) float halo[kHaloMax];
    __shared__ double S1[kNmax];
    __shared__ double Ifloor[kNmax];
    __shared__ double colI[kNmax];
    __shared__ double colD[kNmax];
    __shared__ float wmax[kThreads / 64];
    __shared__ int s_task, s_dead;
    __shared__ double s_gsum;
    __shared__ double cbuf[kCbufMax];
    __shared__ double cext[3 * kNmax];
    rb::Shared sh{halo, S1, Ifloor, colI, colD, wmax, &s_dead, cbuf, cext, &s_gsum};
    const rb::StripCtx none{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float *wsb = A.ws + (size_t)blockIdx.x * rb::kWsPlanes * kSlotCells;
    const int team = blockIdx.x / kTeam, rank = blockIdx.x % kTeam;
    int *abort = A.ctr + 2;

    // ---- phase 1: teams on 512x512 / 256x256 crossbars -----------------------
    if (A.n_big > 0) {
        int ver = 0;
        for (;;) {
            if (threadIdx.x == 0) {
                int tsk;
                if (rank == 0) {
                    tsk = atomicAdd(A.ctr + 0, 1);
                    if (tsk >= A.n_big)
                        tsk = -1;
                    A.team_task[(size_t)team * (A.n_big + 1) + ver] = tsk;
                    __hip_atomic_store(A.team_ver + team, ver + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    if (rb::wait_geq(A.team_ver + team, ver + 1, abort))
                        tsk = A.team_task[(size_t)team * (A.n_big + 1) + ver];
                    else
                        tsk = -2;
                }
                s_task = tsk;
            }
            __syncthreads();
            const int tsk = __builtin_amdgcn_readfirstlane(s_task);
            __syncthreads();
            const int tpar = ver & 1; // the team's consecutive tasks alternate buffers
            ++ver;
            if (tsk == -2)
                return; // aborted
            if (tsk < 0)
                break;
            const int *T = A.big + (size_t)tsk * kTaskInts;
            const int cls = T[0];
            const int grp = cls == 3 ? 0 : rank / 4;
            const int inst = T[1 + grp];
            if (inst < 0)
                continue;
            rb::StripCtx X;
            X.S = cls == 3 ? 16 : 4;
            X.s = cls == 3 ? rank : rank % 4;
            const size_t tb = (size_t)team * 2 + tpar;
            X.rows = A.rows + tb * kRowsPerTeamTask + (size_t)grp * (2 * 4 * 2 * rb::kRowsLd);
            X.part = A.part + tb * kPartPerTeamTask + (size_t)grp * (2 * 4 * 2 * rb::kRowsLd);
            X.setup = A.setup + tb * kSetupPerTeamTask + (size_t)grp * (4 * 4 * rb::kRowsLd);
            X.flag = A.flags + (size_t)tsk * kTeam * 2 + grp * 4 * 2;
            X.cnt = A.cnts + (size_t)tsk * 4 + grp;
            X.abort = abort;
            bool ok;
            if (cls == 3)
                ok = rb::solve<G512, true>(make_inst(A, inst, 512), X, A.ctl[3], wsb, sh);
            else
                ok = rb::solve<G256, true>(make_inst(A, inst, 256), X, A.ctl[2], wsb, sh);
            if (!ok)
                return; // aborted: every member leaves
        }
    }

    // ---- phase 2: single-workgroup crossbars ----------------------------------
    for (;;) {
        if (threadIdx.x == 0)
            s_task = atomicAdd(A.ctr + 1, 1);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_task);
        __syncthreads();
        if (item >= A.n_small)
            break;
        const int inst = A.small[item];
        if (A.cls[inst] == 1)
            rb::solve<G128, false>(make_inst(A, inst, 128), none, A.ctl[1], wsb, sh);
        else
            rb::solve<G64, false>(make_inst(A, inst, 64), none, A.ctl[0], wsb, sh);
    }
}

__global__ void fill_i32(int *p, size_t n, int v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

} // namespace
} // namespace fadesim

using namespace fadesim;

struct fadesim_ragged_plan {
    int device = 0;
    int count = 0;
    int grid = 0, teams = 0;
    int n_big = 0, n_small = 0;
    long long cells = 0, rows_total = 0, cols_total = 0;
    rb::Ctl ctl[4];
    // device buffers, owned by the plan
    void *d_all = nullptr;
    int *d_cls = nullptr;
    long long *d_g_off = nullptr, *d_vin_off = nullptr, *d_out_off = nullptr;
    int *d_big = nullptr, *d_small = nullptr;
    int *d_ints = nullptr; // ctr[4] | team_ver[teams] | team_task[teams][n_big+1] | flags | cnts
    size_t n_ints = 0;
    int *d_ctr = nullptr, *d_team_ver = nullptr, *d_team_task = nullptr, *d_flags = nullptr, *d_cnts = nullptr;
    float *d_rows = nullptr;
    double *d_part = nullptr, *d_setup = nullptr, *d_zeros = nullptr;
    float *d_ws = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;
};

static int size_class(int r, int c)
{
    if (r != c)
        return -1;
    switch (r) {
    case 64: return 0;
    case 128: return 1;
    case 256: return 2;
    case 512: return 3;
    default: return -1;
    }
}

static double class_omega(int n)
{
    const double s = std::sin(M_PI / (2.0 * n + 1.0));
    return 2.0 / (1.0 + 0.83 * s);
}

extern "C" int fadesim_ragged_plan_create(fadesim_ragged_plan **out, int count, const int *rows, const int *cols,
                                          const fadesim_ragged_class_opts *opts)
{
    if (!out)
        return set_error(FADESIM_E_INVALID, "null plan pointer");
    *out = nullptr;
    if (count < 0 || (count > 0 && (!rows || !cols)))
        return set_error(FADESIM_E_INVALID, "count must be >= 0 with rows and cols given");
    auto *P = new fadesim_ragged_plan();
    P->count = count;
    std::vector<int> cls(count);
    std::vector<long long> goff(count), voff(count), ooff(count);
    for (int k = 0; k < count; ++k) {
        cls[k] = size_class(rows[k], cols[k]);
        if (cls[k] < 0) {
            delete P;
            return set_error(FADESIM_E_UNSUPPORTED, "instance " + std::to_string(k) + ": " + std::to_string(rows[k]) +
                                                        " x " + std::to_string(cols[k]) +
                                                        " (mixed-size solver: 64, 128, 256 or 512 square)");
        }
        goff[k] = P->cells;
        voff[k] = P->rows_total;
        ooff[k] = P->cols_total;
        P->cells += (long long)rows[k] * cols[k];
        P->rows_total += rows[k];
        P->cols_total += cols[k];
    }
    for (int c = 0; c < 4; ++c) {
        const int n = 64 << c;
        rb::Ctl &C = P->ctl[c];
        C.omega = (float)class_omega(n);
        C.tol_eff = 5e-7;
        C.switch_tol = 2e-5;
        C.dwr_scale = 16.0;
        C.max_sweeps = 100000;
        C.flags = 0;
        C.check_first = 32;
        C.check_min = 16;
        C.check_max = 256;
        C.ir2s_iters = 0;
        C.corr_iters = 2;
        C.ir2s_omega = 1.0;
        C.ir2s_geo = rb::ir2s_geo(64 << c, 64 << c);
        if (opts) {
            if (opts[c].ir2s_iters < 0 || opts[c].ir2s_iters > FADESIM_MAX_IR2S_ITERS ||
                opts[c].corr_iters < 0 || opts[c].corr_iters > FADESIM_MAX_IR2S_ITERS ||
                !(opts[c].ir2s_omega >= -FADESIM_MAX_IR2S_RHO_SCALE && opts[c].ir2s_omega <= 1.0)) {
                delete P;
                return set_error(FADESIM_E_INVALID, "ir2s_iters and corr_iters must lie in [0, " +
                                                        std::to_string(FADESIM_MAX_IR2S_ITERS) +
                                                        "], ir2s_omega in (0, 1] (0 selects 1) or [-" +
                                                        std::to_string((int)FADESIM_MAX_IR2S_RHO_SCALE) +
                                                        ", 0) (Chebyshev schedule)");
            }
            C.ir2s_iters = opts[c].ir2s_iters;
            if (opts[c].corr_iters > 0)
                C.corr_iters = opts[c].corr_iters;
            if (opts[c].ir2s_omega != 0.0)
                C.ir2s_omega = opts[c].ir2s_omega;
            if (opts[c].omega > 0) {
                if (!(opts[c].omega < 2.0)) {
                    delete P;
                    return set_error(FADESIM_E_INVALID, "omega must lie in (0, 2)");
                }
                C.omega = (float)opts[c].omega;
            }
            if (opts[c].check_first > 0)
                C.check_first = opts[c].check_first;
            if (opts[c].check_min > 0)
                C.check_min = opts[c].check_min;
            if (opts[c].check_max > 0)
                C.check_max = opts[c].check_max;
            if (C.check_max < C.check_min)
                C.check_max = C.check_min;
        }
    }
    // queues, longest first: 512 tasks, then 256 tasks of four; 128 then 64 items
    std::vector<int> big, small;
    for (int k = 0; k < count; ++k)
        if (cls[k] == 3) {
            int T[kTaskInts] = {3, k, -1, -1, -1};
            big.insert(big.end(), T, T + kTaskInts);
        }
    {
        int T[kTaskInts] = {2, -1, -1, -1, -1};
        int fill = 0;
        for (int k = 0; k < count; ++k)
            if (cls[k] == 2) {
                T[1 + fill++] = k;
                if (fill == 4) {
                    big.insert(big.end(), T, T + kTaskInts);
                    fill = 0;
                    T[1] = T[2] = T[3] = T[4] = -1;
                }
            }
        if (fill)
            big.insert(big.end(), T, T + kTaskInts);
    }
    for (int c = 1; c >= 0; --c)
        for (int k = 0; k < count; ++k)
            if (cls[k] == c)
                small.push_back(k);
    P->n_big = (int)(big.size() / kTaskInts);
    P->n_small = (int)small.size();

    if (hipGetDevice(&P->device) != hipSuccess) {
        delete P;
        return set_error(FADESIM_E_HIP, "hipGetDevice failed");
    }
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ragged_kernel, kThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device) != hipSuccess) {
        delete P;
        return set_error(FADESIM_E_HIP, "occupancy query failed");
    }
    // every workgroup must be resident (team members wait on each other), so
    // the grid never exceeds the resident count; teams of 16 when needed
    const int resident = std::max(1, per_cu * cus);
    int grid = resident;
    if (P->n_big > 0) {
        grid -= grid % kTeam;
        if (grid < kTeam) {
            delete P;
            return set_error(FADESIM_E_UNSUPPORTED,
                             "device holds fewer than 16 resident workgroups: no 256/512 crossbars");
        }
    }
    const long long useful = (long long)P->n_big * kTeam + P->n_small;
    if (useful < grid)
        grid = P->n_big > 0 ? (int)(((useful + kTeam - 1) / kTeam) * kTeam) : (int)std::max(1LL, useful);
    P->grid = grid;
    P->teams = std::max(1, grid / kTeam);

    // device memory: metadata, queues, exchange buffers, workspace slots
    P->n_ints = 4 + (size_t)P->teams * (P->n_big + 2) + (size_t)P->n_big * (kTeam * 2 + 4);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_cls = take(sizeof(int) * std::max(count, 1));
    const size_t o_goff = take(sizeof(long long) * std::max(count, 1));
    const size_t o_voff = take(sizeof(long long) * std::max(count, 1));
    const size_t o_ooff = take(sizeof(long long) * std::max(count, 1));
    const size_t o_big = take(sizeof(int) * std::max<size_t>(big.size(), 1));
    const size_t o_small = take(sizeof(int) * std::max<size_t>(small.size(), 1));
    const size_t o_ints = take(sizeof(int) * P->n_ints);
    const size_t nteambufs = P->n_big ? (size_t)P->teams * 2 : 0;
    const size_t o_rows = take(sizeof(float) * nteambufs * kRowsPerTeamTask);
    const size_t o_part = take(sizeof(double) * nteambufs * kPartPerTeamTask);
    const size_t o_setup = take(sizeof(double) * nteambufs * kSetupPerTeamTask);
    const size_t o_zeros = take(sizeof(double) * kNmax);
    const size_t o_ws = take(sizeof(float) * (size_t)P->grid * rb::kWsPlanes * kSlotCells);
    if (hipMalloc(&P->d_all, off) != hipSuccess) {
        delete P;
        return set_error(FADESIM_E_NOMEM, "plan allocation of " + std::to_string(off) + " bytes failed");
    }
    char *base = (char *)P->d_all;
    P->d_cls = (int *)(base + o_cls);
    P->d_g_off = (long long *)(base + o_goff);
    P->d_vin_off = (long long *)(base + o_voff);
    P->d_out_off = (long long *)(base + o_ooff);
    P->d_big = (int *)(base + o_big);
    P->d_small = (int *)(base + o_small);
    P->d_ints = (int *)(base + o_ints);
    P->d_ctr = P->d_ints;
    P->d_team_ver = P->d_ints + 4;
    P->d_team_task = P->d_team_ver + P->teams;
    P->d_flags = P->d_team_task + (size_t)P->teams * (P->n_big + 1);
    P->d_cnts = P->d_flags + (size_t)P->n_big * kTeam * 2;
    P->d_rows = (float *)(base + o_rows);
    P->d_part = (double *)(base + o_part);
    P->d_setup = (double *)(base + o_setup);
    P->d_zeros = (double *)(base + o_zeros);
    P->d_ws = (float *)(base + o_ws);
    std::vector<double> z(kNmax, 0.0);
    bool ok = true;
    if (count) {
        ok &= hipMemcpy(P->d_cls, cls.data(), sizeof(int) * count, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(P->d_g_off, goff.data(), sizeof(long long) * count, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(P->d_vin_off, voff.data(), sizeof(long long) * count, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(P->d_out_off, ooff.data(), sizeof(long long) * count, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!big.empty())
        ok &= hipMemcpy(P->d_big, big.data(), sizeof(int) * big.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!small.empty())
        ok &= hipMemcpy(P->d_small, small.data(), sizeof(int) * small.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(P->d_zeros, z.data(), sizeof(double) * kNmax, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipEventCreate(&P->ev[0]) == hipSuccess && hipEventCreate(&P->ev[1]) == hipSuccess;
    if (!ok) {
        fadesim_ragged_plan_destroy(P);
        return set_error(FADESIM_E_HIP, "plan upload failed");
    }
    *out = P;
    return FADESIM_OK;
}

extern "C" void fadesim_ragged_plan_destroy(fadesim_ragged_plan *P)
{
    if (!P)
        return;
    if (P->d_all) {
        (void)hipDeviceSynchronize();
        (void)hipFree(P->d_all);
    }
    if (P->ev[0])
        (void)hipEventDestroy(P->ev[0]);
    if (P->ev[1])
        (void)hipEventDestroy(P->ev[1]);
    delete P;
}

extern "C" int fadesim_ragged_plan_info(const fadesim_ragged_plan *P, fadesim_ragged_plan_stats *st)
{
    if (!P || !st)
        return set_error(FADESIM_E_INVALID, "null argument");
    st->count = P->count;
    st->cells = P->cells;
    st->rows_total = P->rows_total;
    st->cols_total = P->cols_total;
    st->grid = P->grid;
    st->teams = P->teams;
    st->big_tasks = P->n_big;
    st->small_items = P->n_small;
    for (int c = 0; c < 4; ++c) {
        st->omega[c] = P->ctl[c].omega;
        st->check_first[c] = P->ctl[c].check_first;
        st->check_min[c] = P->ctl[c].check_min;
        st->check_max[c] = P->ctl[c].check_max;
        st->ir2s_iters[c] = P->ctl[c].ir2s_iters;
        st->corr_iters[c] = P->ctl[c].corr_iters;
        st->ir2s_omega[c] = P->ctl[c].ir2s_omega;
    }
    return FADESIM_OK;
}

extern "C" int fadesim_ragged_solve(const fadesim_ragged_plan *P, const fadesim_ragged_desc *d, void *stream)
{
    if (!P || !d)
        return set_error(FADESIM_E_INVALID, "null plan or descriptor");
    if (P->count == 0)
        return FADESIM_OK;
    if (!d->d_g || !d->d_r_wire || !d->d_vin || !d->d_iout)
        return set_error(FADESIM_E_INVALID, "d_g, d_r_wire, d_vin and d_iout are required");
    const bool fixed = (d->flags & FADESIM_F_FIXED_SWEEPS) != 0;
    if (!fixed && !(d->tol > 0.0))
        return set_error(FADESIM_E_INVALID, "tol must be > 0");
    if (d->max_sweeps < 0)
        return set_error(FADESIM_E_INVALID, "max_sweeps must be >= 0");
    const double safety = d->safety > 0 ? d->safety : 2.0;
    if (safety < 1.0)
        return set_error(FADESIM_E_INVALID, "safety must be >= 1");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev != P->device)
        return set_error(FADESIM_E_INVALID, "plan was created on another device");
    hipStream_t s = (hipStream_t)stream;
    RaggedArgs A;
    A.cls = P->d_cls;
    A.g = d->d_g;
    A.g_off = P->d_g_off;
    A.rw = d->d_r_wire;
    A.vin = d->d_vin;
    A.vin_off = P->d_vin_off;
    A.vout = d->d_vout;
    A.out_off = P->d_out_off;
    A.warm = d->d_warm;
    A.field = d->d_field;
    A.iout = d->d_iout;
    A.sweeps = d->d_sweeps;
    A.status = d->d_status;
    A.est = d->d_est;
    A.zeros = P->d_zeros;
    A.big = P->d_big;
    A.n_big = P->n_big;
    A.small = P->d_small;
    A.n_small = P->n_small;
    A.ctr = P->d_ctr;
    A.team_ver = P->d_team_ver;
    A.team_task = P->d_team_task;
    A.flags = P->d_flags;
    A.cnts = P->d_cnts;
    A.rows = P->d_rows;
    A.part = P->d_part;
    A.setup = P->d_setup;
    A.ws = P->d_ws;
    for (int c = 0; c < 4; ++c) {
        A.ctl[c] = P->ctl[c];
        A.ctl[c].tol_eff = d->tol / safety;
        A.ctl[c].switch_tol = d->switch_tol > 0 ? d->switch_tol : std::fmin(1e-3, std::fmax(2e-5, 1e3 * d->tol));
        A.ctl[c].max_sweeps = d->max_sweeps;
        A.ctl[c].flags = d->flags;
        if ((d->flags & FADESIM_F_TEST_FIRST) && A.ctl[c].ir2s_iters > 0)
            A.ctl[c].check_first = 0; // the IR2S guess is tested before the first sweep
        else
            A.ctl[c].flags &= ~FADESIM_F_TEST_FIRST;
    }
    // queues, abort flag, team broadcast slots and per-task counters start at 0
    // (a fill kernel, not a memset: the call may be captured into a graph)
    hipLaunchKernelGGL(fill_i32, dim3((unsigned)std::min<size_t>((P->n_ints + 255) / 256, 4096)), dim3(256), 0, s,
                       P->d_ints, P->n_ints, 0);
    FADESIM_HIP_CHECK(hipGetLastError());
    // every status starts as "aborted": an instance the launch never finishes
    // (a team wait that timed out) keeps it, whatever an earlier call left there
    if (d->d_status) {
        hipLaunchKernelGGL(fill_i32, dim3((unsigned)std::min<size_t>(((size_t)P->count + 255) / 256, 4096)),
                           dim3(256), 0, s, d->d_status, (size_t)P->count, FADESIM_ST_ABORTED);
        FADESIM_HIP_CHECK(hipGetLastError());
    }
    auto *PP = const_cast<fadesim_ragged_plan *>(P);
    const bool timing = (d->flags & FADESIM_F_TIMING) != 0;
    if (timing)
        FADESIM_HIP_CHECK(hipEventRecord(PP->ev[0], s));
    hipLaunchKernelGGL(ragged_kernel, dim3(P->grid), dim3(kThreads), 0, s, A);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (timing) {
        FADESIM_HIP_CHECK(hipEventRecord(PP->ev[1], s));
        PP->timed = true;
    }
    return FADESIM_OK;
}

extern "C" int fadesim_ragged_aborted(const fadesim_ragged_plan *P, int *aborted)
{
    if (!P || !aborted)
        return set_error(FADESIM_E_INVALID, "null argument");
    int h[4] = {0, 0, 0, 0};
    FADESIM_HIP_CHECK(hipMemcpy(h, P->d_ctr, sizeof(h), hipMemcpyDeviceToHost));
    *aborted = h[2];
    return FADESIM_OK;
}

extern "C" double fadesim_ragged_last_ms(const fadesim_ragged_plan *P)
{
    if (!P || !P->timed)
        return -1.0;
    if (hipEventSynchronize(P->ev[1]) != hipSuccess)
        return -2.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, P->ev[0], P->ev[1]) != hipSuccess)
        return -3.0;
    return (double)ms;
}
