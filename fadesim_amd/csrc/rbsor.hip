// Batched IR-drop solver for gfx950, one shape per call: a persistent
// work-queue kernel over the red-black SOR core of rb_core.h (FP32 sweeps on
// drop variables, FP64 adjoint-weighted residual test, one FP64 refinement
// step, FP64 readout), one workgroup per crossbar.  Solves that test their
// guess before sweeping go first through guess_kernel, which runs the part of
// the solve before the first sweep without the sweep loop's register state
// and marks the crossbars its test certified; the sweeping kernel then solves
// only the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "common.h"
#include "fadesim.h"
#include "fadesim_nn.h"
#include "nn_core.h"
#include "rb_core.h"

namespace fadesim {
namespace {

struct RbArgs {
    int batch;
    int *list;         // optional: solve only instances list[k], k < *list_count (those guess_kernel
    int *list_count;   // did not finish); guess_kernel appends to them
    const float *g;
    const int *gidx;   // optional: instance b uses conductance slot gidx[b]
    int *queue;        // work counter (zeroed before the launch)
    const double *rw;
    const double *vin;
    const double *vout; // never null (zeros supplied by the host)
    int vout_ld;        // elements between instances' vout (0: one row shared by all)
    const float *warm;
    const __bf16 *coarse; // optional network output added to warm at load
    const float *cscale;
    int cch, clevel;
    float *ws;         // per workgroup slot [kWsPlanes][M][N]
    rb::Ctl ctl;
    double *iout;
    int *sweeps;
    int *status;
    double *est;
    int *rejected;
    float *field;
    // guess_net_kernel: the network's weights (fragment-major [Kp/32][Cout][32]
    // per layer) and, per instance, the combined start of each crossbar it did
    // not finish ([B][2][M][N], read by rbsor_kernel as a warm start)
    const __bf16 *netw[7];
    float *warm_out;
};

__global__ void fill_zero_f64(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// Persistent work queue: each resident workgroup takes instances until the
// batch is exhausted.
template <class G>
__global__ void __launch_bounds__(G::NT) rbsor_kernel(RbArgs A)
{
    constexpr int M = G::M, N = G::N;
    __shared__ __align__(16) float halo[G::HALO]; // also read as float2 (the coarse warm-start grid)
    __shared__ double S1[N];
    __shared__ double Ifloor[N];
    __shared__ double colI[N];
    __shared__ double colD[N];
    __shared__ float wmax[G::NW];
    __shared__ int s_inst, s_dead;
    __shared__ double s_gsum;
    __shared__ double cbuf[2 * G::TY * N];
    __shared__ double cext[3 * N];
    rb::Shared sh{halo, S1, Ifloor, colI, colD, wmax, &s_dead, cbuf, cext, &s_gsum};
    const rb::StripCtx none{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float *wsb = A.ws + (size_t)blockIdx.x * rb::kWsPlanes * M * N;
    for (;;) {
        if (threadIdx.x == 0)
            s_inst = atomicAdd(A.queue, 1);
        __syncthreads();
        // wave-uniform: keep the instance index and everything derived from it in SGPRs
        int b = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (A.list) {
            if (b >= __builtin_amdgcn_readfirstlane(rb::ldg(A.list_count)))
                break;
            b = __builtin_amdgcn_readfirstlane(rb::ldg(A.list + b));
        } else if (b >= A.batch) {
            break;
        }
        rb::Inst I;
        I.g = A.g + (size_t)(A.gidx ? A.gidx[b] : b) * M * N;
        I.rwire = A.rw[b];
        I.vin = A.vin + (size_t)b * M;
        I.vout = A.vout + (size_t)b * A.vout_ld;
        I.warm = A.warm ? A.warm + (size_t)b * 2 * M * N : nullptr;
        I.field = A.field ? A.field + (size_t)b * 2 * M * N : nullptr;
        I.iout = A.iout + (size_t)b * N;
        I.sweeps = A.sweeps ? A.sweeps + b : nullptr;
        I.status = A.status ? A.status + b : nullptr;
        I.est = A.est ? A.est + b : nullptr;
        I.rejected = A.rejected ? A.rejected + b : nullptr;
        I.rows = M;
        I.coarse = A.coarse ? A.coarse + (size_t)b * (M >> A.clevel) * (N >> A.clevel) * A.cch : nullptr;
        I.cscale = A.coarse ? A.cscale[b] : 0.f;
        I.cch = A.cch;
        I.clevel = A.clevel;
        rb::solve<G, false>(I, none, A.ctl, wsb, sh);
    }
}

// The guess-first part of every solve of the batch (rb::guess_solve), one
// workgroup per crossbar as in rbsor_kernel.  It carries none of the sweep
// loop's state: the drops stay in registers through the IR2S iterations and
// go to LDS for the test, so of the workspace only the adjoint plane is
// written and read.  A crossbar whose solve ended at a test before any sweep
// has its outputs written; every other one is appended to list (list_count
// entries).  kGuessHoldsTile: the IR2S iterations hold the conductance tile in
// registers instead of rereading it from the cache.
constexpr bool kGuessHoldsTile = false;

template <class G>
__global__ void __launch_bounds__(G::NT) guess_kernel(RbArgs A)
{
    constexpr int M = G::M, N = G::N;
    // the tested drops ([2][M][N]) overlay the halo (here only the stage of
    // the coarse combine) and the column buffer, both idle during the test
    constexpr int kHaloB = G::HALO * (int)sizeof(float), kCbufB = 2 * G::TY * N * (int)sizeof(double);
    constexpr int kFldB = 2 * M * N * (int)sizeof(float);
    constexpr int kBigB = kHaloB + kCbufB > kFldB ? kHaloB + kCbufB : kFldB;
    static_assert(kCbufB % 16 == 0, "halo alignment");
    __shared__ __align__(16) unsigned char big[kBigB];
    __shared__ double S1[N];
    __shared__ double Ifloor[N];
    __shared__ double colI[N];
    __shared__ double colD[N];
    __shared__ float wmax[G::NW];
    __shared__ int s_inst, s_dead;
    __shared__ double s_gsum;
    __shared__ double cext[3 * N];
    // the column buffer first: its rows are then in reach of the LDS
    // instructions' immediate offsets from one address register
    double *cbuf = reinterpret_cast<double *>(big);
    float *halo = reinterpret_cast<float *>(big + kCbufB);
    float *fld = reinterpret_cast<float *>(big);
    rb::Shared sh{halo, S1, Ifloor, colI, colD, wmax, &s_dead, cbuf, cext, &s_gsum};
    float *wsb = A.ws + (size_t)blockIdx.x * rb::kWsPlanes * M * N;
    for (;;) {
        if (threadIdx.x == 0)
            s_inst = atomicAdd(A.queue, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (b >= A.batch)
            break;
        rb::Inst I;
        I.g = A.g + (size_t)(A.gidx ? A.gidx[b] : b) * M * N;
        I.rwire = A.rw[b];
        I.vin = A.vin + (size_t)b * M;
        I.vout = A.vout + (size_t)b * A.vout_ld;
        I.warm = A.warm ? A.warm + (size_t)b * 2 * M * N : nullptr;
        I.field = A.field ? A.field + (size_t)b * 2 * M * N : nullptr;
        I.iout = A.iout + (size_t)b * N;
        I.sweeps = A.sweeps ? A.sweeps + b : nullptr;
        I.status = A.status ? A.status + b : nullptr;
        I.est = A.est ? A.est + b : nullptr;
        I.rejected = A.rejected ? A.rejected + b : nullptr;
        I.rows = M;
        I.coarse = A.coarse ? A.coarse + (size_t)b * (M >> A.clevel) * (N >> A.clevel) * A.cch : nullptr;
        I.cscale = A.coarse ? A.cscale[b] : 0.f;
        I.cch = A.cch;
        I.clevel = A.clevel;
        const int start = I.coarse ? (I.warm ? rb::kStartWarmCoarse : rb::kStartBaseCoarse)
                                   : I.warm ? rb::kStartWarm : rb::kStartZero;
        auto load = [&](int st, int t, int i0, int j0, float (&p)[G::TR][G::TC], float (&q)[G::TR][G::TC]) {
            const rb::StripCtx X{1, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
            float gt[G::TR][G::TC];
            int nbar = 0;
            rb::start_tile<G, false, kGuessHoldsTile>(I, X, sh, gt, I.g, 0, i0, j0, t, t % G::TX, t / G::TX, I.rwire,
                                                      nbar, st, p, q);
        };
        const bool done = rb::guess_solve<G, kGuessHoldsTile>(I, A.ctl, wsb, sh, fld, start, load);
        if (!done && threadIdx.x == 0)
            A.list[atomicAdd(A.list_count, 1)] = b;
        // a barrier between this thread-0 append and the next iteration's
        // thread-0 queue update: without one the compiler joins the two into
        // one path for thread 0 around the loop's barriers, so the other lanes
        // of its wave would read s_inst before thread 0 has updated it
        __syncthreads();
    }
}

// guess_kernel with the learned warm start formed in-kernel (128x128, the
// coarse16f network): per crossbar, the pooled features and the base field
// (two undamped wire-solve iterations from zero, nn::pooled_features), the
// network on the workgroup's 8 waves (nn::coarse16_body_pad, every
// activation in LDS), and base + S2 x bilinear(network output) -- the arithmetic of the
// features kernel, the network kernel and the combine at load, term by term,
// so the drops equal those of the three-kernel path bit for bit.  The
// network's buffers overlay `big` (idle between the set-up and the test).  A
// crossbar this kernel does not finish has its combined start recomputed and
// written to warm_out, from which rbsor_kernel then solves it as it would from
// the base field and the network output.
using NetGeo = rb::Geo<4, 8, 32, 16>;
constexpr int kNetWidth = 16;

__global__ void __launch_bounds__(NetGeo::NT) guess_net_kernel(RbArgs A)
{
    using G = NetGeo;
    using Lds = nn::Net16PadLds;
    constexpr int M = G::M, N = G::N, TR = G::TR, TC = G::TC, TX = G::TX, NT = G::NT;
    constexpr int kHaloB = G::HALO * (int)sizeof(float), kCbufB = 2 * G::TY * N * (int)sizeof(double);
    constexpr int kFldB = 2 * M * N * (int)sizeof(float);
    constexpr int kBigB = kHaloB + kCbufB > kFldB ? kHaloB + kCbufB : kFldB;
    // network buffers (zero-bordered layout): bufA, bufS, bufX; the last
    // layer writes its channel pairs to the stage of the combine, the halo
    // region (from byte kCbufB), which lies within bufS, idle by then
    constexpr int kA = Lds::kA * (int)sizeof(__bf16), kS = Lds::kS * (int)sizeof(__bf16);
    static_assert(kA + kS + Lds::kX * (int)sizeof(__bf16) <= kBigB, "network buffers fit the solver's LDS");
    static_assert(kA + kS >= kCbufB, "the features land beyond the column buffer of the wire iterations");
    static_assert(kCbufB >= kA && kCbufB + 2 * 32 * 32 * (int)sizeof(float) <= kA + kS, "stage within bufS");
    static_assert(kA % 16 == 0 && kS % 16 == 0, "16-byte aligned buffers");
    __shared__ __align__(16) unsigned char big[kBigB];
    __shared__ double S1[N];
    __shared__ double Ifloor[N];
    __shared__ double colI[N];
    __shared__ double colD[N];
    __shared__ float wmax[G::NW];
    __shared__ int s_inst, s_dead;
    __shared__ double s_gsum;
    __shared__ double cext[3 * N];
    __shared__ float red[4 * G::NW];
    double *cbuf = reinterpret_cast<double *>(big);
    float *halo = reinterpret_cast<float *>(big + kCbufB);
    float *fld = reinterpret_cast<float *>(big);
    __bf16 *bufA = reinterpret_cast<__bf16 *>(big);
    __bf16 *bufS = bufA + Lds::kA, *bufX = bufS + Lds::kS;
    rb::Shared sh{halo, S1, Ifloor, colI, colD, wmax, &s_dead, cbuf, cext, &s_gsum};
    float *wsb = A.ws + (size_t)blockIdx.x * rb::kWsPlanes * M * N;
    const __bf16 *const w[7] = {A.netw[0], A.netw[1], A.netw[2], A.netw[3], A.netw[4], A.netw[5], A.netw[6]};
    for (;;) {
        if (threadIdx.x == 0)
            s_inst = atomicAdd(A.queue, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (b >= A.batch)
            break;
        rb::Inst I;
        I.g = A.g + (size_t)(A.gidx ? A.gidx[b] : b) * M * N;
        I.rwire = A.rw[b];
        I.vin = A.vin + (size_t)b * M;
        I.vout = A.vout + (size_t)b * A.vout_ld;
        I.warm = nullptr;
        I.field = A.field ? A.field + (size_t)b * 2 * M * N : nullptr;
        I.iout = A.iout + (size_t)b * N;
        I.sweeps = A.sweeps ? A.sweeps + b : nullptr;
        I.status = A.status ? A.status + b : nullptr;
        I.est = A.est ? A.est + b : nullptr;
        I.rejected = A.rejected ? A.rejected + b : nullptr;
        I.rows = M;
        I.coarse = nullptr;
        I.cscale = 0.f;
        I.cch = I.clevel = 0;
        // the combined start: features and base, network, base + S2 x bilinear(out)
        auto net_start = [&](int i0, int j0, float (&p)[TR][TC], float (&q)[TR][TC]) {
            const float S2 = nn::pooled_features<true>(sh, I.g, I.rwire, I.vin, bufX, red, p, q);
            nn::zero_border<NT, 32, 32, nn::kFeatC>(bufX);
            __syncthreads();
            nn::coarse16_body_pad<NT / 64, nn::kOutStage>(bufX, bufA, bufS, w, halo);
            __syncthreads();
            asm volatile("" : "+v"(i0), "+v"(j0));
            rb::add_coarse_staged<G>(halo, 2, S2, M, 0, i0, j0, p, q);
            __syncthreads();
        };
        auto load = [&](int st, int t, int i0, int j0, float (&p)[TR][TC], float (&q)[TR][TC]) {
            if (st == rb::kStartNet)
                net_start(i0, j0, p, q);
            else
                rb::load_start<G, false>(I, sh, t, 0, i0, j0, st, p, q);
        };
        const bool done = rb::guess_solve<G, kGuessHoldsTile>(I, A.ctl, wsb, sh, fld, rb::kStartNet, load);
        if (!done) {
            // rare: the start again, for the sweeping kernel
            int t = threadIdx.x;
            asm volatile("" : "+v"(t));
            const int i0 = TR * (t / TX), j0 = TC * (t % TX);
            float p[TR][TC], q[TR][TC];
            net_start(i0, j0, p, q);
            float *wo = A.warm_out + (size_t)b * 2 * M * N;
#pragma unroll
            for (int ii = 0; ii < TR; ++ii) {
                rb::strow<TC>(wo + (size_t)(i0 + ii) * N + j0, p[ii]);
                rb::strow<TC>(wo + (size_t)M * N + (size_t)(i0 + ii) * N + j0, q[ii]);
            }
            if (threadIdx.x == 0)
                A.list[atomicAdd(A.list_count, 1)] = b;
        }
        // a barrier between this thread-0 append and the next iteration's
        // thread-0 queue update (see guess_kernel)
        __syncthreads();
    }
}

__global__ void zero_counters(int *a, int n)
{
    for (int k = 0; k < n; ++k)
        a[k] = 0;
}

// Scratch of fadesim_solve_batch_noalloc: a zero row for an absent vout
// (kScratchZeroBytes), then the persistent grid's queue counter and
// per-workgroup slots, as launch_rb lays them out in the workspace.  Sized
// for 256 workgroups of the 128x128 geometry (one per CU of an MI355X).
constexpr size_t kScratchZeroBytes = 4096;
constexpr size_t kScratchSlotsBytes = (size_t)256 * rb::kWsPlanes * 128 * 128 * sizeof(float) + 256;
constexpr size_t kScratchBytes = kScratchZeroBytes + kScratchSlotsBytes;
static __device__ __align__(256) char scratch[kScratchBytes];

// t_ev[0] .. t_ev[1]: both kernels of the last timed call; t_ev[2]: the end
// of its guess_kernel (t_guess: that kernel ran)
thread_local hipEvent_t t_ev[3] = {nullptr, nullptr, nullptr};
thread_local bool t_ev_valid = false, t_guess = false;

// Resident workgroups of kernel K (K_ID: its cache entry) on the current
// device: the persistent grid.
template <int K_ID, class KERNEL>
int resident_slots(KERNEL kernel, int threads)
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev >= 64)
        return 256;
    if (!cached[dev]) {
        int per_cu = 0, cus = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
        if (e != hipSuccess || per_cu < 1)
            per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        cached[dev] = per_cu * cus;
    }
    return cached[dev];
}

template <class G>
int sweep_slots() { return resident_slots<0 + 3 * (G::M * 1000 + G::N), void (*)(RbArgs)>(rbsor_kernel<G>, G::NT); }
template <class G>
int guess_slots() { return resident_slots<1 + 3 * (G::M * 1000 + G::N), void (*)(RbArgs)>(guess_kernel<G>, G::NT); }
int guess_net_slots() { return resident_slots<2, void (*)(RbArgs)>(guess_net_kernel, NetGeo::NT); }

// Solves whose guess is tested before any sweep go through guess_kernel
// first: they test a guess (IR2S iterations, a warm start or a coarse
// combine), and no fixed sweep count bypasses the test.
bool takes_guess_path(const RbArgs &args)
{
    const unsigned f = args.ctl.flags;
    return (f & FADESIM_F_TEST_FIRST) && !(f & FADESIM_F_FIXED_SWEEPS) && !(f & FADESIM_F_NO_FASTPATH) &&
           (args.ctl.ir2s_iters > 0 || args.warm || args.coarse);
}

// fixed: the workspace comes from `fixed` (fixed_bytes long) instead of the
// library-owned one; the grids shrink to the slots that fit.
// Layout: [slots][256 B: sweeping queue, guess queue, list count][guess path: list, batch ints].
// The two kernels run one after the other on the stream and share the slots.
// NET: the guess kernel is guess_net_kernel (args.netw, args.warm_out set;
// the crossbars it lists are solved from warm_out).
template <class G, bool NET = false>
int launch_rb(RbArgs args, int M, int N, hipStream_t s, char *fixed = nullptr, size_t fixed_bytes = 0)
{
    // the coarse warm-start grid is staged in the halo planes' LDS
    if (args.coarse && 2 * (M >> args.clevel) * (N >> args.clevel) > G::HALO)
        return set_error(FADESIM_E_UNSUPPORTED, "d_warm_coarse: coarse grid too large for this shape (raise warm_level)");
    const size_t slot_bytes = (size_t)rb::kWsPlanes * G::M * G::N * sizeof(float);
    const size_t list_bytes = ((size_t)args.batch * sizeof(int) + 255) / 256 * 256;
    if (NET && fixed)
        return set_error(FADESIM_E_UNSUPPORTED, "the in-kernel network start needs the library workspace");
    bool guess = NET || takes_guess_path(args);
    int grid = std::min(sweep_slots<G>(), args.batch);
    int ggrid = guess ? std::min(NET ? guess_net_slots() : guess_slots<G>(), args.batch) : 0;
    char *ws;
    if (fixed) {
        if (guess && fixed_bytes < 256 + list_bytes + slot_bytes)
            guess = false; // the list does not fit beside one slot: every crossbar goes to the sweeping kernel
        const size_t fit = (fixed_bytes - 256 - (guess ? list_bytes : 0)) / slot_bytes;
        if (fit < 1)
            return set_error(FADESIM_E_UNSUPPORTED, "scratch too small for one workgroup of this shape");
        grid = (int)std::min((size_t)grid, fit);
        ggrid = (int)std::min((size_t)ggrid, fit);
        ws = fixed;
    } else {
        ws = (char *)workspace_aux((size_t)std::max(grid, ggrid) * slot_bytes + 256 + (guess ? list_bytes : 0));
        if (!ws)
            return FADESIM_E_NOMEM;
    }
    const size_t slots_bytes = (size_t)std::max(grid, ggrid) * slot_bytes;
    args.ws = (float *)ws;
    int *ctr = (int *)(ws + slots_bytes);
    args.queue = ctr;
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(1), 0, s, ctr, guess ? 3 : 1);
    FADESIM_HIP_CHECK(hipGetLastError());
    const bool timing = (args.ctl.flags & FADESIM_F_TIMING) != 0;
    if (timing) {
        if (!t_ev[0]) {
            FADESIM_HIP_CHECK(hipEventCreate(&t_ev[0]));
            FADESIM_HIP_CHECK(hipEventCreate(&t_ev[1]));
            FADESIM_HIP_CHECK(hipEventCreate(&t_ev[2]));
        }
        FADESIM_HIP_CHECK(hipEventRecord(t_ev[0], s));
    }
    if (guess) {
        RbArgs ga = args;
        ga.queue = ctr + 1;
        ga.list = (int *)(ws + slots_bytes + 256);
        ga.list_count = ctr + 2;
        if constexpr (NET)
            hipLaunchKernelGGL(guess_net_kernel, dim3(ggrid), dim3(G::NT), 0, s, ga);
        else
            hipLaunchKernelGGL((guess_kernel<G>), dim3(ggrid), dim3(G::NT), 0, s, ga);
        FADESIM_HIP_CHECK(hipGetLastError());
        if (timing)
            FADESIM_HIP_CHECK(hipEventRecord(t_ev[2], s));
        args.list = ga.list;
        args.list_count = ga.list_count;
        if constexpr (NET)
            args.warm = args.warm_out;
    }
    hipLaunchKernelGGL((rbsor_kernel<G>), dim3(grid), dim3(G::NT), 0, s, args);
    FADESIM_HIP_CHECK(hipGetLastError());
    if (timing) {
        FADESIM_HIP_CHECK(hipEventRecord(t_ev[1], s));
        t_ev_valid = true;
        t_guess = guess;
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

extern "C" double fadesim_last_guess_ms(void)
{
    if (!t_ev_valid || !t_guess)
        return -1.0;
    if (hipEventSynchronize(t_ev[2]) != hipSuccess)
        return -2.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t_ev[0], t_ev[2]) != hipSuccess)
        return -3.0;
    return (double)ms;
}

extern "C" double fadesim_default_omega(int rows, int cols)
{
    const int L = rows > cols ? rows : cols;
    const double s = std::sin(M_PI / (2.0 * L + 1.0));
    return 2.0 / (1.0 + 0.83 * s);
}

namespace {

// Validates d and fills the kernel arguments (everything but the workspace
// and vout).  Returns FADESIM_OK, or an error; *empty = true when there is
// nothing to solve.
int batch_args(const fadesim_batch_desc *d, RbArgs &args, bool *empty)
{
    *empty = false;
    if (!d)
        return set_error(FADESIM_E_INVALID, "null descriptor");
    if (d->struct_size != sizeof(fadesim_batch_desc))
        return set_error(FADESIM_E_INVALID, "fadesim_batch_desc.struct_size is " + std::to_string(d->struct_size) +
                                                ", the library's descriptor has " +
                                                std::to_string(sizeof(fadesim_batch_desc)) + " bytes");
    const int B = d->batch, M = d->rows, N = d->cols;
    if (B < 0)
        return set_error(FADESIM_E_INVALID, "batch must be >= 0");
    if (B == 0) {
        *empty = true;
        return FADESIM_OK;
    }
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
    if (d->ir2s_iters < 0 || d->ir2s_iters > FADESIM_MAX_IR2S_ITERS || d->corr_iters < 0 ||
        d->corr_iters > FADESIM_MAX_IR2S_ITERS)
        return set_error(FADESIM_E_INVALID, "ir2s_iters and corr_iters must lie in [0, " +
                                                std::to_string(FADESIM_MAX_IR2S_ITERS) + "]");
    if (!(d->ir2s_omega >= -FADESIM_MAX_IR2S_RHO_SCALE && d->ir2s_omega <= 1.0))
        return set_error(FADESIM_E_INVALID, "ir2s_omega must lie in (0, 1] (0 selects 1), or in [-" +
                                                 std::to_string((int)FADESIM_MAX_IR2S_RHO_SCALE) +
                                                 ", 0) for the Chebyshev schedule");
    if (d->d_warm_coarse) {
        const int L = d->warm_level;
        if (!d->d_warm_scale || d->warm_coarse_ch < 2 || (d->warm_coarse_ch & 1) || L < 1 || L > 6 || ((M >> L) << L) != M ||
            ((N >> L) << L) != N || (M >> L) < 1 || (N >> L) < 1)
            return set_error(FADESIM_E_INVALID, "d_warm_coarse needs d_warm_scale, an even warm_coarse_ch "
                                                ">= 2 and warm_level in 1..6 dividing rows and cols");
    }
    args.batch = B;
    args.list = nullptr;
    args.list_count = nullptr;
    args.g = d->d_g;
    args.gidx = d->d_gidx;
    args.queue = nullptr;
    args.rw = d->d_r_wire;
    args.vin = d->d_vin;
    args.vout = d->d_vout;
    args.vout_ld = N;
    args.warm = d->d_warm;
    args.coarse = (const __bf16 *)d->d_warm_coarse;
    args.cscale = d->d_warm_scale;
    args.cch = d->warm_coarse_ch;
    args.clevel = d->warm_level;
    args.ws = nullptr;
    args.ctl.omega = (float)w;
    args.ctl.tol_eff = d->tol / safety;
    args.ctl.switch_tol = d->switch_tol > 0 ? d->switch_tol : std::fmin(1e-3, std::fmax(2e-5, 1e3 * d->tol));
    args.ctl.dwr_scale = 16.0;
    args.ctl.max_sweeps = d->max_sweeps;
    args.ctl.flags = d->flags;
    args.ctl.check_first = (d->flags & FADESIM_F_TEST_FIRST) ? 0 : 32;
    args.ctl.check_min = 16;
    args.ctl.check_max = 256;
    args.ctl.ir2s_iters = d->ir2s_iters;
    args.ctl.corr_iters = d->corr_iters > 0 ? d->corr_iters : 2;
    args.ctl.ir2s_omega = d->ir2s_omega != 0.0 ? d->ir2s_omega : 1.0;
    args.ctl.ir2s_geo = rb::ir2s_geo(d->rows, d->cols);
    args.iout = d->d_iout;
    args.sweeps = d->d_sweeps;
    args.status = d->d_status;
    args.est = d->d_est;
    args.rejected = d->d_guess_rejected;
    args.field = d->d_field;
    for (int l = 0; l < 7; ++l)
        args.netw[l] = nullptr;
    args.warm_out = nullptr;
    return FADESIM_OK;
}

int launch_shape(const RbArgs &args, int M, int N, hipStream_t s, char *fixed, size_t fixed_bytes)
{
#define FADESIM_RB_CASE(TRv, TCv, TYv, TXv)                                 \
    if (M == TRv * TYv && N == TCv * TXv)                                   \
        return launch_rb<rb::Geo<TRv, TCv, TYv, TXv>>(args, M, N, s, fixed, fixed_bytes);
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

} // namespace

extern "C" int fadesim_solve_batch(const fadesim_batch_desc *d, void *stream)
{
    RbArgs args;
    bool empty;
    int rc = batch_args(d, args, &empty);
    if (rc != FADESIM_OK || empty)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    const int M = d->rows, N = d->cols;
    if (!args.vout) {
        // one zero row shared by every instance
        double *z = (double *)workspace((size_t)N * sizeof(double));
        if (!z)
            return FADESIM_E_NOMEM;
        hipLaunchKernelGGL(fill_zero_f64, dim3(1), dim3(256), 0, s, z, (size_t)N);
        FADESIM_HIP_CHECK(hipGetLastError());
        args.vout = z;
        args.vout_ld = 0;
    }
    return launch_shape(args, M, N, s, nullptr, 0);
}

namespace {
std::mutex g_scratch_mutex;
char *g_scratch_ptr[64] = {nullptr};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
char *scratch_on_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error(FADESIM_E_HIP, "hipGetDevice failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(FADESIM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (char *)p;
    }
    return g_scratch_ptr[dev];
}
} // namespace

extern "C" int fadesim_solve_batch_noalloc(const fadesim_batch_desc *d, void *stream)
{
    RbArgs args;
    bool empty;
    int rc = batch_args(d, args, &empty);
    if (rc != FADESIM_OK || empty)
        return rc;
    if (d->flags & FADESIM_F_TIMING)
        return set_error(FADESIM_E_INVALID, "fadesim_solve_batch_noalloc: FADESIM_F_TIMING creates events; "
                                            "time the call with the caller's own events");
    if ((size_t)d->cols * sizeof(double) > kScratchZeroBytes)
        return set_error(FADESIM_E_UNSUPPORTED, "fadesim_solve_batch_noalloc: cols beyond the scratch zero row");
    char *base = scratch_on_current_device();
    if (!base)
        return FADESIM_E_HIP;
    hipStream_t s = (hipStream_t)stream;
    if (!args.vout) {
        // one zero row for every instance, cleared by a kernel (no memset node
        // if the caller captures the call)
        double *z = (double *)base;
        hipLaunchKernelGGL(fill_zero_f64, dim3(1), dim3(256), 0, s, z, (size_t)d->cols);
        FADESIM_HIP_CHECK(hipGetLastError());
        args.vout = z;
        args.vout_ld = 0;
    }
    return launch_shape(args, d->rows, d->cols, s, base + kScratchZeroBytes, kScratchSlotsBytes);
}

extern "C" int fadesim_nn_solve_batch(const fadesim_batch_desc *d, const void *const *d_w, const int *kp, void *stream)
{
    RbArgs args;
    bool empty;
    int rc = batch_args(d, args, &empty);
    if (rc != FADESIM_OK)
        return rc;
    if (!d_w || !kp)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_solve_batch: d_w and kp are required");
    const int cin[7] = {8, kNetWidth, kNetWidth, 2 * kNetWidth, 2 * kNetWidth, kNetWidth, kNetWidth};
    for (int l = 0; l < 7; ++l)
        if (!d_w[l] || kp[l] != nn::net_kpad(cin[l]))
            return set_error(FADESIM_E_UNSUPPORTED, "fadesim_nn_solve_batch: the weights are not the coarse16f "
                                                    "network's (layer " + std::to_string(l) + ")");
    if (d->rows != NetGeo::M || d->cols != NetGeo::N)
        return set_error(FADESIM_E_UNSUPPORTED, "fadesim_nn_solve_batch: 128 x 128 crossbars only");
    if (d->d_warm || d->d_warm_coarse || d->d_vout)
        return set_error(FADESIM_E_INVALID, "fadesim_nn_solve_batch: d_warm, d_warm_coarse and d_vout must be NULL "
                                            "(the start is formed in-kernel, for grounded bitlines)");
    const unsigned f = d->flags;
    if (!(f & FADESIM_F_TEST_FIRST) || (f & (FADESIM_F_FIXED_SWEEPS | FADESIM_F_NO_FASTPATH)))
        return set_error(FADESIM_E_INVALID, "fadesim_nn_solve_batch: needs FADESIM_F_TEST_FIRST, without "
                                            "FADESIM_F_FIXED_SWEEPS or FADESIM_F_NO_FASTPATH");
    if (empty)
        return FADESIM_OK;
    hipStream_t s = (hipStream_t)stream;
    constexpr int M = NetGeo::M, N = NetGeo::N;
    double *z = (double *)workspace((size_t)N * sizeof(double));
    if (!z)
        return FADESIM_E_NOMEM;
    hipLaunchKernelGGL(fill_zero_f64, dim3(1), dim3(256), 0, s, z, (size_t)N);
    FADESIM_HIP_CHECK(hipGetLastError());
    args.vout = z;
    args.vout_ld = 0;
    for (int l = 0; l < 7; ++l)
        args.netw[l] = (const __bf16 *)d_w[l];
    args.warm_out = (float *)workspace_warm((size_t)args.batch * 2 * M * N * sizeof(float));
    if (!args.warm_out)
        return FADESIM_E_NOMEM;
    return launch_rb<NetGeo, true>(args, M, N, s);
}
