/*
 * fadesim.h -- C ABI of the MI355X-native FADESim IR-drop solver.
 *
 * Every pointer argument named d_* is a device pointer (HBM) owned by the
 * caller; every size is in elements.  `stream` is a hipStream_t passed as
 * void* (NULL = the default stream).  Calls are stream-ordered and do not
 * synchronise the host unless stated.  Functions return 0 on success or a
 * negative FADESIM_E_* code; fadesim_last_error() returns the message of the
 * last failure on the calling thread.
 *
 * Network (both entry points): an n x m crossbar; wordline ("up") nodes u_ij
 * are driven from the left through segment 0 by vin_i, bitline ("down")
 * nodes d_ij sink at the bottom through segment n into vout_j; cell (i,j)
 * joins u_ij and d_ij with conductance g_ij.  The bitline current read out
 * is I_j = sum_i g_ij (u_ij - d_ij).
 */
#ifndef FADESIM_H
#define FADESIM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FADESIM_OK 0
#define FADESIM_E_INVALID -1   /* bad argument (size, pointer, omega ...) */
#define FADESIM_E_UNSUPPORTED -2 /* shape not handled by any kernel */
#define FADESIM_E_HIP -3       /* HIP runtime error */
#define FADESIM_E_NOMEM -4     /* workspace allocation failed */

/* per-instance status written by fadesim_solve_batch */
#define FADESIM_ST_CONVERGED 0
#define FADESIM_ST_MAXIT 1
#define FADESIM_ST_NONFINITE 2
#define FADESIM_ST_ABORTED 3      /* fadesim_ragged_solve: the launch was aborted before this
                                   * instance was solved (its other outputs are not valid) */

/* flags of fadesim_batch_desc */
#define FADESIM_F_FIXED_SWEEPS 1u /* run exactly max_sweeps sweeps, no test, no refinement */
#define FADESIM_F_NO_REFINE 2u    /* skip the FP64 iterative-refinement step */
#define FADESIM_F_DWR 4u          /* return I - (adjoint-weighted residual) and stop on it */
#define FADESIM_F_TIMING 8u       /* record HIP events around the solve kernel */
#define FADESIM_F_ZERO_CORRECTION 16u /* start the refinement correction from zero instead of
                                        * the wire Green's-function (IR2S) estimate */
#define FADESIM_F_TEST_FIRST 32u  /* test the initial guess (and the refinement correction's
                                   * initial guess) before sweeping: for accurate guesses */
#define FADESIM_F_NO_FASTPATH 64u /* solve every crossbar on the sweeping kernel alone (by default a
                                   * solve with FADESIM_F_TEST_FIRST and a guess to test runs its
                                   * steps up to the first sweep in a kernel of twice the occupancy,
                                   * and only the crossbars that test does not certify reach the
                                   * sweeping kernel; the outputs are the same bit for bit) */
#define FADESIM_MAX_IR2S_ITERS 64
/* Largest |ir2s_omega| of the Chebyshev schedule (ir2s_omega < 0). */
#define FADESIM_MAX_IR2S_RHO_SCALE 4.0

const char *fadesim_last_error(void);
int fadesim_version(void);

/*
 * Batched IR-drop solve, FP32 red-black SOR on drop variables with an FP64
 * residual-based convergence test, one FP64 iterative-refinement step and an
 * FP64 readout.  All instances of one call share rows x cols.
 *
 * Replaces the per-crossbar solve of
 *   NodebasedGSMethod::nodeBaseGS   crossbarHspice/V2/concrete_command.h:1136-1361
 *   array2D::nodeBasedGS            crossbarHspice/2dArray.cc:1377-1687
 * and the sweep loop of ir_drop_gs  pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/
 *                                   split_arr_matmul.cu:204-333
 * for uniform wire resistance per instance, converged to a relative I_out
 * tolerance instead of a fixed or max|dv| sweep rule.
 */
typedef struct fadesim_batch_desc {
    size_t struct_size;     /* sizeof(fadesim_batch_desc) as the caller compiled it; a call whose
                             * size differs from the library's is refused (FADESIM_E_INVALID) */
    int batch;              /* B instances */
    int rows, cols;         /* n, m: multiples of 4, 8..128 */
    const float *d_g;       /* [B][n][m] cell conductance (S), or [S][n][m] with d_gidx */
    const int *d_gidx;      /* [B] conductance slot of each instance, or NULL (slot b) */
    const double *d_r_wire; /* [B] wire segment resistance (ohm) */
    const double *d_vin;    /* [B][n] wordline source voltage (V) */
    const double *d_vout;   /* [B][m] bitline sink voltage (V), or NULL = 0 V */
    const float *d_warm;    /* [B][2][n][m] initial drops (vin-u, d-vout), or NULL = 0 */
    double omega;           /* SOR factor in (0,2); <= 0 selects the built-in rule */
    double tol;             /* target relative I_out error (e.g. 1e-6) */
    double safety;          /* stop when estimate * safety <= tol (>= 1; 0 -> 2) */
    double switch_tol;      /* estimate at which the refinement step is taken
                             * (0 -> min(1e-3, max(2e-5, 1e3 tol))) */
    int max_sweeps;         /* sweep budget (FIXED_SWEEPS: exact count) */
    unsigned flags;         /* FADESIM_F_* */
    double *d_iout;         /* [B][m] bitline currents (A), FP64 */
    int *d_sweeps;          /* [B] sweeps performed, or NULL */
    int *d_status;          /* [B] FADESIM_ST_*, or NULL */
    double *d_est;          /* [B] final relative error estimate, or NULL */
    float *d_field;         /* [B][2][n][m] final drops, or NULL */
    int ir2s_iters;         /* wire-solve (IR2S) iterations applied to the initial guess
                             * before the first sweep (0..FADESIM_MAX_IR2S_ITERS); each solves
                             * every wordline and bitline exactly for the cell currents of the
                             * present drops (crossbarHspice/ir_drop_fast_solver.h:24-80) */
    int corr_iters;         /* IR2S iterations of the refinement correction's initial
                             * guess (0 -> 2) */
    double ir2s_omega;      /* damping w of the IR2S iterations, x <- (1-w) x + w x_IR2S,
                             * in (0, 1]; 0 -> 1.  With the coupling's spectral radius rho
                             * (~0.8 r gbar rows cols) the contraction is rho/(2+rho) at
                             * w = 2/(2+rho), against rho undamped.  Negative, in
                             * [-FADESIM_MAX_IR2S_RHO_SCALE, 0): a Chebyshev schedule, step k
                             * of K with w_k = 1 / (1 + (rho/2)(1 + cos(pi (2k+1) / (2K)))),
                             * rho = |ir2s_omega| r gbar (1/(4 sin^2(pi/(4 cols+2))) +
                             * 1/(4 sin^2(pi/(4 rows+2)))), gbar the instance's mean
                             * conductance: the K steps reduce the error by
                             * 1/T_K((2+rho)/rho) (T_K: Chebyshev polynomial), against
                             * (rho/(2+rho))^K.  The same schedule serves corr_iters. */
    /* Warm start given as base + d_warm_scale[b] x bilinear(network output):
     * when d_warm_coarse is not NULL the initial drops are formed at load,
     * with the arithmetic of fadesim_nn_combine_up (channel 0 -> wordline drops,
     * 1 -> bitline drops), so that kernel's field is never written.  The base
     * is d_warm or, when d_warm is NULL, formed in the kernel: two undamped
     * IR2S iterations from zero drops (the base fadesim_nn_features_pooled
     * writes, bit for bit, when the bitlines are grounded).  warm_level 1..6
     * with the coarse grid small enough for the kernel's LDS staging:
     * 2 (n >> level)(m >> level) <= 12800 floats, so every level >= 1 fits at
     * 128x128. */
    const void *d_warm_coarse;  /* [B][n >> warm_level][m >> warm_level][warm_coarse_ch] BF16, or NULL */
    const float *d_warm_scale;  /* [B] */
    int warm_coarse_ch;         /* even, >= 2 */
    int warm_level;
    /* A given initial guess (d_warm, with or without d_warm_coarse) is judged at
     * the first convergence test: when its estimated error exceeds that of the
     * zero-drop guess (estimated from the same test), the solve restarts from
     * zero drops, with ir2s_iters IR2S iterations on them. */
    int *d_guess_rejected;      /* [B] out: 1 where that happened, else 0; or NULL */
} fadesim_batch_desc;

int fadesim_solve_batch(const fadesim_batch_desc *desc, void *stream);

/*
 * fadesim_solve_batch that allocates nothing: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * results as fadesim_solve_batch.  Its workspace is a file-scope device array
 * of 83,890,432 bytes (80 MiB + 4.25 KiB) that every GPU the library runs on
 * holds until the process ends: a zero row for an absent d_vout, the queue
 * counter and the slots of up to 256 resident workgroups at 128x128 (more at
 * smaller shapes; a larger batch is worked through by the same workgroups).  Every call on one GPU shares that array, so calls that overlap on
 * one GPU race: order them on one stream.  FADESIM_F_TIMING is refused (it
 * would create events).  Nothing is cleared with a memset node.
 */
int fadesim_solve_batch_noalloc(const fadesim_batch_desc *desc, void *stream);

/* Duration (ms) of the solve kernels of the last fadesim_solve_batch call
 * made with FADESIM_F_TIMING on this thread, from HIP events on the call's
 * stream; waits for them.  Negative when no timed call was made. */
double fadesim_last_kernel_ms(void);

/* Of that call, the duration (ms) of its guess-first kernel alone (the part of
 * a FADESIM_F_TEST_FIRST solve before its first sweep); negative when the call
 * did not run one.  No reference counterpart: a timing query. */
double fadesim_last_guess_ms(void);

/* The built-in relaxation factor for an n x m array (used when omega <= 0). */
double fadesim_default_omega(int rows, int cols);

/*
 * Sinh-nonlinear cell conductance for a crossbar state, the G update of the
 * Picard iteration around fadesim_solve_batch.  For each instance b and cell
 * (i,j): V = (vin_i - p_ij) - (vout_j + q_ij) with the drops (p, q) of
 * d_field (NULL = zero drops), |V| clamped to >= 1e-5,
 *   R = (1-z)|V|/(b_on sinh(a_on|V|)) + z|V|/(b_off sinh(a_off|V|)),
 *   z = max(0, (1/g_lin - Ron)/(Roff - Ron)),  d_g = 1/R  (FP64, stored FP32).
 * params (host) = {Ron, Roff, a_on, b_on, a_off, b_off}.  The cell model of
 *   deal_v / get_nonlinear_ratio_z / nonlinear_R_tensor
 *   pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/split_arr_matmul.cu:142-196
 * d_glin is [B][n][m], or [S][n][m] slots with d_gidx [B]; d_g is [B][n][m].
 */
int fadesim_cell_conductance(int batch, int rows, int cols, const float *d_glin, const int *d_gidx,
                             const double *d_vin, const double *d_vout, const float *d_field,
                             const double *params, float *d_g, void *stream);

/*
 * Mixed-size batched solve: crossbars of 64x64, 128x128, 256x256 and 512x512
 * cells, in any order, in one persistent launch (same method and stopping
 * rule as fadesim_solve_batch; 256x256 and 512x512 crossbars are solved by
 * teams of workgroups, one strip of rows each).  Replaces, for uneven sizes,
 * the per-crossbar sweep loops named at fadesim_solve_batch.
 *
 * A plan fixes the sizes, the queues and the per-size solve controls and owns
 * all device memory the solve needs (allocated at create), so
 * fadesim_ragged_solve allocates nothing and can be captured into a HIP graph.
 * Packed layout, instance k in input order:
 *   d_g, d_warm, d_field: conductances [rows][cols] at g_off[k] = sum of the
 *     earlier rows*cols (d_warm / d_field: [2][rows][cols] at 2*g_off[k]);
 *   d_vin at sum of the earlier rows; d_vout, d_iout at sum of the earlier cols.
 * Calls with one plan are stream-ordered and must not overlap (the plan's
 * queues and exchange buffers are reused).
 */
typedef struct fadesim_ragged_plan fadesim_ragged_plan;

/* Per-size solve controls, index 0..3 = 64, 128, 256, 512; 0 = built-in. */
typedef struct fadesim_ragged_class_opts {
    double omega;     /* SOR factor in (0,2) */
    int check_first;  /* sweeps before the first convergence test */
    int check_min;    /* bounds of the predicted interval between tests */
    int check_max;
    int ir2s_iters;   /* IR2S iterations on the initial guess (0 = none) */
    int corr_iters;   /* IR2S iterations of the refinement correction's guess (0 = built-in, 2) */
    double ir2s_omega; /* damping of the IR2S iterations in (0, 1] (0 = built-in, 1), or the
                        * Chebyshev schedule's spectral scale, negative (fadesim_batch_desc) */
} fadesim_ragged_class_opts;

typedef struct fadesim_ragged_plan_stats {
    int count;
    long long cells, rows_total, cols_total;
    int grid, teams, big_tasks, small_items;
    double omega[4];
    int check_first[4], check_min[4], check_max[4];
    int ir2s_iters[4], corr_iters[4];
    double ir2s_omega[4];
} fadesim_ragged_plan_stats;

typedef struct fadesim_ragged_desc {
    const float *d_g;        /* packed conductances (S) */
    const double *d_r_wire;  /* [K] wire segment resistance (ohm) */
    const double *d_vin;     /* packed wordline drives (V) */
    const double *d_vout;    /* packed bitline sink voltages, or NULL = 0 V */
    const float *d_warm;     /* packed initial drops, or NULL = 0 */
    double tol, safety, switch_tol; /* as in fadesim_batch_desc */
    int max_sweeps;
    unsigned flags;          /* FADESIM_F_* */
    double *d_iout;          /* packed bitline currents (A) */
    int *d_sweeps;           /* [K] or NULL */
    int *d_status;           /* [K] or NULL */
    double *d_est;           /* [K] or NULL */
    float *d_field;          /* packed final drops, or NULL */
} fadesim_ragged_desc;

/* rows, cols: host arrays [count].  opts: host array [4] or NULL. */
int fadesim_ragged_plan_create(fadesim_ragged_plan **plan, int count, const int *rows, const int *cols,
                               const fadesim_ragged_class_opts *opts);
void fadesim_ragged_plan_destroy(fadesim_ragged_plan *plan);
int fadesim_ragged_plan_info(const fadesim_ragged_plan *plan, fadesim_ragged_plan_stats *stats);
int fadesim_ragged_solve(const fadesim_ragged_plan *plan, const fadesim_ragged_desc *desc, void *stream);
/* 1 in *aborted when the last solve's team waits timed out (results invalid:
 * every instance the launch did not finish has status FADESIM_ST_ABORTED);
 * synchronises with the device. */
int fadesim_ragged_aborted(const fadesim_ragged_plan *plan, int *aborted);
/* Kernel duration (ms) of the plan's last solve made with FADESIM_F_TIMING. */
double fadesim_ragged_last_ms(const fadesim_ragged_plan *plan);

/*
 * One crossbar, FP64 SOR in node voltages with per-segment wire resistances
 * and the reference's stopping rule (stop after the first sweep whose
 * max |dv| < tol; fixed iter_number sweeps without a break).  The solver
 * behind the `sim` CLI command
 *   nodebasedGSMethod <iter> <omega> <mode> [yes <tol>]
 * replacing NodebasedGSMethod::nodeBaseGS
 *   crossbarHspice/V2/concrete_command.h:1136-1361 (command :1097-1127)
 * and array2D::nodeBasedGS  crossbarHspice/2dArray.cc:1377-1687.
 * order FADESIM_GS64_LEXICOGRAPHIC (0, the default) is the reference's own
 * sweep order and arithmetic: the same iterates, break sweep and output bits
 * (rows <= 512); FADESIM_GS64_REDBLACK colours the nodes instead: the same
 * fixed point, different iterates.
 */
#define FADESIM_GS64_LEXICOGRAPHIC 0
#define FADESIM_GS64_REDBLACK 1
typedef struct fadesim_gs64_desc {
    size_t struct_size;    /* sizeof(fadesim_gs64_desc) as the caller compiled it */
    int rows, cols;        /* n, m >= 2 */
    const double *d_g;     /* [n][m] cell conductance (S) */
    const double *d_ru;    /* [n][ru_ld] wordline segments (ohm), j in [0, m] */
    int ru_ld;             /* >= m + 1 */
    const double *d_rd;    /* [m][rd_ld] bitline segments (ohm), i in [0, n] */
    int rd_ld;             /* >= n + 1 */
    const double *d_vin;   /* [n] left (wordline) drive (V) */
    const double *d_vout;  /* [m] bottom (bitline) sink voltage (V) */
    int iter_number;       /* sweep budget */
    double omega;          /* SOR factor */
    int enable_break;      /* reference break rule on/off */
    double tol;            /* break threshold on max |dv| (V) */
    double *d_vup;         /* [n][m] out: wordline node voltages */
    double *d_vdown;       /* [n][m] out: bitline node voltages */
    double *d_iout;        /* [m] out: sum_i (vup - vdown) g */
    int *d_record;         /* out: 0-based sweep of the break, or -1 */
    int *d_sweeps;         /* out: sweeps run */
    int general;           /* 1: general boundary, replacing array2D::nodeBasedGeneralGS
                            * (crossbarHspice/2dArray.cc:1209-1374): each line end driven
                            * or open by the flags below; initial u = vin on driven rows
                            * (0 elsewhere), d = 0; d_iout = sink wire currents
                            * (d_n-1,j - vout_j) / rd[j][n] of driven bottoms, 0 elsewhere */
    const int *d_on_top, *d_on_left, *d_on_bottom, *d_on_right; /* general: [m], [n], [m], [n] */
    const double *d_v_top;   /* general: [m] top drive voltages (left = d_vin, bottom = d_vout) */
    const double *d_v_right; /* general: [n] right drive voltages */
    int order;               /* FADESIM_GS64_LEXICOGRAPHIC or FADESIM_GS64_REDBLACK */
    const double *d_init_up;   /* [n][m] initial wordline node voltages, or NULL (the reference's
                                * initial state); with d_init_down: V1 mode bit 0's guess */
    const double *d_init_down; /* [n][m] initial bitline node voltages, or NULL */
} fadesim_gs64_desc;

int fadesim_gs64_solve(const fadesim_gs64_desc *desc, void *stream);

/*
 * The PimTorch hook, with the reference's iteration: line Gauss-Seidel/SOR,
 * all wordlines left->right then all bitlines bottom->top per sweep, source
 * and sink conductance g_wire, initial state u = input, d = 0; a fixed
 * iter_times, or a break when no node of the whole batch moved by more than
 * break_threshold; sinh-nonlinear cells by the lagged-conductance Picard
 * update.  Replaces ir_drop_gs and its kernels
 *   pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/split_arr_matmul.cu:27-140, 204-333
 * (pybind11 entry :1431).  dtype_bytes selects float (4) or double (8) for
 * input, G and out.  Shapes: input [batch][numX][row], linear G
 * [numX][row][numY*col]; out [batch][numX][numY*col] currents, or
 * [batch][numX][numY][row][col] voltage differences with ret_v.
 */
typedef struct fadesim_hook_desc {
    int dtype_bytes;
    const void *d_input;
    const void *d_G;
    int batch_size, row_size, col_size, numX, numY;
    int iter_times;
    double g_wire, beta;
    int enable_break;
    double break_threshold;
    int nonlinear;
    double nonlinear_params[6]; /* Ron, Roff, a_on, b_on, a_off, b_off */
    int ret_v;
    void *d_out;
    int *d_sweeps; /* out (device): sweeps run, or NULL */
    double g_load_left, g_load_down; /* source / sink load conductances (<= 0: g_wire),
                                      * ir_gs_ou_local (split_arr_matmul.cu:690-815) */
    int ou_rows, ou_cols; /* > 0: OU mode, ir_gs_ou (split_arr_matmul.cu:1086-1327): every
                           * ou_rows x ou_cols block solved as its own crossbar, source load
                           * g_wire/(c0+1), sink load g_wire/(row_size - last row); out
                           * [batch][numX][row_size/ou_rows][numY*col] currents */
} fadesim_hook_desc;

int fadesim_hook_ir_drop_gs(const fadesim_hook_desc *desc, void *stream);

/*
 * IR2S with the reference's own semantics: every wordline and bitline solved
 * exactly for the cell currents of the present state, both layers from the
 * same currents, initial state u = vin, d = 0; a fixed iter_times, or a break
 * rule; sinh-nonlinear cells by the lagged-conductance update.  Replaces
 *   ir2s (PimTorch hook)  pimtorch/pimfixedpoint/fixedPoint/SplitArrCuda/
 *                         split_arr_matmul.cu:336-545 (pybind11 entry :1432)
 *   ir_drop_fastSolve_cpu_singleMat  crossbarHspice/ir_drop_fast_solver.h:24-163
 *                         behind the V1 command `fastsolve` (2dArray.cc:318-327, 1691-1900)
 * The hook's arithmetic: layout FADESIM_IR2S_HOOK, divide_by_g 0,
 * carry_residue 1, r_source = r_sink = r_wire, vout NULL, break
 * FADESIM_IR2S_BREAK_BATCH.  The CPU solver's: layout FADESIM_IR2S_PLAIN,
 * divide_by_g 1, carry_residue 0, break FADESIM_IR2S_BREAK_EACH.
 */
#define FADESIM_IR2S_HOOK 0   /* vin [batch][numX][row], G [numX][row][numY*col] shared by the batch */
#define FADESIM_IR2S_PLAIN 1  /* vin [batch][row], G [batch][row][col]; numX = numY = 1 */
#define FADESIM_IR2S_BREAK_BATCH 1 /* stop after the first iteration in which no node of the
                                    * whole call moved by more than break_threshold */
#define FADESIM_IR2S_BREAK_EACH 2  /* per crossbar: stop after the first iteration whose largest
                                    * move is below break_threshold */
typedef struct fadesim_ir2s_desc {
    int dtype_bytes;        /* 4 (float) or 8 (double): vin, G, vout, out and the fields */
    int layout;             /* FADESIM_IR2S_HOOK or FADESIM_IR2S_PLAIN */
    const void *d_vin;
    const void *d_G;
    const void *d_vout;     /* [crossbars][col] bitline sink voltages, or NULL = 0 V */
    int batch_size, row_size, col_size, numX, numY; /* crossbars = batch_size * numX * numY */
    int iter_times;
    double r_wire;          /* inner wire segments (ohm) */
    double r_source, r_sink; /* first wordline / last bitline segment; <= 0 -> r_wire */
    int divide_by_g;        /* inner drop as I / (1/r) (CPU solver) instead of I * r (hook) */
    int carry_residue;      /* keep each line's rounding residue of the running current sum
                             * across iterations (the hook's never-reset I_row_sum) */
    int enable_break;       /* 0, FADESIM_IR2S_BREAK_BATCH or FADESIM_IR2S_BREAK_EACH */
    double break_threshold; /* volts */
    int nonlinear;
    double nonlinear_params[6]; /* Ron, Roff, a_on, b_on, a_off, b_off */
    int ret_v;              /* out: [crossbars][row][col] u - d instead of currents [crossbars][col] */
    void *d_out;
    void *d_vup, *d_vdown;  /* optional [crossbars][row][col] final node voltages, or NULL */
    int *d_iters;           /* optional (device): BREAK_EACH: [crossbars] counts as the CPU solver
                             * returns them (break sweep + 1, or iter_times + 1 when never met);
                             * otherwise one int: iterations run */
    int ou_rows, ou_cols;   /* > 0 (hook layout): OU mode, ir2s_ou (split_arr_matmul.cu:816-1085):
                             * every ou_rows x ou_cols block solved as its own crossbar, source
                             * r_wire (c0+1), sink r_wire (row_size - last row); out
                             * [batch][numX][row_size/ou_rows][numY*col] currents */
} fadesim_ir2s_desc;

int fadesim_ir2s(const fadesim_ir2s_desc *desc, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FADESIM_H */
