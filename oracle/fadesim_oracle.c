/*
 * CPU oracle for the FADESim IR-drop solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fadesim_amd/, include/)
 * links, loads or calls this file; only tests/, the smoke check and the CPU
 * baseline leg of bench.py do, and there only as the checker.
 *
 * Contents
 *   oracle_nodegs      lexicographic point SOR, a statement-by-statement
 *                      restatement of NodebasedGSMethod::nodeBaseGS
 *                      (crossbarHspice/V2/concrete_command.h:1136-1361) and
 *                      array2D::nodeBasedGS (crossbarHspice/2dArray.cc:1377-1687).
 *                      Pinned bitwise against golden vectors written by the
 *                      reference `sim` binaries (tests/golden/).
 *   oracle_scichina_guess, oracle_nodegs_init  V1 mode bit 0: the initial
 *                      state of crossbarHspice/2dArray.cc:1425-1535 and the same
 *                      SOR from it; pinned against golden vectors of `sim_v1`.
 *   oracle_exact       exact FP64 solve of the same linear network by banded
 *                      Cholesky (the matrix is SPD).  Pinned against
 *                      scipy.sparse.linalg.spsolve, 2x2 closed forms and the
 *                      KCL/monotonicity invariants (tests/test_oracle.py).
 *   oracle_exact_nonlinear_batch  sinh-nonlinear cells: Picard iteration of
 *                      the hook's cell model around exact solves, to the
 *                      nonlinear fixed point; pinned against the converged
 *                      hook replay and a Newton solve (tests/test_oracle.py).
 *   oracle_exact_slots_f32g  the same solve for many drives over shared
 *                      conductance slots (one factorisation per slot); the
 *                      checker of the PimTorch-level crossbar currents.
 *   oracle_errisum     the partial-KCL "error I sum" of concrete_command.h:1293-1319.
 *   oracle_rb_f64      FP64 red-black SOR replay of the GPU drop-in kernel
 *                      (reference stopping rule max|dv| < tol).
 *   oracle_rb_f32      FP32 red-black SOR replay of the GPU batched kernel on
 *                      drop variables (fixed sweep count, for bitwise checks).
 *   oracle_hook_gs_*   the PimTorch hook's line-GS ordering
 *                      (pimtorch/.../SplitArrCuda/split_arr_matmul.cu:27-140,204-333),
 *                      linear and sinh-nonlinear cells, FP32 and FP64.
 *   oracle_nodegs_general  general-boundary lexicographic SOR, a statement-by-
 *                      statement restatement of array2D::nodeBasedGeneralGS
 *                      (crossbarHspice/2dArray.cc:1209-1374): every edge driven
 *                      or open; pinned against golden vectors of the
 *                      reference `sim_v1` binary (tests/golden/).
 *   oracle_exact_general  exact FP64 solve of the general-boundary network
 *                      (the same banded Cholesky); pinned against scipy and
 *                      against oracle_exact where the two networks coincide.
 *   oracle_ir2s        the IR2S wire-solve iteration, a statement-by-statement
 *                      restatement of ir_drop_fastSolve_cpu_singleMat
 *                      (crossbarHspice/ir_drop_fast_solver.h:24-163), behind
 *                      the V1 `fastsolve` command (2dArray.cc:318-327,
 *                      1691-1900).  Pinned bitwise against golden vectors of
 *                      the reference `sim_v1` binary (tests/golden/).
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so that no multiply-add is
 * fused: the reference binary is built for baseline x86-64 which has no FMA).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, m) ((size_t)(i) * (size_t)(m) + (size_t)(j))

/* ------------------------------------------------------------------------ */
/* Lexicographic SOR (reference CLI solver)                                  */
/* ------------------------------------------------------------------------ */

/*
 * RU: wordline segment resistances, RU[i*ru_ld + j], j in [0, m]; segment 0
 *     joins the left source to node (i,0), segment j joins (i,j-1)-(i,j).
 * RD: bitline segment resistances, RD[j*rd_ld + i], i in [0, n]; segment i
 *     joins (i-1,j)-(i,j), segment n joins (n-1,j) to the bottom sink.
 * V2 passes RD for both (concrete_command.h:1214 aliases RULine to RDLine).
 * g:  cell conductances n*m (the reference computes 1.0/cellR_states[state],
 *     concrete_command.h:1163-1166; callers pass that same quotient).
 * Returns 0, or -1 if n<2 or m<2 (concrete_command.h:1199-1202).
 * vup/vdown/iout are the final iterate and the readout sum_i (vup-vdown)*g
 * (concrete_command.h:1344-1348); *record is the 0-based break sweep or -1.
 */
int oracle_nodegs_init(int n, int m, const double *RU, int ru_ld, const double *RD,
                       int rd_ld, const double *g, const double *vin,
                       const double *vout, const double *init_up, const double *init_down,
                       int iter_number, double w, int enable_break, double tol, double *vup_out,
                       double *vdown_out, double *iout, int *record_out);

int oracle_nodegs(int n, int m, const double *RU, int ru_ld, const double *RD,
                  int rd_ld, const double *g, const double *vin,
                  const double *vout, int iter_number, double w,
                  int enable_break, double tol, double *vup_out,
                  double *vdown_out, double *iout, int *record_out)
{
    return oracle_nodegs_init(n, m, RU, ru_ld, RD, rd_ld, g, vin, vout, NULL, NULL, iter_number, w,
                              enable_break, tol, vup_out, vdown_out, iout, record_out);
}

/*
 * The initial state of V1 mode bit 0 ("sci_china way", crossbarHspice/
 * 2dArray.cc:1425-1535): equivalent resistances of the bitline ladders above
 * and below each cell, the wordline voltages of a ladder of those, and the
 * bitline voltages under them.  r is the wire resistance RDLine[0][0] (the
 * reference uses g_wire = 1/RDLine[0][0] for every segment here).
 */
void oracle_scichina_guess(int M, int N, double r_seg, const double *g, const double *vin,
                           double *vup, double *vdown)
{
    const double g_wire = 1.0 / r_seg;
    const double r_wire = 1 / g_wire, Rs = 1 / g_wire;
    const size_t nm = (size_t)M * N;
    double *R = malloc(nm * sizeof(double)), *RR = malloc(nm * sizeof(double));
    double *RRR = malloc(nm * sizeof(double)), *Vwl = malloc(nm * sizeof(double));
    double *Rup = malloc(nm * sizeof(double)), *Rdn = malloc(nm * sizeof(double));
    double *Req = malloc(nm * sizeof(double)), *gs = calloc(N, sizeof(double));
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            R[IDX(i, j, N)] = 1.0 / g[IDX(i, j, N)];
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i)
            gs[j] += g[IDX(i, j, N)];
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < N; ++j)
            RR[IDX(k, j, N)] = R[IDX(k, j, N)] + (M - k) * r_wire + (gs[j] * R[IDX(k, j, N)]) * Rs;
    for (int k = 0; k < M; ++k) {
        RRR[IDX(k, N - 1, N)] = RR[IDX(k, N - 1, N)] + r_wire;
        for (int j = N - 2; j >= 0; --j) {
            const double a = RR[IDX(k, j, N)], b = RRR[IDX(k, j + 1, N)];
            RRR[IDX(k, j, N)] = (a * b) / (a + b) + r_wire;
        }
    }
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < N; ++j) {
            const double x = RRR[IDX(k, j, N)];
            Vwl[IDX(k, j, N)] = (x - r_wire) / x * (j == 0 ? vin[k] : Vwl[IDX(k, j - 1, N)]);
        }
#define PAR(x, y) ((x) * (y) / ((x) + (y)))
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            Rup[IDX(i, j, N)] = i == 0 ? r_wire + R[IDX(i, j, N)] : PAR(Rup[IDX(i - 1, j, N)], R[IDX(i, j, N)]) + r_wire;
    for (int i = M - 1; i >= 0; --i)
        for (int j = 0; j < N; ++j)
            Rdn[IDX(i, j, N)] = i != M - 1 ? r_wire + PAR(Rdn[IDX(i + 1, j, N)], R[IDX(i + 1, j, N)]) : Rs;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            Req[IDX(i, j, N)] = i == 0 ? Rdn[IDX(i, j, N)] : PAR(Rup[IDX(i - 1, j, N)], Rdn[IDX(i, j, N)]);
#undef PAR
    for (size_t k = 0; k < nm; ++k) {
        vup[k] = Vwl[k];
        vdown[k] = Vwl[k] * (Req[k]) / (Req[k] + R[k]);
    }
    free(R); free(RR); free(RRR); free(Vwl); free(Rup); free(Rdn); free(Req); free(gs);
}

/* oracle_nodegs from a given initial state (NULL: the reference's u = vin,
 * d = vout), as V1 mode bit 0 starts from oracle_scichina_guess. */
int oracle_nodegs_init(int n, int m, const double *RU, int ru_ld, const double *RD,
                       int rd_ld, const double *g, const double *vin,
                       const double *vout, const double *init_up, const double *init_down,
                       int iter_number, double w, int enable_break, double tol, double *vup_out,
                       double *vdown_out, double *iout, int *record_out)
{
    if (!(m >= 2 && n >= 2))
        return -1;
    size_t nm = (size_t)n * m;
    double *vup[2], *vdown[2];
    vup[0] = (double *)malloc(nm * sizeof(double));
    vup[1] = (double *)malloc(nm * sizeof(double));
    vdown[0] = (double *)malloc(nm * sizeof(double));
    vdown[1] = (double *)malloc(nm * sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            vup[1][IDX(i, j, m)] = vup[0][IDX(i, j, m)] = init_up ? init_up[IDX(i, j, m)] : vin[i];
            vdown[1][IDX(i, j, m)] = vdown[0][IDX(i, j, m)] = init_down ? init_down[IDX(i, j, m)] : vout[j];
        }
#define RUL(i, j) RU[(size_t)(i) * ru_ld + (j)]
#define RDL(j, i) RD[(size_t)(j) * rd_ld + (i)]
    int record = -1;
    for (int iter = 0; iter < iter_number; ++iter) {
        int x = iter & 1, y = 1 - x;
        double *ux = vup[x], *uy = vup[y], *dx = vdown[x], *dy = vdown[y];
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < m; ++j) {
                double sumg;
                double l = (j == 0 ? vin[i] : ux[IDX(i, j - 1, m)]), r,
                       o = dy[IDX(i, j, m)];
                double gij = g[IDX(i, j, m)];
                if (j == m - 1) {
                    sumg = 1.0 / RUL(i, j) + gij;
                    ux[IDX(i, j, m)] = w * (l / RUL(i, j) + gij * o) / sumg +
                                       (1 - w) * uy[IDX(i, j, m)];
                } else {
                    r = uy[IDX(i, j + 1, m)];
                    sumg = 1.0 / RUL(i, j) + 1.0 / RUL(i, j + 1) + gij;
                    ux[IDX(i, j, m)] =
                        w * (l / RUL(i, j) + r / RUL(i, j + 1) + gij * o) / sumg +
                        (1 - w) * uy[IDX(i, j, m)];
                }
                r = (i == n - 1 ? vout[j] : dy[IDX(i + 1, j, m)]);
                o = ux[IDX(i, j, m)];
                if (i == 0) {
                    sumg = 1.0 / RDL(j, 1) + gij;
                    dx[IDX(i, j, m)] = w * (r / RDL(j, 1) + gij * o) / sumg +
                                       (1 - w) * dy[IDX(i, j, m)];
                } else {
                    l = dx[IDX(i - 1, j, m)];
                    sumg = 1.0 / RDL(j, i) + 1.0 / RDL(j, i + 1) + gij;
                    dx[IDX(i, j, m)] =
                        w * (l / RDL(j, i) + r / RDL(j, i + 1) + gij * o) / sumg +
                        (1 - w) * dy[IDX(i, j, m)];
                }
            }
        }
        if (enable_break) {
            double mx = 0;
            for (size_t k = 0; k < nm; ++k) {
                mx = fmax(mx, fabs(dx[k] - dy[k]));
                mx = fmax(mx, fabs(ux[k] - uy[k]));
            }
            if (mx < tol) {
                record = iter;
                break;
            }
        }
    }
#undef RUL
#undef RDL
    int final_i = 1 - (iter_number & 1);
    if (record != -1)
        final_i = record & 1;
    memcpy(vup_out, vup[final_i], nm * sizeof(double));
    memcpy(vdown_out, vdown[final_i], nm * sizeof(double));
    for (int j = 0; j < m; ++j)
        iout[j] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j)
            iout[j] += (vup_out[IDX(i, j, m)] - vdown_out[IDX(i, j, m)]) *
                       g[IDX(i, j, m)];
    *record_out = record;
    free(vup[0]);
    free(vup[1]);
    free(vdown[0]);
    free(vdown[1]);
    return 0;
}

/* Partial-KCL residual sum printed as "error I sum"
 * (concrete_command.h:1293-1319; 2dArray.cc:1620-1645). */
double oracle_errisum(int n, int m, double g_wire, const double *g,
                      const double *vup, const double *vdown)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < m - 1; ++j)
            s += fabs((vup[IDX(i, j - 1, m)] - vup[IDX(i, j, m)] +
                       vup[IDX(i, j + 1, m)] - vup[IDX(i, j, m)]) * g_wire +
                      (vdown[IDX(i, j, m)] - vup[IDX(i, j, m)]) * g[IDX(i, j, m)]);
        if (m > 1)
            s += fabs((vup[IDX(i, m - 2, m)] - vup[IDX(i, m - 1, m)]) * g_wire +
                      (vdown[IDX(i, m - 1, m)] - vup[IDX(i, m - 1, m)]) *
                          g[IDX(i, m - 1, m)]);
    }
    for (int j = 0; j < m; ++j)
        s += fabs((vup[IDX(0, j, m)] - vdown[IDX(0, j, m)]) * g[IDX(0, j, m)] +
                  (n > 1 ? vdown[IDX(1, j, m)] - vdown[IDX(0, j, m)] : 0) * g_wire);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 0; j < m; ++j)
            s += fabs((vup[IDX(i, j, m)] - vdown[IDX(i, j, m)]) * g[IDX(i, j, m)] +
                      (vdown[IDX(i - 1, j, m)] - vdown[IDX(i, j, m)] +
                       vdown[IDX(i + 1, j, m)] - vdown[IDX(i, j, m)]) * g_wire);
    return s;
}

/* ------------------------------------------------------------------------ */
/* Exact FP64 solve: banded Cholesky                                         */
/* ------------------------------------------------------------------------ */
/*
 * Unknown k = 2*(i*m+j)+t, t=0 wordline node u_ij, t=1 bitline node d_ij.
 * KCL (the equations nodeBaseGS relaxes):
 *   (g + a_j + [j<m-1]a_{j+1}) u_ij - a_j u_i,j-1 - [j<m-1]a_{j+1} u_i,j+1 - g d_ij
 *        = [j=0] a_0 vin_i                       (a_j = 1/RU[i][j], u_i,-1 = vin_i)
 *   (g + [i>0]b_i + b_{i+1}) d_ij - [i>0]b_i d_i-1,j - [i<n-1]b_{i+1} d_i+1,j - g u_ij
 *        = [i=n-1] b_n vout_j                    (b_i = 1/RD[j][i])
 * vout may be NULL (all bitlines grounded).  Half bandwidth kd = 2m.  Row-major lower band: L[k*(kd+1) + (kd - (k-c))]
 * holds entry (k, c) for c in [k-kd, k].
 */
/* Assembles the banded SPD matrix (lower band, width kd = 2m) of the network
 * and factors it in place (A = L L^T).  Unknown order: u_ij at 2(i m + j),
 * d_ij at 2(i m + j) + 1.  Returns 0, or -3 if the matrix is not SPD. */
/* Edge drives of the general-boundary network: on[e] flags (NULL = the
 * nodeBaseGS network: left and bottom driven, top and right open), e = 0 top
 * (bitline ends above row 0, [m]), 1 left ([n]), 2 bottom ([m]), 3 right ([n]). */
static int edge_on(const int *const *on, int e, int k)
{
    if (!on || !on[e])
        return e == 1 || e == 2;
    return on[e][k] != 0;
}

static int band_factor_general(int n, int m, const double *RU, int ru_ld, const double *RD,
                               int rd_ld, const double *g, const int *const *on, double *L);

static int band_factor(int n, int m, const double *RU, int ru_ld, const double *RD,
                       int rd_ld, const double *g, double *L)
{
    return band_factor_general(n, m, RU, ru_ld, RD, rd_ld, g, NULL, L);
}

static int band_factor_general(int n, int m, const double *RU, int ru_ld, const double *RD,
                               int rd_ld, const double *g, const int *const *on, double *L)
{
    const long N = 2L * n * m;
    const long kd = 2L * m;
    const long W = kd + 1;
#define BAND(r, c) L[(size_t)(r) * W + (kd - ((r) - (c)))]
#define RUL(i, j) RU[(size_t)(i) * ru_ld + (j)]
#define RDL(j, i) RD[(size_t)(j) * rd_ld + (i)]
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            double gij = g[IDX(i, j, m)];
            long ku = 2L * (i * (long)m + j), kdn = ku + 1;
            double aj = (j > 0 || edge_on(on, 1, i)) ? 1.0 / RUL(i, j) : 0.0;
            double aj1 = (j < m - 1 || edge_on(on, 3, i)) ? 1.0 / RUL(i, j + 1) : 0.0;
            BAND(ku, ku) = gij + aj + aj1;
            if (j > 0)
                BAND(ku, ku - 2) = -aj;
            double bi = (i > 0 || edge_on(on, 0, j)) ? 1.0 / RDL(j, i) : 0.0;
            double bi1 = (i < n - 1 || edge_on(on, 2, j)) ? 1.0 / RDL(j, i + 1) : 0.0;
            BAND(kdn, kdn) = gij + bi + bi1;
            BAND(kdn, ku) = -gij;
            if (i > 0)
                BAND(kdn, kdn - 2L * m) = -bi;
        }
#undef RUL
#undef RDL
    for (long k = 0; k < N; ++k) {
        long c0 = k - kd < 0 ? 0 : k - kd;
        /* off-diagonal entries of row k */
        for (long c = c0; c < k; ++c) {
            long lo = (c - kd < 0 ? 0 : c - kd);
            if (lo < c0)
                lo = c0;
            double s = BAND(k, c);
            const double *rk = &BAND(k, lo);
            const double *rc = &BAND(c, lo);
            long len = c - lo;
#pragma omp simd reduction(- : s)
            for (long t = 0; t < len; ++t)
                s -= rk[t] * rc[t];
            BAND(k, c) = s / BAND(c, c);
        }
        double s = BAND(k, k);
        const double *rk = &BAND(k, c0);
        long len = k - c0;
#pragma omp simd reduction(- : s)
        for (long t = 0; t < len; ++t)
            s -= rk[t] * rk[t];
        if (!(s > 0))
            return -3;
        BAND(k, k) = sqrt(s);
    }
    return 0;
}

/* Solves L L^T x = b in place for the factor of band_factor. */
static void band_solve(int n, int m, const double *L, double *b)
{
    const long N = 2L * n * m;
    const long kd = 2L * m;
    const long W = kd + 1;
    for (long k = 0; k < N; ++k) {
        long c0 = k - kd < 0 ? 0 : k - kd;
        double s = b[k];
        for (long c = c0; c < k; ++c)
            s -= BAND(k, c) * b[c];
        b[k] = s / BAND(k, k);
    }
    for (long k = N - 1; k >= 0; --k) {
        b[k] /= BAND(k, k);
        double xk = b[k];
        long c0 = k - kd < 0 ? 0 : k - kd;
        for (long c = c0; c < k; ++c)
            b[c] -= BAND(k, c) * xk;
    }
#undef BAND
}

/* Right-hand side of the network for drives vin (left) and vout (bottom, NULL = 0). */
static void band_rhs(int n, int m, const double *RU, int ru_ld, const double *RD, int rd_ld,
                     const double *vin, const double *vout, double *b)
{
    for (long k = 0; k < 2L * n * m; ++k)
        b[k] = 0.0;
    for (int i = 0; i < n; ++i)
        b[2L * (i * (long)m)] += (1.0 / RU[(size_t)i * ru_ld]) * vin[i];
    if (vout)
        for (int j = 0; j < m; ++j)
            b[2L * ((n - 1) * (long)m + j) + 1] += (1.0 / RD[(size_t)j * rd_ld + n]) * vout[j];
}

/* Exact FP64 solve of the general-boundary network (nodeBasedGeneralGS,
 * crossbarHspice/2dArray.cc:1209-1374): edge e driven where on[e][k] != 0, at
 * the voltage v[e][k] through its end segment (top RD[j][0], left RU[i][0],
 * bottom RD[j][n], right RU[i][m]), open otherwise.  Readout as the reference:
 * iout_j = (d_{n-1,j} - v_bottom_j) / RD[j][n] for driven bottoms, else 0.
 * Returns 0, -2 (memory) or -3 (singular: a floating line). */
int oracle_exact_general(int n, int m, const double *RU, int ru_ld, const double *RD, int rd_ld, const double *g,
                         const int *on_top, const int *on_left, const int *on_bottom, const int *on_right,
                         const double *v_top, const double *v_left, const double *v_bottom, const double *v_right,
                         double *u_out, double *d_out, double *iout)
{
    const int *on[4] = {on_top, on_left, on_bottom, on_right};
    const long N = 2L * n * m;
    const long W = 2L * m + 1;
    double *L = (double *)calloc((size_t)N * W, sizeof(double));
    double *b = (double *)calloc((size_t)N, sizeof(double));
    if (!L || !b) {
        free(L);
        free(b);
        return -2;
    }
    int e = band_factor_general(n, m, RU, ru_ld, RD, rd_ld, g, on, L);
    if (e) {
        free(L);
        free(b);
        return e;
    }
    for (int i = 0; i < n; ++i) {
        if (on_left[i])
            b[2L * (i * (long)m)] += (1.0 / RU[(size_t)i * ru_ld]) * v_left[i];
        if (on_right[i])
            b[2L * (i * (long)m + m - 1)] += (1.0 / RU[(size_t)i * ru_ld + m]) * v_right[i];
    }
    for (int j = 0; j < m; ++j) {
        if (on_top[j])
            b[2L * j + 1] += (1.0 / RD[(size_t)j * rd_ld]) * v_top[j];
        if (on_bottom[j])
            b[2L * ((n - 1) * (long)m + j) + 1] += (1.0 / RD[(size_t)j * rd_ld + n]) * v_bottom[j];
    }
    band_solve(n, m, L, b);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            u_out[IDX(i, j, m)] = b[2 * IDX(i, j, m)];
            d_out[IDX(i, j, m)] = b[2 * IDX(i, j, m) + 1];
        }
    for (int j = 0; j < m; ++j)
        iout[j] = on_bottom[j] ? (d_out[IDX(n - 1, j, m)] - v_bottom[j]) / RD[(size_t)j * rd_ld + n] : 0.0;
    free(L);
    free(b);
    return 0;
}

int oracle_exact(int n, int m, const double *RU, int ru_ld, const double *RD,
                 int rd_ld, const double *g, const double *vin,
                 const double *vout, double *u_out, double *d_out, double *iout)
{
    if (n < 1 || m < 1)
        return -1;
    const long N = 2L * n * m;
    const long W = 2L * m + 1;
    double *L = (double *)calloc((size_t)N * W, sizeof(double));
    double *b = (double *)calloc((size_t)N, sizeof(double));
    if (!L || !b) {
        free(L);
        free(b);
        return -2;
    }
    int e = band_factor(n, m, RU, ru_ld, RD, rd_ld, g, L);
    if (e) {
        free(L);
        free(b);
        return e;
    }
    band_rhs(n, m, RU, ru_ld, RD, rd_ld, vin, vout, b);
    band_solve(n, m, L, b);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            u_out[IDX(i, j, m)] = b[2 * IDX(i, j, m)];
            d_out[IDX(i, j, m)] = b[2 * IDX(i, j, m) + 1];
        }
    for (int j = 0; j < m; ++j)
        iout[j] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j)
            iout[j] += (u_out[IDX(i, j, m)] - d_out[IDX(i, j, m)]) * g[IDX(i, j, m)];
    free(L);
    free(b);
    return 0;
}

/*
 * Sinh-nonlinear cells, converged: the fixed point of G <- G(V) with exact
 * FP64 solves of the network at each G (uniform wire resistance r[b], vout =
 * 0) -- an independent route to the state that the reference's lagged-
 * conductance sweeps (split_arr_matmul.cu:241-262) and the GPU's Picard
 * solve both converge to.  The cell model is the hook's (:142-196):
 *   V = u - d, |V| < 1e-5 clamped to 1e-5,
 *   R(V) = (1-z)|V|/(b_on sinh(a_on|V|)) + z|V|/(b_off sinh(a_off|V|)),
 *   z = max(0, (1/g_lin - Ron)/(Roff - Ron)),  G = 1/R,
 * params = {Ron, Roff, a_on, b_on, a_off, b_off}.  The first G is that of
 * u = vin, d = 0 (the hook's initial state).  Iterates until consecutive
 * readouts agree to tol relative in every column, at most max_iter solves.
 * glin [B][n][m] FP32, vin [B][n]; iout [B][m]; iters [B] = solves used, or
 * -1 (not converged), -2 (memory), -3 (singular).  OpenMP over instances.
 */
void oracle_exact_nonlinear_batch(int B, int n, int m, const float *glin, const double *r, const double *vin,
                                  const double *nl, double tol, int max_iter, double *iout, int *iters)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < B; ++b) {
        const size_t nm = (size_t)n * m;
        const long N = 2L * n * m, W = 2L * m + 1;
        double *RU = malloc((size_t)n * (m + 1) * sizeof(double));
        double *RD = malloc((size_t)m * (n + 1) * sizeof(double));
        double *z = malloc(nm * sizeof(double)), *G = malloc(nm * sizeof(double));
        double *u = malloc(nm * sizeof(double)), *d = malloc(nm * sizeof(double));
        double *L = malloc((size_t)N * W * sizeof(double)), *rhs = malloc((size_t)N * sizeof(double));
        double *io = iout + (size_t)b * m, *prev = malloc((size_t)m * sizeof(double));
        const float *gl = glin + (size_t)b * nm;
        const double *vi = vin + (size_t)b * n;
        int status = -2;
        if (RU && RD && z && G && u && d && L && rhs && prev) {
            for (size_t k = 0; k < (size_t)n * (m + 1); ++k)
                RU[k] = r[b];
            for (size_t k = 0; k < (size_t)m * (n + 1); ++k)
                RD[k] = r[b];
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < m; ++j) {
                    const double zz = (1.0 / (double)gl[IDX(i, j, m)] - nl[0]) / (nl[1] - nl[0]);
                    z[IDX(i, j, m)] = zz < 0 ? 0 : zz;
                    u[IDX(i, j, m)] = vi[i];
                    d[IDX(i, j, m)] = 0.0;
                }
            status = -1;
            for (int it = 1; it <= max_iter; ++it) {
                for (size_t k = 0; k < nm; ++k) {
                    double V = u[k] - d[k];
                    if (V < 1e-5 && V > -1e-5)
                        V = V > 0 ? 1e-5 : -1e-5;
                    V = fabs(V);
                    const double R_ = (1.0 - z[k]) * V / (nl[3] * sinh(V * nl[2])) + z[k] * V / (nl[5] * sinh(V * nl[4]));
                    G[k] = 1.0 / R_;
                }
                memset(L, 0, (size_t)N * W * sizeof(double));
                if (band_factor(n, m, RU, m + 1, RD, n + 1, G, L)) {
                    status = -3;
                    break;
                }
                band_rhs(n, m, RU, m + 1, RD, n + 1, vi, NULL, rhs);
                band_solve(n, m, L, rhs);
                for (size_t k = 0; k < nm; ++k) {
                    u[k] = rhs[2 * k];
                    d[k] = rhs[2 * k + 1];
                }
                for (int j = 0; j < m; ++j)
                    io[j] = 0.0;
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < m; ++j)
                        io[j] += (u[IDX(i, j, m)] - d[IDX(i, j, m)]) * G[IDX(i, j, m)];
                if (it > 1) {
                    double ch = 0.0;
                    for (int j = 0; j < m; ++j) {
                        const double den = fabs(io[j]) > 0 ? fabs(io[j]) : 1e-300;
                        const double c = fabs(io[j] - prev[j]) / den;
                        ch = c > ch ? c : ch;
                    }
                    if (ch <= tol) {
                        status = it;
                        break;
                    }
                }
                memcpy(prev, io, (size_t)m * sizeof(double));
            }
        }
        iters[b] = status;
        free(RU); free(RD); free(z); free(G); free(u); free(d); free(L); free(rhs); free(prev);
    }
}

/* Exact solves of K drives over S shared FP32 conductance slots with uniform
 * wire resistance r[s]: instance k uses slot[k] and vin[k][n] (vout = 0).
 * One factorisation per slot, OpenMP over slots.  Output iout[k*m + j]. */
int oracle_exact_slots_f32g(int S, int n, int m, const float *g32, const double *r,
                            int K, const int *slot, const double *vin, double *iout)
{
    int err = 0;
    const long N = 2L * n * m;
    const long W = 2L * m + 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int s = 0; s < S; ++s) {
        size_t nm = (size_t)n * m;
        double *RU = (double *)malloc((size_t)n * (m + 1) * sizeof(double));
        double *RD = (double *)malloc((size_t)m * (n + 1) * sizeof(double));
        double *g = (double *)malloc(nm * sizeof(double));
        double *L = (double *)calloc((size_t)N * W, sizeof(double));
        double *b = (double *)malloc((size_t)N * sizeof(double));
        if (!RU || !RD || !g || !L || !b) {
            err |= 1;
        } else {
            for (size_t t = 0; t < (size_t)n * (m + 1); ++t)
                RU[t] = r[s];
            for (size_t t = 0; t < (size_t)m * (n + 1); ++t)
                RD[t] = r[s];
            for (size_t t = 0; t < nm; ++t)
                g[t] = (double)g32[(size_t)s * nm + t];
            if (band_factor(n, m, RU, m + 1, RD, n + 1, g, L)) {
                err |= 1;
            } else {
                for (int k = 0; k < K; ++k) {
                    if (slot[k] != s)
                        continue;
                    band_rhs(n, m, RU, m + 1, RD, n + 1, vin + (size_t)k * n, NULL, b);
                    band_solve(n, m, L, b);
                    double *io = iout + (size_t)k * m;
                    for (int j = 0; j < m; ++j)
                        io[j] = 0;
                    for (int i = 0; i < n; ++i)
                        for (int j = 0; j < m; ++j)
                            io[j] += (b[2 * IDX(i, j, m)] - b[2 * IDX(i, j, m) + 1]) * g[IDX(i, j, m)];
                }
            }
        }
        free(RU);
        free(RD);
        free(g);
        free(L);
        free(b);
    }
    return err ? -1 : 0;
}

/* Batched exact solve for uniform wire resistance r[k] per instance, FP32
 * conductances (the batched kernel's input format), OpenMP over instances.
 * Output: iout[k*m + j]. */
int oracle_exact_batch_f32g(int B, int n, int m, const float *g32,
                            const double *r, const double *vin,
                            const double *vout, double *iout)
{
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int k = 0; k < B; ++k) {
        size_t nm = (size_t)n * m;
        double *RU = (double *)malloc((size_t)n * (m + 1) * sizeof(double));
        double *RD = (double *)malloc((size_t)m * (n + 1) * sizeof(double));
        double *g = (double *)malloc(nm * sizeof(double));
        double *u = (double *)malloc(nm * sizeof(double));
        double *d = (double *)malloc(nm * sizeof(double));
        for (size_t t = 0; t < (size_t)n * (m + 1); ++t)
            RU[t] = r[k];
        for (size_t t = 0; t < (size_t)m * (n + 1); ++t)
            RD[t] = r[k];
        for (size_t t = 0; t < nm; ++t)
            g[t] = (double)g32[(size_t)k * nm + t];
        int e = oracle_exact(n, m, RU, m + 1, RD, n + 1, g, vin + (size_t)k * n,
                             vout ? vout + (size_t)k * m : NULL, u, d,
                             iout + (size_t)k * m);
        if (e)
            err |= 1;
        free(RU);
        free(RD);
        free(g);
        free(u);
        free(d);
    }
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 red-black SOR: replay of the GPU drop-in kernel (rb_f64)             */
/* ------------------------------------------------------------------------ */
/*
 * Colours: c(u_ij) = (i+j)&1, c(d_ij) = (i+j+1)&1; half-sweep 0 then 1.
 * Node update (no fused multiply-add; same order on the GPU):
 *   u: num = (aL*L + aR*R) + g*d ; u' = wD*num + (1-w)*u
 *      aL = 1/RU[i][j], aR = [j<m-1] 1/RU[i][j+1], L = j? u_i,j-1 : vin_i,
 *      R = j<m-1 ? u_i,j+1 : 0, wD = w/((aL+aR)+g)
 *   d: num = (bU*U + bD*Dn) + g*u ; d' = wD*num + (1-w)*d
 *      bU = [i>0] 1/RD[j][i], bD = 1/RD[j][i+1], U = i? d_i-1,j : 0,
 *      Dn = i<n-1 ? d_i+1,j : vout_j, wD = w/((bU+bD)+g)
 * Stop after the first sweep whose max |x' - x| < tol (reference rule,
 * concrete_command.h:1256-1272).  iout_j = sum_{i ascending} (u-d)*g.
 * Returns sweeps performed; *record = 0-based break sweep or -1.
 */
/* The FP64 red-black replay with general edges: e = top, left, bottom, right
 * driven where on[e][k] != 0 at v[e][k] (NULL on: the nodeBaseGS network).
 * general = 1 takes nodeBasedGeneralGS's initial state (u = v_left on driven
 * rows, else 0; d = 0) and readout (sink wire currents of driven bottoms);
 * general = 0 the nodeBaseGS ones (u = vin, d = vout; sum (u - d) g). */
static int rb_f64_impl(int n, int m, const double *RU, int ru_ld, const double *RD, int rd_ld, const double *g,
                       const int *const *on, const double *vt, const double *vin, const double *vout,
                       const double *vr, int general, int iter_number, double w, int enable_break, double tol,
                       double *u, double *d, double *iout, int *record_out)
{
    size_t nm = (size_t)n * m;
    double *wDu = (double *)malloc(nm * sizeof(double));
    double *wDd = (double *)malloc(nm * sizeof(double));
    double *aL = (double *)malloc(nm * sizeof(double));
    double *aR = (double *)malloc(nm * sizeof(double));
    double *bU = (double *)malloc(nm * sizeof(double));
    double *bD = (double *)malloc(nm * sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            size_t k = IDX(i, j, m);
            aL[k] = (j > 0 || edge_on(on, 1, i)) ? 1.0 / RU[(size_t)i * ru_ld + j] : 0.0;
            aR[k] = (j < m - 1 || edge_on(on, 3, i)) ? 1.0 / RU[(size_t)i * ru_ld + j + 1] : 0.0;
            bU[k] = (i > 0 || edge_on(on, 0, j)) ? 1.0 / RD[(size_t)j * rd_ld + i] : 0.0;
            bD[k] = (i < n - 1 || edge_on(on, 2, j)) ? 1.0 / RD[(size_t)j * rd_ld + i + 1] : 0.0;
            wDu[k] = w / ((aL[k] + aR[k]) + g[k]);
            wDd[k] = w / ((bU[k] + bD[k]) + g[k]);
            u[k] = general ? (edge_on(on, 1, i) ? vin[i] : 0.0) : vin[i];
            d[k] = general ? 0.0 : vout[j];
        }
    double omw = 1.0 - w;
    int record = -1, it;
    for (it = 0; it < iter_number; ++it) {
        double mx = 0;
        for (int c = 0; c < 2; ++c)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < m; ++j) {
                    size_t k = IDX(i, j, m);
                    if (((i + j) & 1) == c) {
                        double L = j ? u[k - 1] : vin[i];
                        double R = (j < m - 1) ? u[k + 1] : (vr ? vr[i] : 0.0);
                        double num = (aL[k] * L + aR[k] * R) + g[k] * d[k];
                        double nv = wDu[k] * num + omw * u[k];
                        mx = fmax(mx, fabs(nv - u[k]));
                        u[k] = nv;
                    } else {
                        double U = i ? d[k - m] : (vt ? vt[j] : 0.0);
                        double Dn = (i < n - 1) ? d[k + m] : vout[j];
                        double num = (bU[k] * U + bD[k] * Dn) + g[k] * u[k];
                        double nv = wDd[k] * num + omw * d[k];
                        mx = fmax(mx, fabs(nv - d[k]));
                        d[k] = nv;
                    }
                }
        if (enable_break && mx < tol) {
            record = it;
            ++it;
            break;
        }
    }
    for (int j = 0; j < m; ++j) {
        if (general) {
            iout[j] = edge_on(on, 2, j) ? (d[IDX(n - 1, j, m)] - vout[j]) / RD[(size_t)j * rd_ld + n] : 0.0;
        } else {
            double s = 0;
            for (int i = 0; i < n; ++i)
                s += (u[IDX(i, j, m)] - d[IDX(i, j, m)]) * g[IDX(i, j, m)];
            iout[j] = s;
        }
    }
    *record_out = record;
    free(wDu);
    free(wDd);
    free(aL);
    free(aR);
    free(bU);
    free(bD);
    return it;
}

int oracle_rb_f64(int n, int m, const double *RU, int ru_ld, const double *RD,
                  int rd_ld, const double *g, const double *vin,
                  const double *vout, int iter_number, double w,
                  int enable_break, double tol, double *u, double *d,
                  double *iout, int *record_out)
{
    return rb_f64_impl(n, m, RU, ru_ld, RD, rd_ld, g, NULL, NULL, vin, vout, NULL, 0, iter_number, w, enable_break,
                       tol, u, d, iout, record_out);
}

int oracle_rb_f64_general(int n, int m, const double *RU, int ru_ld, const double *RD, int rd_ld, const double *g,
                          const int *on_top, const int *on_left, const int *on_bottom, const int *on_right,
                          const double *v_top, const double *v_left, const double *v_bottom, const double *v_right,
                          int iter_number, double w, int enable_break, double tol, double *u, double *d,
                          double *iout, int *record_out)
{
    const int *on[4] = {on_top, on_left, on_bottom, on_right};
    return rb_f64_impl(n, m, RU, ru_ld, RD, rd_ld, g, on, v_top, v_left, v_bottom, v_right, 1, iter_number, w,
                       enable_break, tol, u, d, iout, record_out);
}

/* ------------------------------------------------------------------------ */
/* FP32 red-black SOR on drop variables: replay of the GPU batched kernel    */
/* ------------------------------------------------------------------------ */
/*
 * Drop variables p = vin_i - u_ij, q = d_ij - vout_j, c = vin_i - vout_j,
 * uniform wire conductance a = 1/r, hcoef = w/(2a) with w the FP32 omega:
 *   hn  = (float)(-(g * hcoef)),  rho = (float)((g * c) * hcoef)
 *   tt  = fmaf(hn, p + q, rho)
 *   x'  = fmaf(cB, N1 + N2, fmaf(cA, x, tt)),  cA = 1-w, cB = w/2
 *   p: N1 = j ? p_i,j-1 : 0 (source),  N2 = j<m-1 ? p_i,j+1 : fmaf(2/w, tt, N1)
 *   q: N2 = i<n-1 ? q_i+1,j : 0 (sink), N1 = i ? q_i-1,j : fmaf(2/w, tt, N2)
 * (the ghost at an open end makes the update the exact SOR step there).
 * Cell (i,j) updates p in half-sweep (i+j)&1 and q in the other one.
 * Runs exactly `sweeps` sweeps from the given p, q (in/out).
 */
void oracle_rb_f32(int n, int m, const float *g, double r, const double *vin,
                   const double *vout, double w, int sweeps, float *p, float *q)
{
    size_t nm = (size_t)n * m;
    float *hn = (float *)malloc(nm * sizeof(float));
    float *rho = (float *)malloc(nm * sizeof(float));
    const float wf = (float)w;
    const double hcoef = 0.5 * (double)wf * r;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const double gd = (double)g[IDX(i, j, m)];
            const double c = vin[i] - (vout ? vout[j] : 0.0);
            hn[IDX(i, j, m)] = (float)(-(gd * hcoef));
            rho[IDX(i, j, m)] = (float)((gd * c) * hcoef);
        }
    const float cA = 1.0f - wf;
    const float cB = 0.5f * wf;
    const float c2 = 2.0f / wf;
    for (int s = 0; s < sweeps; ++s)
        for (int col = 0; col < 2; ++col)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < m; ++j) {
                    size_t k = IDX(i, j, m);
                    const float sm = p[k] + q[k];
                    const float tt = fmaf(hn[k], sm, rho[k]);
                    if (((i + j) & 1) == col) {
                        const float n1 = j ? p[k - 1] : 0.0f;
                        const float n2 = (j < m - 1) ? p[k + 1] : fmaf(c2, tt, n1);
                        p[k] = fmaf(cB, n1 + n2, fmaf(cA, p[k], tt));
                    } else {
                        const float n2 = (i < n - 1) ? q[k + m] : 0.0f;
                        const float n1 = i ? q[k - m] : fmaf(c2, tt, n2);
                        q[k] = fmaf(cB, n1 + n2, fmaf(cA, q[k], tt));
                    }
                }
    free(hn);
    free(rho);
}

/* ------------------------------------------------------------------------ */
/* PimTorch hook line-GS (split_arr_matmul.cu:27-140, 204-333)               */
/* ------------------------------------------------------------------------ */
/*
 * Crossbar (b, x, y) of a [bsize, numX, numY] grid; row_size x col_size
 * cells; input[b][x][row]; linear G[x][row][y*col + c] shared over the batch.
 * Sweep: wordlines left->right (new left, old right, old bitline), then
 * bitlines bottom->top (new lower, old upper, new wordline); source and sink
 * conductances equal g_wire (split_arr_matmul.cu:274-275, 290-291).
 * Nonlinear cells (split_arr_matmul.cu:246-262): before each sweep the cell
 * conductance becomes 1/R(V) of the voltage difference held in the buffer the
 * sweep is about to overwrite (two sweeps old; the initial state for the
 * first two sweeps), with |V| < 1e-5 clamped to +-1e-5 (deal_v :142-166),
 *   R(V) = (1-z)|V|/(b_on sinh(a_on|V|)) + z|V|/(b_off sinh(a_off|V|)),
 *   z = max(0, (1/G - Ron)/(Roff - Ron))      (:176-196).
 * Break: after a sweep, stop when no node moved by more than thr (:295-306).
 * Output: ret_v ? (u-d)[b][x][y][row][col] : I[b][x][y*col + c].
 * Returns the number of sweeps run.
 */
/* Unit xb of the grid: tile (b, x, y), or in OU mode block (oux, ouy) of it,
 * rows [r0, r0 + R) and columns [c0, c0 + C) of the tile. */
#define HOOK_UNIT                                                                      \
    const int nOuX = ou_rows > 0 ? row / ou_rows : 1, nOuY = ou_rows > 0 ? col / ou_cols : 1; \
    const int R = ou_rows > 0 ? ou_rows : row, C = ou_rows > 0 ? ou_cols : col;        \
    const size_t cells = (size_t)R * C;                                                \
    const size_t nxb = (size_t)bsize * numX * numY * nOuX * nOuY;
#define UNIT_DECODE(xb)                                                                \
    const int ouy = (int)((xb) % nOuY), oux = (int)(((xb) / nOuY) % nOuX);             \
    const size_t t2_ = (xb) / ((size_t)nOuY * nOuX);                                   \
    const int y = (int)(t2_ % numY), x = (int)((t2_ / numY) % numX), b = (int)(t2_ / ((size_t)numY * numX)); \
    const int r0 = oux * R, c0 = ouy * C;                                              \
    (void)b;
#define GLIN(i, c) G[((size_t)x * row + r0 + (i)) * (size_t)(numY * col) + (size_t)y * col + c0 + (c)]
#define INP(i) input[((size_t)b * numX + x) * row + r0 + (i)]

#define HOOK_GS_BODY(T, SINH, FABS)                                                    \
    HOOK_UNIT                                                                          \
    T *u0 = (T *)malloc(nxb * cells * sizeof(T)), *u1 = (T *)malloc(nxb * cells * sizeof(T)); \
    T *d0 = (T *)calloc(nxb * cells, sizeof(T)), *d1 = (T *)calloc(nxb * cells, sizeof(T));  \
    T *Gc = (T *)malloc(nxb * cells * sizeof(T));                                      \
    T *z = nonlinear ? (T *)malloc(nxb * cells * sizeof(T)) : NULL;                    \
    const T gw = (T)g_wire, be = (T)beta, omb = (T)1 - be;                             \
    for (size_t xb = 0; xb < nxb; ++xb) {                                              \
        UNIT_DECODE(xb)                                                                \
        for (int i = 0; i < R; ++i)                                                    \
            for (int c = 0; c < C; ++c) {                                              \
                const size_t k = xb * cells + (size_t)i * C + c;                       \
                u0[k] = u1[k] = INP(i);                                                \
                Gc[k] = GLIN(i, c);                                                    \
                if (nonlinear) {                                                       \
                    T zz = ((T)1 / GLIN(i, c) - (T)nl[0]) / (T)(nl[1] - nl[0]);        \
                    z[k] = zz < 0 ? 0 : zz;                                            \
                }                                                                      \
            }                                                                          \
    }                                                                                  \
    /* cur/prev buffers: the sweep writes cur, reads prev; then swap */                \
    T *ucur = u0, *uprev = u1, *dcur = d0, *dprev = d1;                                \
    int it;                                                                            \
    for (it = 0; it < iter_times; ++it) {                                              \
        if (nonlinear)                                                                 \
            for (size_t k = 0; k < nxb * cells; ++k) {                                 \
                T V = ucur[k] - dcur[k];                                               \
                if (V < (T)1e-5 && V > -(T)1e-5)                                       \
                    V = V > 0 ? (T)1e-5 : -(T)1e-5;                                    \
                V = FABS(V);                                                           \
                T Ion = (T)nl[3] * SINH(V * (T)nl[2]);                                 \
                T Ioff = (T)nl[5] * SINH(V * (T)nl[4]);                                \
                T R_ = ((T)1 - z[k]) * V / Ion + z[k] * V / Ioff;                      \
                Gc[k] = (T)1 / R_;                                                     \
            }                                                                          \
        for (size_t xb = 0; xb < nxb; ++xb) {                                          \
            UNIT_DECODE(xb)                                                            \
            const T gl = ou_rows > 0 ? gw / (T)(c0 + 1) : (T)g_load_left;              \
            const T gd = ou_rows > 0 ? gw / (T)(row - (r0 + R - 1)) : (T)g_load_down;  \
            T *uc = ucur + xb * cells, *up = uprev + xb * cells;                       \
            T *dc = dcur + xb * cells, *dp = dprev + xb * cells;                       \
            const T *Gk = Gc + xb * cells;                                             \
            for (int i = 0; i < R; ++i) {                                              \
                T *ucr = uc + (size_t)i * C, *upr = up + (size_t)i * C;                \
                const T *dpr = dp + (size_t)i * C, *Gr = Gk + (size_t)i * C;           \
                T left = INP(i);                                                       \
                T sumg = gw + gl + Gr[0];                                              \
                ucr[0] = be * (gl * left + gw * upr[1] + Gr[0] * dpr[0]) / sumg + omb * upr[0]; \
                int j;                                                                 \
                for (j = 1; j < C - 1; ++j) {                                          \
                    sumg = gw * 2 + Gr[j];                                             \
                    ucr[j] = be * (gw * (ucr[j - 1] + upr[j + 1]) + Gr[j] * dpr[j]) / sumg + omb * upr[j]; \
                }                                                                      \
                sumg = gw + Gr[j];                                                     \
                ucr[j] = be * (gw * ucr[j - 1] + Gr[j] * dpr[j]) / sumg + omb * upr[j];\
            }                                                                          \
            for (int c = 0; c < C; ++c) {                                              \
                int i = R - 1;                                                         \
                T sumg = gw + gd + Gk[(size_t)i * C + c];                              \
                dc[(size_t)i * C + c] = be * (gw * dp[(size_t)(i - 1) * C + c] + gd * (T)0 + \
                    Gk[(size_t)i * C + c] * uc[(size_t)i * C + c]) / sumg + omb * dp[(size_t)i * C + c]; \
                for (i = R - 2; i > 0; --i) {                                          \
                    sumg = gw * 2 + Gk[(size_t)i * C + c];                             \
                    dc[(size_t)i * C + c] = be * (gw * (dp[(size_t)(i - 1) * C + c] + dc[(size_t)(i + 1) * C + c]) + \
                        Gk[(size_t)i * C + c] * uc[(size_t)i * C + c]) / sumg + omb * dp[(size_t)i * C + c]; \
                }                                                                      \
                sumg = gw + Gk[c];                                                     \
                dc[c] = be * (gw * dc[(size_t)C + c] + Gk[c] * uc[c]) / sumg + omb * dp[c]; \
            }                                                                          \
        }                                                                              \
        T *t = ucur; ucur = uprev; uprev = t;                                          \
        t = dcur; dcur = dprev; dprev = t;                                             \
        if (enable_break) {                                                            \
            long cnt = 0;                                                              \
            for (size_t k = 0; k < nxb * cells; ++k) {                                 \
                cnt += FABS(ucur[k] - uprev[k]) > (T)thr;                              \
                cnt += FABS(dcur[k] - dprev[k]) > (T)thr;                              \
            }                                                                          \
            if (cnt == 0) { ++it; break; }                                             \
        }                                                                              \
    }                                                                                  \
    if (ret_v) {                                                                       \
        for (size_t k = 0; k < nxb * cells; ++k)                                       \
            out[k] = uprev[k] - dprev[k];                                              \
    } else {                                                                           \
        for (size_t xb = 0; xb < nxb; ++xb) {                                          \
            UNIT_DECODE(xb)                                                            \
            for (int c = 0; c < C; ++c) {                                              \
                T s = 0;                                                               \
                for (int i = 0; i < R; ++i) {                                          \
                    size_t k = xb * cells + (size_t)i * C + c;                         \
                    s += (uprev[k] - dprev[k]) * Gc[k];                                \
                }                                                                      \
                out[((((size_t)b * numX + x) * nOuX + oux) * numY + y) * col + c0 + c] = s; \
            }                                                                          \
        }                                                                              \
    }                                                                                  \
    free(u0); free(u1); free(d0); free(d1); free(Gc); free(z);                         \
    return it;

/* With separate source / sink loads (ir_gs_ou_local, split_arr_matmul.cu:690-815)
 * and OU mode (ir_gs_ou, :1086-1327: every ou_rows x ou_cols block its own
 * crossbar with loads g_wire/(c0+1), g_wire/(row - last row); output
 * I[b][x][oux][y*col + c0 + c]). */
int oracle_hook_gs_ou_f64(const double *input, const double *G, int bsize, int row, int col, int numX, int numY,
                          int iter_times, double g_wire, double g_load_left, double g_load_down, double beta,
                          int enable_break, double thr, int nonlinear, const double *nl, int ret_v, int ou_rows,
                          int ou_cols, double *out)
{
    HOOK_GS_BODY(double, sinh, fabs)
}

int oracle_hook_gs_ou_f32(const float *input, const float *G, int bsize, int row, int col, int numX, int numY,
                          int iter_times, double g_wire, double g_load_left, double g_load_down, double beta,
                          int enable_break, double thr, int nonlinear, const double *nl, int ret_v, int ou_rows,
                          int ou_cols, float *out)
{
    HOOK_GS_BODY(float, sinhf, fabsf)
}

int oracle_hook_gs_f64(const double *input, const double *G, int bsize, int row,
                       int col, int numX, int numY, int iter_times,
                       double g_wire, double beta, int enable_break, double thr,
                       int nonlinear, const double *nl, int ret_v, double *out)
{
    return oracle_hook_gs_ou_f64(input, G, bsize, row, col, numX, numY, iter_times, g_wire, g_wire, g_wire, beta,
                                 enable_break, thr, nonlinear, nl, ret_v, 0, 0, out);
}

int oracle_hook_gs_f32(const float *input, const float *G, int bsize, int row,
                       int col, int numX, int numY, int iter_times,
                       double g_wire, double beta, int enable_break, double thr,
                       int nonlinear, const double *nl, int ret_v, float *out)
{
    return oracle_hook_gs_ou_f32(input, G, bsize, row, col, numX, numY, iter_times, g_wire, g_wire, g_wire, beta,
                                 enable_break, thr, nonlinear, nl, ret_v, 0, 0, out);
}

/* CPU baseline: the lexicographic restatement (oracle_nodegs) over a batch of
 * instances with FP32 conductances, one wire resistance each and grounded
 * bitlines, OpenMP over instances.  Returns the number of threads used. */
int oracle_nodegs_batch_f32g(int B, int n, int m, const float *g32, const double *r,
                             const double *vin, int iters, double w, double tol,
                             double *iout, int *sweeps)
{
    int nthreads = 1;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            nthreads = omp_get_num_threads();
#endif
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < B; ++k) {
        size_t nm = (size_t)n * m;
        double *RU = (double *)malloc((size_t)n * (m + 1) * sizeof(double));
        double *RD = (double *)malloc((size_t)m * (n + 1) * sizeof(double));
        double *g = (double *)malloc(nm * sizeof(double));
        double *vo = (double *)calloc((size_t)m, sizeof(double));
        double *u = (double *)malloc(nm * sizeof(double));
        double *d = (double *)malloc(nm * sizeof(double));
        for (size_t t = 0; t < (size_t)n * (m + 1); ++t)
            RU[t] = r[k];
        for (size_t t = 0; t < (size_t)m * (n + 1); ++t)
            RD[t] = r[k];
        for (size_t t = 0; t < nm; ++t)
            g[t] = (double)g32[(size_t)k * nm + t];
        int rec = -1;
        oracle_nodegs(n, m, RU, m + 1, RD, n + 1, g, vin + (size_t)k * n, vo, iters, w, 1, tol, u, d,
                      iout + (size_t)k * m, &rec);
        sweeps[k] = rec >= 0 ? rec + 1 : iters;
        free(RU);
        free(RD);
        free(g);
        free(vo);
        free(u);
        free(d);
    }
    return nthreads;
}


/* ---------------------------------------------------------------- IR2S
 * One sweep: cell currents of the present voltages, then every wordline and
 * every bitline solved exactly for those currents (the source / sink segment
 * r_load, inner segments 1/g_wire): wordline i carries, in segment j, the sum
 * of the cell currents at columns >= j; bitline j carries, in segment i (below
 * row i), the sum of the currents at rows <= i.  Accumulation orders and the
 * division by g_wire follow ir_drop_fast_solver.h:24-80 (no break) and
 * :81-163 (break when the largest voltage change of a sweep is < break_th).
 * vup / vdown hold the starting voltages on entry.  Returns the reference's
 * count: `times` without the break, else (0-based break sweep or times) + 1.
 */
int oracle_ir2s(int n, int m, const double *g, const double *vin, const double *vout, double g_wire,
                double r_load, int times, int enable_break, double break_th, double *vup, double *vdown,
                double *out)
{
    double *cur = (double *)malloc(sizeof(double) * (size_t)n * m);
    double *wl = (double *)malloc(sizeof(double) * (size_t)n);
    double *bl = (double *)malloc(sizeof(double) * (size_t)m);
    if (!cur || !wl || !bl) {
        free(cur); free(wl); free(bl);
        return -1;
    }
    int record = times;
    for (int k = 0; k < times; ++k) {
        double dmax = 0.0;
        for (int i = 0; i < n; ++i) wl[i] = 0.0;
        for (int j = 0; j < m; ++j) bl[j] = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                const double c = (vup[i * m + j] - vdown[i * m + j]) * g[i * m + j];
                cur[i * m + j] = c;
                wl[i] += c;
                bl[j] += c;
            }
        for (int i = 0; i < n; ++i) {
            double carried = wl[i], v = vin[i] - carried * r_load;
            dmax = fmax(dmax, fabs(v - vup[i * m]));
            vup[i * m] = v;
            carried -= cur[i * m];
            for (int j = 1; j < m; ++j) {
                v = vup[i * m + j - 1] - carried / g_wire;
                dmax = fmax(dmax, fabs(v - vup[i * m + j]));
                vup[i * m + j] = v;
                carried -= cur[i * m + j];
            }
        }
        for (int j = 0; j < m; ++j) {
            const double v = bl[j] * r_load + vout[j];
            dmax = fmax(dmax, fabs(v - vdown[(n - 1) * m + j]));
            vdown[(n - 1) * m + j] = v;
            bl[j] -= cur[(n - 1) * m + j];
        }
        for (int i = n - 2; i >= 0; --i)
            for (int j = 0; j < m; ++j) {
                const double v = vdown[(i + 1) * m + j] + bl[j] / g_wire;
                dmax = fmax(dmax, fabs(v - vdown[i * m + j]));
                vdown[i * m + j] = v;
                bl[j] -= cur[i * m + j];
            }
        if (enable_break && dmax < break_th) {
            record = k;
            break;
        }
    }
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i)
            s += (vup[i * m + j] - vdown[i * m + j]) * g[i * m + j];
        out[j] = s;
    }
    free(cur); free(wl); free(bl);
    return enable_break ? record + 1 : times;
}

/*
 * The PimTorch hook's ir2s (split_arr_matmul.cu:336-545): the same grid,
 * layout and nonlinear lag as oracle_hook_gs_*; per iteration the wordline
 * kernel (:336-375) and the bitline kernel (:377-414) both read the newest
 * state ("prev") and write the other buffer, each line's running current
 * sum (I_row_sum / I_col_sum) persisting across iterations with its rounding
 * residue; inner segments multiply by r_wire, source and sink segments too.
 * Break: after an iteration, stop when no node of the whole grid moved by
 * more than thr (:517-527).  Returns the number of iterations run.
 */
#define HOOK_IR2S_BODY(T, SINH, FABS)                                                  \
    HOOK_UNIT                                                                          \
    T *ua = (T *)malloc(nxb * cells * sizeof(T)), *ub = (T *)malloc(nxb * cells * sizeof(T)); \
    T *da = (T *)calloc(nxb * cells, sizeof(T)), *db = (T *)calloc(nxb * cells, sizeof(T));  \
    T *Gk = (T *)malloc(nxb * cells * sizeof(T)), *Glin = (T *)malloc(nxb * cells * sizeof(T)); \
    T *wsum = (T *)calloc(nxb * R, sizeof(T)), *bsum = (T *)calloc(nxb * C, sizeof(T)); \
    const T rw = (T)r_wire;                                                            \
    for (size_t xb = 0; xb < nxb; ++xb) {                                              \
        UNIT_DECODE(xb)                                                                \
        for (int i = 0; i < R; ++i)                                                    \
            for (int c = 0; c < C; ++c) {                                              \
                const size_t k = xb * cells + (size_t)i * C + c;                       \
                ua[k] = ub[k] = INP(i);                                                \
                Gk[k] = Glin[k] = GLIN(i, c);                                          \
            }                                                                          \
    }                                                                                  \
    T *ucur = ua, *uprev = ub, *dcur = da, *dprev = db;                                \
    int it;                                                                            \
    for (it = 0; it < iter_times; ++it) {                                              \
        if (nonlinear)                                                                 \
            for (size_t k = 0; k < nxb * cells; ++k) {                                 \
                T V = ucur[k] - dcur[k];                                               \
                if (V < (T)1e-5 && V > -(T)1e-5)                                       \
                    V = V > 0 ? (T)1e-5 : -(T)1e-5;                                    \
                V = FABS(V);                                                           \
                T zz = ((T)1 / Glin[k] - (T)nl[0]) / (T)(nl[1] - nl[0]);               \
                if (zz < 0) zz = 0;                                                    \
                T R_ = ((T)1 - zz) * V / ((T)nl[3] * SINH(V * (T)nl[2])) +             \
                       zz * V / ((T)nl[5] * SINH(V * (T)nl[4]));                       \
                Gk[k] = (T)1 / R_;                                                     \
            }                                                                          \
        for (size_t xb = 0; xb < nxb; ++xb) {                                          \
            UNIT_DECODE(xb)                                                            \
            const T rsrc = ou_rows > 0 ? rw * (T)(c0 + 1) : (T)r_source;               \
            const T rsnk = ou_rows > 0 ? rw * (T)(row - (r0 + R - 1)) : (T)r_sink;     \
            const size_t o = xb * cells;                                               \
            for (int i = 0; i < R; ++i) {                                              \
                T *s = wsum + xb * R + i;                                              \
                const size_t ri = o + (size_t)i * C;                                   \
                for (int j = 0; j < C; ++j)                                            \
                    *s += (uprev[ri + j] - dprev[ri + j]) * Gk[ri + j];                \
                ucur[ri] = INP(i) - *s * rsrc;                                         \
                *s -= (uprev[ri] - dprev[ri]) * Gk[ri];                                \
                for (int j = 1; j < C; ++j) {                                          \
                    ucur[ri + j] = ucur[ri + j - 1] - *s * rw;                         \
                    *s -= (uprev[ri + j] - dprev[ri + j]) * Gk[ri + j];                \
                }                                                                      \
            }                                                                          \
            for (int c = 0; c < C; ++c) {                                              \
                T *s = bsum + xb * C + c;                                              \
                for (int i = 0; i < R; ++i) {                                          \
                    const size_t k = o + (size_t)i * C + c;                            \
                    *s += (uprev[k] - dprev[k]) * Gk[k];                               \
                }                                                                      \
                size_t k = o + (size_t)(R - 1) * C + c;                                \
                dcur[k] = *s * rsnk;                                                   \
                *s -= (uprev[k] - dprev[k]) * Gk[k];                                   \
                for (int i = R - 2; i >= 0; --i) {                                     \
                    k = o + (size_t)i * C + c;                                         \
                    dcur[k] = dcur[k + C] + *s * rw;                                   \
                    *s -= (uprev[k] - dprev[k]) * Gk[k];                               \
                }                                                                      \
            }                                                                          \
        }                                                                              \
        T *t = ucur; ucur = uprev; uprev = t;                                          \
        t = dcur; dcur = dprev; dprev = t;                                             \
        if (enable_break) {                                                            \
            long cnt = 0;                                                              \
            for (size_t k = 0; k < nxb * cells; ++k) {                                 \
                cnt += FABS(ucur[k] - uprev[k]) > (T)thr;                              \
                cnt += FABS(dcur[k] - dprev[k]) > (T)thr;                              \
            }                                                                          \
            if (cnt == 0) { ++it; break; }                                             \
        }                                                                              \
    }                                                                                  \
    if (ret_v) {                                                                       \
        for (size_t k = 0; k < nxb * cells; ++k)                                       \
            out[k] = uprev[k] - dprev[k];                                              \
    } else {                                                                           \
        for (size_t xb = 0; xb < nxb; ++xb) {                                          \
            UNIT_DECODE(xb)                                                            \
            for (int c = 0; c < C; ++c) {                                              \
                T s = 0;                                                               \
                for (int i = 0; i < R; ++i) {                                          \
                    const size_t k = xb * cells + (size_t)i * C + c;                   \
                    s += (uprev[k] - dprev[k]) * Gk[k];                                \
                }                                                                      \
                out[((((size_t)b * numX + x) * nOuX + oux) * numY + y) * col + c0 + c] = s; \
            }                                                                          \
        }                                                                              \
    }                                                                                  \
    free(ua); free(ub); free(da); free(db); free(Gk); free(Glin); free(wsum); free(bsum); \
    return it;

/* With separate source / sink resistances (ir2s_ou_local, split_arr_matmul.cu:546-688)
 * and OU mode (ir2s_ou, :816-1085: loads r_wire (c0+1), r_wire (row - last row)). */
int oracle_hook_ir2s_ou_f64(const double *input, const double *G, int bsize, int row, int col, int numX, int numY,
                            int iter_times, double r_wire, double r_source, double r_sink, int enable_break,
                            double thr, int nonlinear, const double *nl, int ret_v, int ou_rows, int ou_cols,
                            double *out)
{
    HOOK_IR2S_BODY(double, sinh, fabs)
}

int oracle_hook_ir2s_ou_f32(const float *input, const float *G, int bsize, int row, int col, int numX, int numY,
                            int iter_times, double r_wire, double r_source, double r_sink, int enable_break,
                            double thr, int nonlinear, const double *nl, int ret_v, int ou_rows, int ou_cols,
                            float *out)
{
    HOOK_IR2S_BODY(float, sinhf, fabsf)
}

int oracle_hook_ir2s_f64(const double *input, const double *G, int bsize, int row, int col, int numX, int numY,
                         int iter_times, double r_wire, int enable_break, double thr, int nonlinear,
                         const double *nl, int ret_v, double *out)
{
    return oracle_hook_ir2s_ou_f64(input, G, bsize, row, col, numX, numY, iter_times, r_wire, r_wire, r_wire,
                                   enable_break, thr, nonlinear, nl, ret_v, 0, 0, out);
}

int oracle_hook_ir2s_f32(const float *input, const float *G, int bsize, int row, int col, int numX, int numY,
                         int iter_times, double r_wire, int enable_break, double thr, int nonlinear,
                         const double *nl, int ret_v, float *out)
{
    return oracle_hook_ir2s_ou_f32(input, G, bsize, row, col, numX, numY, iter_times, r_wire, r_wire, r_wire,
                                   enable_break, thr, nonlinear, nl, ret_v, 0, 0, out);
}


/* nodeBasedGeneralGS (crossbarHspice/2dArray.cc:1209-1374), statement by
 * statement: lexicographic SOR over (i, j) updating u_ij then d_ij, edges
 * driven or open per on[e] with the voltages v[e] (e = top, left, bottom,
 * right), double-buffered iterates (the new left / upper neighbour, the old
 * right / lower one), initial u = v_left on driven rows (0 elsewhere), d = 0;
 * break after the first sweep whose max |change| < tol.  Outputs the final
 * buffers, the reference's readout (sink wire currents of driven bottoms)
 * and the 0-based break sweep (-1 without a break).  Returns 0. */
int oracle_nodegs_general(int n, int m, const double *RU, int ru_ld, const double *RD, int rd_ld, const double *g,
                          const int *on_top, const int *on_left, const int *on_bottom, const int *on_right,
                          const double *v_top, const double *v_left, const double *v_bottom, const double *v_right,
                          int iters, double w, int enable_break, double tol, double *vup, double *vdown,
                          double *iout, int *record)
{
    const size_t nm = (size_t)n * m;
    double *U[2], *D[2];
    U[0] = (double *)malloc(nm * sizeof(double));
    U[1] = (double *)malloc(nm * sizeof(double));
    D[0] = (double *)calloc(nm, sizeof(double));
    D[1] = (double *)calloc(nm, sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j)
            U[0][IDX(i, j, m)] = U[1][IDX(i, j, m)] = on_left[i] ? v_left[i] : 0.0;
#define RUL(i, j) RU[(size_t)(i) * ru_ld + (j)]
#define RDL(j, i) RD[(size_t)(j) * rd_ld + (i)]
    int rec = -1;
    for (int it = 0; it < iters; ++it) {
        const int x = it & 1, y = 1 - x;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                const size_t k = IDX(i, j, m);
                double sumg = g[k] + ((j != 0 || on_left[i] == 1) ? 1.0 / RUL(i, j) : 0) +
                              ((j != m - 1 || on_right[i] == 1) ? 1.0 / RUL(i, j + 1) : 0);
                double l_side = (j == 0 ? (on_left[i] == 0 ? 0 : v_left[i]) : U[x][k - 1]) / RUL(i, j);
                double r_side = (j == m - 1 ? (on_right[i] == 0 ? 0 : v_right[i]) : U[y][k + 1]) / RUL(i, j + 1);
                double other = D[y][k] * g[k];
                U[x][k] = w * (l_side + r_side + other) / sumg + (1 - w) * U[y][k];
                sumg = g[k] + ((i != 0 || on_top[j] == 1) ? 1.0 / RDL(j, i) : 0) +
                       ((i != n - 1 || on_bottom[j] == 1) ? 1.0 / RDL(j, i + 1) : 0);
                l_side = (i == 0 ? (on_top[j] == 0 ? 0 : v_top[j]) : D[x][k - m]) / RDL(j, i);
                r_side = (i == n - 1 ? (on_bottom[j] == 0 ? 0 : v_bottom[j]) : D[y][k + m]) / RDL(j, i + 1);
                other = U[x][k] * g[k];
                D[x][k] = w * (l_side + r_side + other) / sumg + (1 - w) * D[y][k];
            }
        if (enable_break) {
            double mx = 0;
            for (size_t k = 0; k < nm; ++k) {
                mx = fmax(mx, fabs(D[x][k] - D[y][k]));
                mx = fmax(mx, fabs(U[x][k] - U[y][k]));
            }
            if (mx < tol) {
                rec = it;
                break;
            }
        }
    }
    const int fin = rec != -1 ? (rec & 1) : 1 - (iters & 1);
    memcpy(vup, U[fin], nm * sizeof(double));
    memcpy(vdown, D[fin], nm * sizeof(double));
    for (int j = 0; j < m; ++j)
        iout[j] = on_bottom[j] == 1 ? (vdown[IDX(n - 1, j, m)] - v_bottom[j]) / RDL(j, n) : 0.0;
#undef RUL
#undef RDL
    *record = rec;
    free(U[0]); free(U[1]); free(D[0]); free(D[1]);
    return 0;
}

/* The general solve's "error I sum" (2dArray.cc:1316-1345): interior wordline
 * nodes j >= 1 and bitline rows 0 .. n-2, wire conductance 1/RD[0][0]. */
double oracle_errisum_general(int n, int m, double g_wire, const double *g, const double *vup, const double *vdown)
{
    double e = 0;
#define U(i, j) vup[IDX(i, j, m)]
#define D(i, j) vdown[IDX(i, j, m)]
#define G(i, j) g[IDX(i, j, m)]
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < m - 1; ++j)
            e += fabs((U(i, j - 1) - U(i, j) + U(i, j + 1) - U(i, j)) * g_wire + (D(i, j) - U(i, j)) * G(i, j));
        if (m > 1)
            e += fabs((U(i, m - 2) - U(i, m - 1)) * g_wire + (D(i, m - 1) - U(i, m - 1)) * G(i, m - 1));
    }
    for (int j = 0; j < m; ++j)
        e += fabs((U(0, j) - D(0, j)) * G(0, j) + (n > 1 ? D(1, j) - D(0, j) : 0) * g_wire);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 0; j < m; ++j)
            e += fabs((U(i, j) - D(i, j)) * G(i, j) + (D(i - 1, j) - D(i, j) + D(i + 1, j) - D(i, j)) * g_wire);
#undef U
#undef D
#undef G
    return e;
}
