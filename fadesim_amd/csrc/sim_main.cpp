// `sim <config>` -- the crossbar command language of the FADESim CLI, with
// nodebasedGSMethod solved on the GPU through fadesim_gs64_solve and the V1
// IR2S command fastsolve through fadesim_ir2s.
//
// Command semantics follow the reference CLI (V2 registry
// crossbarHspice/V2/main.cpp:68-117, command bodies
// crossbarHspice/V2/concrete_command.h:51-1127; V1 spellings of
// crossbarHspice/2dArray.cc:24-873 are accepted too): one command per line,
// whitespace-separated arguments, `//`, `%` and `#` comment lines.  Output
// files keep the reference formats: fastmode_nodebasedGS.out
// (concrete_command.h:1321-1353) and the operation log default.log
// (crossbarHspice/V2/main.cpp:61-66, abstract_array.h:17-29).  Like the
// reference, an error prints "Error(s) occured in line N: ..." and the log,
// and the process exits 0 (FADESIM_STRICT_EXIT=1 makes it exit 1).
//
// Extension: `seed <n>` makes randSetR / randSetInputV reproducible (the
// reference seeds them from the clock or std::random_device).
// Commands that only drive SPICE netlist generation or other solvers are
// accepted and logged as not handled by this build.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fadesim.h"

namespace {

struct CmdError : std::runtime_error {
    explicit CmdError(const std::string &s) : std::runtime_error(s) {}
};

struct Array2D {
    int n = 0, m = 0;
    std::vector<std::vector<int>> arr;           // cell state index
    std::vector<std::vector<double>> RU, RD;     // RU[n][m+1], RD[m][n+1]
    std::vector<double> v[4];                    // up, left, down, right
    std::vector<int> bv[4];
    double r_load[4] = {-1, -1, -1, -1};
    std::vector<double> cellR_states, inputV_states;
    std::vector<std::string> history;
    bool seeded = false;
    std::mt19937_64 rng;

    void setSize(int nn, int mm)
    {
        n = nn;
        m = mm;
        arr.assign(n, std::vector<int>(m, 0));
        RU.assign(n, std::vector<double>(m + 1, 0.0));
        RD.assign(m, std::vector<double>(n + 1, 0.0));
        v[0] = v[2] = std::vector<double>(m, 0.0);
        v[1] = v[3] = std::vector<double>(n, 0.0);
        bv[0] = bv[2] = std::vector<int>(m, 0);
        bv[1] = bv[3] = std::vector<int>(n, 0);
    }
    void log(const std::string &s) { history.push_back(s); }
    void printLog(std::ostream &os) const
    {
        os << std::endl << "--------Operation Log---------" << std::endl;
        for (auto &h : history)
            os << h << std::endl;
        os << "--------Log end--------" << std::endl;
    }
};

template <typename T>
T parse(const std::string &s)
{
    std::istringstream iss(s);
    T v;
    iss >> v;
    if (iss.fail() || !iss.eof())
        throw CmdError("Failed to parse argument: " + s);
    return v;
}

int getDir(const std::string &d)
{
    if (d == "up") return 0;
    if (d == "left") return 1;
    if (d == "down") return 2;
    if (d == "right") return 3;
    throw CmdError("Arg str not matched Direction should be up|left|down|right.");
}

bool checkYes(std::string s)
{
    for (auto &c : s)
        c = (char)::tolower(c);
    return s == "yes" || s == "1" || s == "true" || s == "y" || s == "t";
}

void need_size(const Array2D &a)
{
    if (a.n == 0 || a.m == 0)
        throw CmdError("Please set arraySize before this command");
}

std::vector<int> sample(Array2D &a, const std::vector<double> &p, size_t len)
{
    double sum = 0;
    for (double x : p)
        sum += x;
    if (std::abs(sum - 1.0) > 1e-6)
        throw CmdError("Error: p sum is not 1");
    if (!a.seeded) {
        std::random_device rd;
        a.rng.seed(((uint64_t)rd() << 32) ^ rd());
        a.seeded = true;
    }
    std::discrete_distribution<int> dist(p.begin(), p.end());
    std::vector<int> out(len);
    for (auto &x : out)
        x = dist(a.rng);
    return out;
}

#define HIPCHK(e)                                                                  \
    do {                                                                           \
        hipError_t r_ = (e);                                                       \
        if (r_ != hipSuccess)                                                      \
            throw CmdError(std::string("HIP error: ") + hipGetErrorString(r_));    \
    } while (0)

// V1's dispatcher (crossbarHspice/2dArray.cc:84-110): a driven top or right
// edge, or an undriven left or bottom line, needs the general solver.
bool general_boundary(const Array2D &a)
{
    for (int x : a.bv[0])
        if (x == 1)
            return true;
    for (int x : a.bv[1])
        if (x == 0)
            return true;
    for (int x : a.bv[2])
        if (x == 0)
            return true;
    for (int x : a.bv[3])
        if (x == 1)
            return true;
    return false;
}

// Command-line options: -v1 selects the V1 CLI's semantics
// (crossbarHspice/2dArray.cc: omega parsed with stof, wordline segments from
// RULine, mode bit 0); the default is V2's (concrete_command.h: stod, the
// wordline segments of a square array read from RDLine, :1214).  The
// environment variable FADESIM_GS64_REDBLACK=1 solves with red-black order
// instead of the reference's lexicographic one.
bool g_v1 = false;

struct GsRun {
    std::vector<double> vup, vdown, iout;
    int record = -1, sweeps = 0;
    double seconds = 0;
    bool redblack = false;
};

// One fadesim_gs64_solve call: stages the host arrays on the device, solves,
// and returns the fields, the readout and the break record.
GsRun run_gs64(const Array2D &a, const std::vector<double> &g, const std::vector<double> &ru,
               const std::vector<double> &rd, int iter_time, double w, bool enable_break, double tol, bool general,
               const std::vector<double> *init_u, const std::vector<double> *init_d)
{
    const int n = a.n, m = a.m;
    const size_t nm = (size_t)n * m;
    GsRun res;
    res.vup.resize(nm);
    res.vdown.resize(nm);
    res.iout.resize(m);
    const char *rbenv = std::getenv("FADESIM_GS64_REDBLACK");
    res.redblack = (rbenv && std::string(rbenv) == "1") || n > 512;
    std::vector<void *> owned;
    auto up = [&](const void *src, size_t bytes) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        owned.push_back(p);
        if (src)
            HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        return p;
    };
    struct Guard {
        std::vector<void *> &v;
        ~Guard()
        {
            for (void *p : v)
                (void)hipFree(p);
        }
    } guard{owned};
    fadesim_gs64_desc d{};
    d.struct_size = sizeof d;
    d.rows = n;
    d.cols = m;
    d.d_g = (const double *)up(g.data(), nm * 8);
    d.d_ru = (const double *)up(ru.data(), ru.size() * 8);
    d.ru_ld = m + 1;
    d.d_rd = (const double *)up(rd.data(), rd.size() * 8);
    d.rd_ld = n + 1;
    d.d_vin = (const double *)up(a.v[1].data(), n * 8);
    d.d_vout = (const double *)up(a.v[2].data(), m * 8);
    d.iter_number = iter_time;
    d.omega = w;
    d.enable_break = enable_break;
    d.tol = tol;
    d.d_vup = (double *)up(nullptr, nm * 8);
    d.d_vdown = (double *)up(nullptr, nm * 8);
    d.d_iout = (double *)up(nullptr, m * 8);
    int *d_int = (int *)up(nullptr, 2 * sizeof(int));
    d.d_record = d_int;
    d.d_sweeps = d_int + 1;
    if (general) {
        d.general = 1;
        int *on[4];
        for (int e = 0; e < 4; ++e)
            on[e] = (int *)up(a.bv[e].data(), a.bv[e].size() * sizeof(int));
        d.d_on_top = on[0];
        d.d_on_left = on[1];
        d.d_on_bottom = on[2];
        d.d_on_right = on[3];
        d.d_v_top = (const double *)up(a.v[0].data(), m * 8);
        d.d_v_right = (const double *)up(a.v[3].data(), n * 8);
    }
    if (init_u) {
        d.d_init_up = (const double *)up(init_u->data(), nm * 8);
        d.d_init_down = (const double *)up(init_d->data(), nm * 8);
    }
    d.order = res.redblack ? FADESIM_GS64_REDBLACK : FADESIM_GS64_LEXICOGRAPHIC;
    HIPCHK(hipDeviceSynchronize());
    auto t0 = std::chrono::steady_clock::now();
    if (fadesim_gs64_solve(&d, nullptr) != FADESIM_OK)
        throw CmdError(std::string("solver: ") + fadesim_last_error());
    HIPCHK(hipDeviceSynchronize());
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int hi[2];
    HIPCHK(hipMemcpy(res.vup.data(), d.d_vup, nm * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(res.vdown.data(), d.d_vdown, nm * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(res.iout.data(), d.d_iout, m * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hi, d_int, 2 * sizeof(int), hipMemcpyDeviceToHost));
    res.record = hi[0];
    res.sweeps = hi[1];
    return res;
}

// Segment arrays in the solver's layout: ru [n][m+1] and rd [m][n+1].  V2
// reads the wordline segments of its (square) arrays from RDLine.
void segments(const Array2D &a, std::vector<double> &ru, std::vector<double> &rd, bool v2_alias)
{
    const int n = a.n, m = a.m;
    ru.assign((size_t)n * (m + 1), 0.0);
    rd.assign((size_t)m * (n + 1), 0.0);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i <= n; ++i)
            rd[(size_t)j * (n + 1) + i] = a.RD[j][i];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= m; ++j)
            ru[(size_t)i * (m + 1) + j] = v2_alias ? a.RD[i][j] : a.RU[i][j];
}

// The initial state of V1 mode bit 0 (crossbarHspice/2dArray.cc:1425-1535):
// an equivalent-resistance estimate of the wordline voltages and the bitline
// voltages under them, each expression in the reference's evaluation order.
void scichina_guess(const Array2D &a, const std::vector<double> &g, std::vector<double> &vup,
                    std::vector<double> &vdown)
{
    const int M = a.n, N = a.m;
    const double g_wire = 1.0 / a.RD[0][0];
    const double r_wire = 1 / g_wire, Rs = 1 / g_wire;
    std::vector<std::vector<double>> R(M, std::vector<double>(N)), RR = R, RRR = R, Vwl = R, Rup = R, Rdown = R,
                                                                    Req = R;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            R[i][j] = 1.0 / g[(size_t)i * N + j];
    std::vector<double> gsum(N, 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i)
            gsum[j] += g[(size_t)i * N + j];
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < N; ++j)
            RR[k][j] = R[k][j] + (M - k) * r_wire + (gsum[j] * R[k][j]) * Rs;
    for (int k = 0; k < M; ++k)
        for (int j = N - 1; j >= 0; --j)
            RRR[k][j] = (j == N - 1) ? RR[k][N - 1] + r_wire
                                     : (RR[k][j] * RRR[k][j + 1]) / (RR[k][j] + RRR[k][j + 1]) + r_wire;
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < N; ++j)
            Vwl[k][j] = (RRR[k][j] - r_wire) / RRR[k][j] * (j == 0 ? a.v[1][k] : Vwl[k][j - 1]);
    auto parallel = [](double x, double y) { return x * y / (x + y); };
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            Rup[i][j] = (i == 0) ? r_wire + R[i][j] : parallel(Rup[i - 1][j], R[i][j]) + r_wire;
    for (int i = M - 1; i >= 0; --i)
        for (int j = 0; j < N; ++j)
            Rdown[i][j] = (i != M - 1) ? r_wire + parallel(Rdown[i + 1][j], R[i + 1][j]) : Rs;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j)
            Req[i][j] = (i == 0) ? Rdown[i][j] : parallel(Rup[i - 1][j], Rdown[i][j]);
    vup.assign((size_t)M * N, 0.0);
    vdown.assign((size_t)M * N, 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            vup[(size_t)i * N + j] = Vwl[i][j];
            vdown[(size_t)i * N + j] = Vwl[i][j] * (Req[i][j]) / (Req[i][j] + R[i][j]);
        }
}

// The general-boundary solve (array2D::nodeBasedGeneralGS, 2dArray.cc:1209-1374)
// on the GPU through fadesim_gs64_solve(general = 1); output
// fastmode_nodebasedGeneralGS.out in V1's format (precision 6, width 14).
void generalGS(Array2D &a, const std::vector<double> &g, int iter_time, double w, bool enable_break, double tol,
               std::ostringstream &os)
{
    const int n = a.n, m = a.m;
    std::vector<double> ru, rd;
    segments(a, ru, rd, false);
    GsRun r = run_gs64(a, g, ru, rd, iter_time, (double)(float)w /* V1 reads omega with stof */, enable_break, tol,
                       true, nullptr, nullptr);
    std::cout << "evalutate ir drop using general GS-method, w = " << (float)w
              << " , print to file fastmode_nodebasedGeneralGS.out" << std::endl;
    std::cout << "time = " << r.seconds << " s" << std::endl;
    if (r.record != -1)
        std::cout << "threshould break, iter times =  " << r.record << std::endl;
    // the general solve's "error I sum" (2dArray.cc:1316-1345): interior
    // wordline nodes and bitline rows 0 .. n-2
    const double gw = 1.0 / a.RD[0][0];
    auto U = [&](int i, int j) { return r.vup[(size_t)i * m + j]; };
    auto D = [&](int i, int j) { return r.vdown[(size_t)i * m + j]; };
    auto G = [&](int i, int j) { return g[(size_t)i * m + j]; };
    double err = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < m - 1; ++j)
            err += std::fabs((U(i, j - 1) - U(i, j) + U(i, j + 1) - U(i, j)) * gw + (D(i, j) - U(i, j)) * G(i, j));
        if (m > 1)
            err += std::fabs((U(i, m - 2) - U(i, m - 1)) * gw + (D(i, m - 1) - U(i, m - 1)) * G(i, m - 1));
    }
    for (int j = 0; j < m; ++j)
        err += std::fabs((U(0, j) - D(0, j)) * G(0, j) + (n > 1 ? D(1, j) - D(0, j) : 0) * gw);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 0; j < m; ++j)
            err += std::fabs((U(i, j) - D(i, j)) * G(i, j) + (D(i - 1, j) - D(i, j) + D(i + 1, j) - D(i, j)) * gw);
    std::cout << "error I sum = " << err << std::endl;
    std::ofstream of("fastmode_nodebasedGeneralGS.out");
    of << "error I sum = " << err << std::endl;
    of << "vup = " << std::endl;
    of << std::scientific << std::setprecision(6);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(14) << U(i, j);
        of << std::endl;
    }
    of << std::endl << "vdown = " << std::endl;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(14) << D(i, j);
        of << std::endl;
    }
    of << std::endl << "iout = " << std::endl;
    for (int j = 0; j < m; ++j)
        of << std::setw(14) << r.iout[j];
    of << std::endl;
    os << "general GS (GPU, " << (r.redblack ? "red-black" : "lexicographic") << "), sweeps = " << r.sweeps
       << ", error I sum = " << err << std::endl;
}

void nodebasedGS(Array2D &a, const std::vector<std::string> &args)
{
    if (args.size() < 3)
        throw CmdError("args number not matched (<3)");
    const int iter_time = parse<int>(args[0]);
    // V1 parses omega with stof (2dArray.cc:65), V2 with stod (concrete_command.h:1111)
    const double w = g_v1 ? (double)std::stof(args[1]) : std::stod(args[1]);
    const int mode = parse<int>(args[2]);
    bool enable_break = false;
    double tol = 0;
    if (args.size() > 3) {
        if (args.size() != 5 && !g_v1)
            throw CmdError("args number not matched (!=5)");
        enable_break = checkYes(args[3]);
        if (enable_break) {
            if (args.size() < 5)
                throw CmdError("args number not matched (!=5)");
            tol = std::stod(args[4]);
        }
    }
    if ((mode & 1) && !g_v1)
        throw CmdError("mode bit 0 is not supported now");
    if (!(a.m >= 2 && a.n >= 2))
        throw CmdError("m and n should greater than 2");
    if (mode & 4)
        throw CmdError("mode bit 2 is not supported now");
    const int n = a.n, m = a.m;
    const size_t nm = (size_t)n * m;
    std::vector<double> g(nm), ru, rd;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const int s = a.arr[i][j];
            if (s < 0 || s >= (int)a.cellR_states.size())
                throw CmdError("cell state " + std::to_string(s) + " has no resistance (cellRStates)");
            g[(size_t)i * m + j] = 1.0 / a.cellR_states[s];
        }
    segments(a, ru, rd, !g_v1 && n == m);
    std::ostringstream os;
    std::vector<double> vup(nm), vdown(nm), iout(m);
    std::vector<double> gu, gd;
    if (mode & 1)
        scichina_guess(a, g, gu, gd);
    int record = -1, sweeps = 0;
    bool redblack = false;
    if ((mode & 2) == 0) {
        GsRun r = run_gs64(a, g, ru, rd, iter_time, w, enable_break, tol, false, (mode & 1) ? &gu : nullptr,
                           (mode & 1) ? &gd : nullptr);
        vup = std::move(r.vup);
        vdown = std::move(r.vdown);
        iout = std::move(r.iout);
        record = r.record;
        sweeps = r.sweeps;
        redblack = r.redblack;
        os << "Duration time = " << r.seconds << " s" << std::endl;
    } else {
        // mode bit 1: no sweeps, the initial state is reported
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < m; ++j) {
                vup[(size_t)i * m + j] = (mode & 1) ? gu[(size_t)i * m + j] : a.v[1][i];
                vdown[(size_t)i * m + j] = (mode & 1) ? gd[(size_t)i * m + j] : a.v[2][j];
            }
        for (int j = 0; j < m; ++j) {
            double s = 0;
            for (int i = 0; i < n; ++i)
                s += (vup[(size_t)i * m + j] - vdown[(size_t)i * m + j]) * g[(size_t)i * m + j];
            iout[j] = s;
        }
    }
    os << "GS method mem usage  = "
       << (sizeof(int) * 5 + sizeof(double) * 3 + nm * sizeof(double) * 5 + m * sizeof(double) * 2 +
           n * sizeof(double))
       << " bytes (Compuated by simply multiplying the number"
          " of variables by the variable space)"
       << std::endl;
    if (record != -1)
        os << "threshould break, iter times =  " << record << std::endl;
    else
        os << "iter times = " << iter_time << std::endl;
    // partial-KCL residual sum, as concrete_command.h:1293-1319
    const double gw = 1.0 / a.RD[0][0];
    auto U = [&](int i, int j) { return vup[(size_t)i * m + j]; };
    auto D = [&](int i, int j) { return vdown[(size_t)i * m + j]; };
    auto G = [&](int i, int j) { return g[(size_t)i * m + j]; };
    double err = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < m - 1; ++j)
            err += std::fabs((U(i, j - 1) - U(i, j) + U(i, j + 1) - U(i, j)) * gw + (D(i, j) - U(i, j)) * G(i, j));
        if (m > 1)
            err += std::fabs((U(i, m - 2) - U(i, m - 1)) * gw + (D(i, m - 1) - U(i, m - 1)) * G(i, m - 1));
    }
    for (int j = 0; j < m; ++j)
        err += std::fabs((U(0, j) - D(0, j)) * G(0, j) + (n > 1 ? D(1, j) - D(0, j) : 0) * gw);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 0; j < m; ++j)
            err += std::fabs((U(i, j) - D(i, j)) * G(i, j) + (D(i - 1, j) - D(i, j) + D(i + 1, j) - D(i, j)) * gw);
    std::ofstream of("fastmode_nodebasedGS.out");
    of << "error I sum = " << err << std::endl;
    os << "error I sum = " << err << std::endl;
    of << "vup = " << std::endl;
    of << std::scientific << std::setprecision(12);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(20) << U(i, j);
        of << std::endl;
    }
    of << std::endl << "vdown = " << std::endl;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(20) << D(i, j);
        of << std::endl;
    }
    of << std::endl << "iout = " << std::endl;
    for (int j = 0; j < m; ++j)
        of << std::setw(20) << iout[j];
    of << std::endl;
    std::cout << "evaluate ir drop using GS-method (GPU, " << (redblack ? "red-black" : "lexicographic")
              << "), w = " << w << ", sweeps = " << sweeps << ", print to file fastmode_nodebasedGS.out" << std::endl;
    if (record != -1)
        std::cout << "threshould break, iter times =  " << record << std::endl;
    if (general_boundary(a)) // V1's dispatcher (2dArray.cc:84-110); V2 has no general solver
        generalGS(a, g, iter_time, w, enable_break, tol, os);
    a.log(os.str());
}

// `fastsolve <iter_times> [table 0] [enable_break break_th]`: IR2S with the
// V1 CPU solver's semantics (crossbarHspice/2dArray.cc:318-327, 1691-1900;
// ir_drop_fast_solver.h:24-163) on the GPU through fadesim_ir2s; output
// fastmode.out in the reference's format.
void fastsolve(Array2D &a, const std::vector<std::string> &args)
{
    if (args.empty())
        throw CmdError("fastsolve needs iter_times");
    const int iter_times = parse<int>(args[0]);
    int from_table = 0, enable_break = 0;
    double break_th = 1e-9;
    if (args.size() > 1)
        from_table = parse<int>(args[1]);
    if (args.size() > 2)
        enable_break = parse<int>(args[2]);
    if (args.size() > 3)
        break_th = std::stod(args[3]);
    if (from_table)
        throw CmdError("fastsolve: initial voltages from a table are not supported");
    if (!(a.m >= 2 && a.n >= 2))
        throw CmdError("m and n should greater than 2");
    const int n = a.n, m = a.m;
    const size_t nm = (size_t)n * m;
    std::vector<double> g(nm);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const int st = a.arr[i][j];
            if (st < 0 || st >= (int)a.cellR_states.size())
                throw CmdError("cell state " + std::to_string(st) + " has no resistance (cellRstates)");
            g[(size_t)i * m + j] = 1.0 / a.cellR_states[st];
        }
    const double r_wire = a.RD[0][0]; // one wire resistance: g_wire = 1 / RDLine[0][0]
    const double gw = 1.0 / r_wire;
    const double r_load = a.r_load[1] < 0 ? 1.0 / gw : a.r_load[1];
    std::vector<double> vup(nm), vdown(nm), iout(m);
    int count = 0;
    double *d_g, *d_vin, *d_vout, *d_vup, *d_vdn, *d_iout;
    int *d_cnt;
    HIPCHK(hipMalloc(&d_g, nm * 8));
    HIPCHK(hipMalloc(&d_vin, n * 8));
    HIPCHK(hipMalloc(&d_vout, m * 8));
    HIPCHK(hipMalloc(&d_vup, nm * 8));
    HIPCHK(hipMalloc(&d_vdn, nm * 8));
    HIPCHK(hipMalloc(&d_iout, m * 8));
    HIPCHK(hipMalloc(&d_cnt, sizeof(int)));
    HIPCHK(hipMemcpy(d_g, g.data(), nm * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_vin, a.v[1].data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_vout, a.v[2].data(), m * 8, hipMemcpyHostToDevice));
    fadesim_ir2s_desc d{};
    d.dtype_bytes = 8;
    d.layout = FADESIM_IR2S_PLAIN;
    d.d_vin = d_vin;
    d.d_G = d_g;
    d.d_vout = d_vout;
    d.batch_size = d.numX = d.numY = 1;
    d.row_size = n;
    d.col_size = m;
    d.iter_times = iter_times;
    d.r_wire = r_wire;
    d.r_source = d.r_sink = r_load;
    d.divide_by_g = 1;
    d.carry_residue = 0;
    d.enable_break = enable_break ? FADESIM_IR2S_BREAK_EACH : 0;
    d.break_threshold = break_th;
    d.d_out = d_iout;
    d.d_vup = d_vup;
    d.d_vdown = d_vdn;
    d.d_iters = d_cnt;
    HIPCHK(hipDeviceSynchronize());
    auto t0 = std::chrono::steady_clock::now();
    int rc = fadesim_ir2s(&d, nullptr);
    if (rc != FADESIM_OK)
        throw CmdError(std::string("solver: ") + fadesim_last_error());
    HIPCHK(hipDeviceSynchronize());
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    HIPCHK(hipMemcpy(vup.data(), d_vup, nm * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vdown.data(), d_vdn, nm * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(iout.data(), d_iout, m * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&count, d_cnt, sizeof(int), hipMemcpyDeviceToHost));
    for (void *p : {(void *)d_g, (void *)d_vin, (void *)d_vout, (void *)d_vup, (void *)d_vdn, (void *)d_iout,
                    (void *)d_cnt})
        HIPCHK(hipFree(p));
    // the V1 "error I sum" (2dArray.cc:1830-1866): every node's KCL with the end
    // segments at g_wire, the sink at 0 V
    auto U = [&](int i, int j) { return vup[(size_t)i * m + j]; };
    auto D = [&](int i, int j) { return vdown[(size_t)i * m + j]; };
    auto G = [&](int i, int j) { return g[(size_t)i * m + j]; };
    const std::vector<double> &vin = a.v[1];
    double err = 0;
    for (int i = 0; i < n; ++i) {
        err += std::fabs((vin[i] - U(i, 0) + (m > 1 ? U(i, 1) - U(i, 0) : 0)) * gw + (D(i, 0) - U(i, 0)) * G(i, 0));
        for (int j = 1; j < m - 1; ++j)
            err += std::fabs((U(i, j - 1) - U(i, j) + U(i, j + 1) - U(i, j)) * gw + (D(i, j) - U(i, j)) * G(i, j));
        if (m > 1)
            err += std::fabs((U(i, m - 2) - U(i, m - 1)) * gw + (D(i, m - 1) - U(i, m - 1)) * G(i, m - 1));
    }
    for (int j = 0; j < m; ++j)
        err += std::fabs((U(0, j) - D(0, j)) * G(0, j) + (n > 1 ? D(1, j) - D(0, j) : 0) * gw);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 0; j < m; ++j)
            err += std::fabs((U(i, j) - D(i, j)) * G(i, j) + (D(i - 1, j) - D(i, j) + D(i + 1, j) - D(i, j)) * gw);
    if (n > 1)
        for (int j = 0; j < m; ++j)
            err += std::fabs((0 - D(n - 1, j) + D(n - 2, j) - D(n - 1, j)) * gw + (U(n - 1, j) - D(n - 1, j)) * G(n - 1, j));
    std::ofstream of("fastmode.out");
    of << "error I sum = " << err << std::endl;
    of << "vup = " << std::endl;
    of << std::scientific << std::setprecision(12);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(20) << U(i, j);
        of << std::endl;
    }
    of << std::endl << "vdown = " << std::endl;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j)
            of << std::setw(20) << D(i, j);
        of << std::endl;
    }
    of << std::endl << "iout = " << std::endl;
    for (int j = 0; j < m; ++j)
        of << std::setw(20) << iout[j];
    of << std::endl;
    const int record = enable_break ? count : iter_times;
    std::cout << "start fast solve" << std::endl << "ir2s fast mode time = " << seconds << " s" << std::endl;
    std::cout << "ir2s fast iter times = " << record << std::endl;
    std::cout << "error I sum = " << err << std::endl;
    std::cout << "fast solve, iter_times = " << iter_times << ", print to file fastmode.out" << std::endl;
    std::ostringstream os;
    os << "fastsolve " << iter_times << " (GPU IR2S), iter times = " << record << ", error I sum = " << err;
    a.log(os.str());
}

void run_line(Array2D &a, const std::string &line)
{
    std::istringstream iss(line);
    std::vector<std::string> tok;
    for (std::string s; iss >> s;)
        tok.push_back(s);
    if (tok.empty())
        return;
    const std::string cmd = tok[0];
    std::vector<std::string> args(tok.begin() + 1, tok.end());
    if (cmd == "%" || cmd == "#" || cmd.rfind("//", 0) == 0 || cmd.rfind("%", 0) == 0)
        return;
    std::ostringstream os;
    auto nargs = [&](size_t k) {
        if (args.size() != k)
            throw CmdError("Failed to parse argument due to the number of args not match.");
    };
    if (cmd == "arraySize") {
        nargs(2);
        const int n = parse<int>(args[0]), m = parse<int>(args[1]);
        if (n < 1 || m < 1)
            throw CmdError("array size must be positive");
        a.setSize(n, m);
        os << "Set array size to ( " << n << ", " << m << " )";
    } else if (cmd == "line_resistance" || (cmd == "setLineR" && args.size() == 1)) {
        nargs(1);
        need_size(a);
        const double r = parse<double>(args[0]);
        for (auto &row : a.RU)
            for (auto &x : row)
                x = r;
        for (auto &row : a.RD)
            for (auto &x : row)
                x = r;
        os << "Set all line resistance to " << r;
    } else if (cmd == "setLineR") {
        nargs(4);
        need_size(a);
        const int x = parse<int>(args[1]), y = parse<int>(args[2]);
        const double r = parse<double>(args[3]);
        if (args[0] == "up") {
            if (x < 0 || x >= a.n || y < 0 || y > a.m || r < 0)
                throw CmdError("Arg Direction not matched:setLineR (up|down) x:int<n y:int<m r:double>0");
            a.RU[x][y] = r;
            os << "RULine (" << x << ", " << y << ") to " << r;
        } else if (args[0] == "down") {
            if (x < 0 || x >= a.m || y < 0 || y > a.n || r < 0)
                throw CmdError("Arg Direction not matched:setLineR (up|down) x:int<n y:int<m r:double>0");
            a.RD[x][y] = r;
            os << "RDLine (" << x << ", " << y << ") to " << r;
        } else {
            throw CmdError("Arg Direction not matched:setLineR up|down x:int y:int r:double");
        }
    } else if (cmd == "rload") {
        nargs(2);
        need_size(a);
        const std::string md = args[0];
        const double r = parse<double>(args[1]);
        if (md == "all") {
            for (auto &row : a.RU)
                row.front() = row.back() = r;
            for (auto &row : a.RD)
                row.front() = row.back() = r;
            for (double &x : a.r_load)
                x = r;
        } else if (md == "left" || md == "right") {
            for (auto &row : a.RU)
                (md == "left" ? row.front() : row.back()) = r;
            a.r_load[md == "left" ? 1 : 3] = r;
        } else if (md == "up" || md == "down") {
            for (auto &row : a.RD)
                (md == "up" ? row.front() : row.back()) = r;
            a.r_load[md == "up" ? 0 : 2] = r;
        } else {
            throw CmdError("Arg str not matched");
        }
        os << "set " << md << " line r_load to " << r << " omega";
    } else if (cmd == "setUseLine" || cmd == "setNotUseLine" || cmd == "setUseLine++") {
        need_size(a);
        if (args.empty())
            throw CmdError("Failed to parse argument with 0 args.");
        const int use = cmd == "setNotUseLine" ? 0 : (cmd == "setUseLine" ? 1 : 2);
        const int d = getDir(args[0]);
        for (size_t k = 1; k < args.size(); ++k) {
            const int x = parse<int>(args[k]);
            if (x == -1)
                for (auto &b : a.bv[d])
                    b = use;
            else
                a.bv[d].at(x) = use;
        }
        os << (use ? "set use Direction " : "set not use Direction ") << d;
    } else if (cmd == "setLineV") {
        need_size(a);
        if (args.size() != 3)
            throw CmdError("Args size not matched!");
        const int d = getDir(args[0]);
        const int x = parse<int>(args[1]);
        const double val = std::stod(args[2]);
        if (x == -1)
            for (auto &vv : a.v[d])
                vv = val;
        else {
            if (x < 0 || x >= (int)a.v[d].size())
                throw CmdError("line index out of range");
            a.v[d][x] = val;
        }
        os << "set direction " << d << (x == -1 ? " all line" : " line No." + std::to_string(x))
           << " to voltage " << val << " [d=(0, 1, 2, 3)->(up, left, down, right)]";
    } else if (cmd == "setCellR") {
        nargs(3);
        need_size(a);
        const int x = parse<int>(args[0]), y = parse<int>(args[1]), z = parse<int>(args[2]);
        os << "setCellR " << x << ' ' << y << ' ' << z;
        for (int i = 0; i < a.n; ++i) {
            if (x != -1 && i != x)
                continue;
            for (int j = 0; j < a.m; ++j)
                if (y == -1 || j == y)
                    a.arr[i][j] = z;
        }
        if ((x != -1 && (x < 0 || x >= a.n)) || (y != -1 && (y < 0 || y >= a.m)))
            throw CmdError("cell index out of range");
    } else if (cmd == "cellRStates" || cmd == "cellRstates") {
        try {
            const int k = std::stoi(args.at(0));
            a.cellR_states.assign(k, 0.0);
            for (int i = 0; i < k; ++i)
                a.cellR_states.at(i) = std::stod(args.at(i + 1));
        } catch (const std::exception &) {
            throw CmdError("cell r set fail!");
        }
    } else if (cmd == "inputVStates" || cmd == "inputVstates") {
        try {
            const int k = std::stoi(args.at(0));
            a.inputV_states.assign(k, 0.0);
            for (int i = 0; i < k; ++i)
                a.inputV_states.at(i) = std::stod(args.at(i + 1));
        } catch (const std::exception &) {
            throw CmdError("input set fail!");
        }
        os << "set cell input v states, note that, this is useful when calling randSetInputV";
    } else if (cmd == "seed") {
        nargs(1);
        a.rng.seed(parse<unsigned long long>(args[0]));
        a.seeded = true;
        os << "random seed " << args[0];
    } else if (cmd == "randSetR") {
        need_size(a);
        std::vector<double> p;
        for (auto &s : args)
            p.push_back(parse<double>(s));
        if (p.size() != a.cellR_states.size())
            throw CmdError("p size should equal to item size");
        auto seq = sample(a, p, (size_t)a.n * a.m);
        for (int i = 0; i < a.n; ++i)
            for (int j = 0; j < a.m; ++j)
                a.arr[i][j] = seq[(size_t)i * a.m + j];
        os << "randSetR finish";
    } else if (cmd == "randSetInputV") {
        need_size(a);
        if (args.empty())
            throw CmdError("Failed to parse argument with 0 args.");
        const int d = getDir(args[0]);
        std::vector<double> p;
        for (size_t k = 1; k < args.size(); ++k)
            p.push_back(parse<double>(args[k]));
        if (p.size() != a.inputV_states.size())
            throw CmdError("p size should equal to item size");
        auto seq = sample(a, p, a.v[d].size());
        for (size_t i = 0; i < seq.size(); ++i)
            a.v[d][i] = a.inputV_states[seq[i]];
        os << "randSetInputV " << args[0] << " finish";
    } else if (cmd == "senseBitlineI" || cmd == "senseCellV" || cmd == "senseWordlineI" ||
               cmd == "useRtypeReRAMCell" || cmd == "dc" || cmd == "ac") {
        os << cmd << " noted";
    } else if (cmd == "selector") {
        if (!args.empty() && checkYes(args[0]))
            throw CmdError("selector cells are not supported by nodebasedGSMethod");
        os << "selector no";
    } else if (cmd == "topString" || cmd == "bottomString" || cmd == "build" || cmd == "capacity" ||
               cmd == "capacitance" || cmd == "setLineV++" || cmd == "simpleWriteOne" ||
               cmd == "simpleWriteZero" || cmd == "simpleReadOut" || cmd == "fastmode" ||
               cmd.rfind("ir_", 0) == 0) {
        os << cmd << ": not handled by this build (SPICE generation and other solvers); ignored";
        std::cout << os.str() << std::endl;
    } else if (cmd == "nodebasedGSMethod") {
        nodebasedGS(a, args);
        return;
    } else if (cmd == "fastsolve") {
        fastsolve(a, args);
        return;
    } else {
        throw CmdError("Unknown command: " + cmd);
    }
    a.log(os.str());
}

} // namespace

int main(int argc, char *argv[])
{
    std::string input = "config";
    bool have_input = false;
    for (int k = 1; k < argc; ++k) {
        const std::string s = argv[k];
        if (s == "-v1")
            g_v1 = true;
        else if (s == "-debug" || s == "-3DH" || s == "-3dh")
            std::cout << "option " << s << " is not handled by this build; ignored" << std::endl;
        else if (!have_input) {
            input = s;
            have_input = true;
        }
    }
    Array2D arr;
    int line = 0;
    const bool strict = std::getenv("FADESIM_STRICT_EXIT") && std::string(std::getenv("FADESIM_STRICT_EXIT")) == "1";
    try {
        std::ifstream fin(input);
        if (!fin)
            throw CmdError("cannot open config file " + input);
        for (std::string s; std::getline(fin, s);) {
            ++line;
            run_line(arr, s);
        }
        std::cout << "Log print to default.log" << std::endl;
        std::ofstream os("default.log");
        arr.printLog(os);
        std::cout << "--------Run end && Succeed---------" << std::endl;
    } catch (const std::exception &e) {
        std::cout << "Error(s) occured in line " << line << ": " << e.what() << std::endl;
        arr.printLog(std::cout);
        return strict ? 1 : 0;
    }
    return 0;
}
