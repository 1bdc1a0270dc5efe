// C entry points of the oracle for ctypes (tests/, smoke, bench cpu_baseline).
// Test infrastructure only.
#include <chrono>
#include <cstring>
#include <exception>
#include <string>

#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;
int fail(const std::exception& e) { g_err = e.what(); return -1; }
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

int oracle_validation(int center, const int* bc6, int kernel, int N, int lda, const char* kdir, double* err) {
    try {
        validation_case(center, bc6, kernel, N, lda, kdir ? kdir : ".", err);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int oracle_vtube(int vdir, int symx, int symy, int N, int kernel, int order, double sigma, const char* kdir, double* err) {
    try {
        vtube_case(vdir, symx, symy, N, kernel, order, sigma, kdir ? kdir : ".", err);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// manufactured rhs / exact solution of a validation case, and its grid
int oracle_validation_fields(int center, const int* bc6, int N, int lda, int* nglob, double* h, double* rhs, double* sol) {
    try {
        validation_geometry(center, bc6, N, nglob, h);
        if (rhs && sol) validation_fields(center, bc6, N, lda, rhs, sol);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}
// vortex-tube rhs / exact solution and its [3][2][3] boundary conditions
int oracle_vtube_fields(int vdir, int symx, int symy, int N, double sigma, int* bc, double* rhs, double* sol) {
    try {
        vtube_case(vdir, symx, symy, N, 0, 1, sigma, ".", nullptr, bc, rhs, sol);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// bc: [3][2][lda] ints; rhs/field: dense, dim 0 fastest, components slowest
int oracle_solve(const int* nglob, const int* bc, int lda, const double* h, const double* L, int center, int odiff, int kernel,
                 double alpha, const char* kdir, int type, const double* rhs, double* field) {
    try {
        Solver s(nglob, bc, lda, h, L, center == 0, (DiffType)odiff);
        s.kernel_dir = kdir ? kdir : ".";
        s.set_green((GreenType)kernel);
        s.set_alpha(alpha);
        s.setup();
        s.solve(field, rhs, (SolverType)type);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// repeated solves on one set-up solver (for the CPU baseline timing); returns seconds of the last `reps` solves
int oracle_solve_timed(const int* nglob, const int* bc, int lda, const double* h, const double* L, int center, int odiff,
                       int kernel, double alpha, const char* kdir, int type, const double* rhs, double* field, int warm,
                       int reps, double* seconds) {
    try {
        Solver s(nglob, bc, lda, h, L, center == 0, (DiffType)odiff);
        s.kernel_dir = kdir ? kdir : ".";
        s.set_green((GreenType)kernel);
        s.set_alpha(alpha);
        s.setup();
        for (int i = 0; i < warm; ++i) s.solve(field, rhs, (SolverType)type);
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < reps; ++i) s.solve(field, rhs, (SolverType)type);
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// the same with a time budget: `warm` solves, then timed solves until `budget_s` seconds have
// passed (at least min_reps, at most max_reps); *reps_done and *seconds report the timed part
int oracle_solve_timed_for(const int* nglob, const int* bc, int lda, const double* h, const double* L, int center,
                           int odiff, int kernel, double alpha, const char* kdir, int type, const double* rhs,
                           double* field, int warm, double budget_s, int min_reps, int max_reps, int* reps_done,
                           double* seconds) {
    try {
        Solver s(nglob, bc, lda, h, L, center == 0, (DiffType)odiff);
        s.kernel_dir = kdir ? kdir : ".";
        s.set_green((GreenType)kernel);
        s.set_alpha(alpha);
        s.setup();
        for (int i = 0; i < warm; ++i) s.solve(field, rhs, (SolverType)type);
        const auto t0 = std::chrono::steady_clock::now();
        int r = 0;
        double el = 0;
        while (r < max_reps && (r < min_reps || el < budget_s)) {
            s.solve(field, rhs, (SolverType)type);
            ++r;
            el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        *reps_done = r;
        *seconds = el;
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// spectral-space data after the forward transforms (for stage-wise comparisons)
int oracle_spectral_size(const int* nglob, const int* bc, int lda, const double* h, const double* L, int center, int* n_out,
                         int* nf) {
    try {
        Solver s(nglob, bc, lda, h, L, center == 0, NOD);
        Topo t = s.hat[s.ndim - 1];
        if (s.fwd[s.ndim - 1].r2c) t.to_complex(s.fwd[s.ndim - 1].dim);
        for (int d = 0; d < 3; ++d) n_out[d] = t.n[d];
        *nf = t.nf;
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int oracle_green(const int* nglob, const int* bc, const double* h, const double* L, int center, int kernel, double alpha,
                 const char* kdir, double* green /* sized by oracle_spectral_size with lda = 1 */) {
    try {
        Solver s(nglob, bc, 1, h, L, center == 0, NOD);
        s.kernel_dir = kdir ? kdir : ".";
        s.set_green((GreenType)kernel);
        s.set_alpha(alpha);
        s.setup();
        std::memcpy(green, s.green.data(), s.green.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int oracle_plan_info(const int* nglob, const int* bc, int lda, const double* h, const double* L, int center, double* normfact,
                     double* volfact, double* kfact, double* koffset, double* symstart, int* order) {
    try {
        Solver s(nglob, bc, lda, h, L, center == 0, NOD);
        *normfact = s.normfact;
        *volfact = s.volfact;
        s.spectral_info(kfact, koffset, symstart);
        for (int i = 0; i < 3; ++i) order[i] = s.fwd[i].dim;
        return s.ndim;
    } catch (const std::exception& e) { return fail(e); }
}

int oracle_c2c(int n, int sign, const double* in, double* out) {
    try { fft_c2c(n, sign, (const cplx*)in, (cplx*)out); return 0; } catch (const std::exception& e) { return fail(e); }
}
int oracle_r2c(int n, const double* in, double* out) {
    try { fft_r2c(n, in, (cplx*)out); return 0; } catch (const std::exception& e) { return fail(e); }
}
int oracle_c2r(int n, const double* in, double* out) {
    try { fft_c2r(n, (const cplx*)in, out); return 0; } catch (const std::exception& e) { return fail(e); }
}
int oracle_r2r(int n, int kind, const double* in, double* out) {
    try { fft_r2r(n, (R2R)kind, in, out); return 0; } catch (const std::exception& e) { return fail(e); }
}
double oracle_lgf1d(int order, int n, double c) {
    try { return lgf1d(order, n, c); } catch (const std::exception& e) { fail(e); return 0.0 / 0.0; }
}
double oracle_lgf_symbol(int order, double kh) { return lgf_symbol(order, kh); }
double oracle_bessel_k0(double x) { return bessel_k0(x); }
double oracle_bessel_k1(double x) { return bessel_k1(x); }
double oracle_expint_e1(double x) { return expint_e1(x); }
double oracle_sine_integral(double x) { return sine_integral(x); }
double oracle_ji0c(double x) { return ji0c(x); }
double oracle_hej_1unb(int type, double r, double k, double sig, int k0) {
    return hej_1unb((GreenType)type, r, k, sig, k0 != 0);
}
double oracle_mehr_1d(int kind, int n, double k1, double k2) { return mehr_1d(kind, n, k1, k2); }

}  // extern "C"
