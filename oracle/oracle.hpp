// CPU FP64 oracle of the FLUPS solve path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
// path is compared against; nothing in flups_amd/ links, includes or calls it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// It restates, in its own code, the algorithm of the reference:
//   plan rules        src/FFTW_plan_dim.cpp:59-161, FFTW_plan_dim_cell.cpp:19-391,
//                     FFTW_plan_dim_node.cpp:19-462
//   dry run / topos   src/Solver.cpp:589-811
//   solve             src/Solver.cpp:1207-1272, 1402-1462, 1471-1542
//   multiply          src/dothemagic_std.ipp:56-83, src/dothemagic_rot.ipp:102-349
//   Green's function  src/Solver.cpp:980-1154, src/green_functions.cpp:27-708
// FFTW's transforms are replaced by the oracle's own mixed-radix Stockham FFT
// and the textbook definitions of the r2r kinds (FFTW manual, "What FFTW
// Really Computes").
#pragma once
#include <complex>
#include <string>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

enum BC { EVEN = 0, ODD = 1, PER = 3, UNB = 4, NONE = 9 };
enum R2R { REDFT00 = 0, REDFT10, REDFT01, REDFT11, RODFT00, RODFT10, RODFT01, RODFT11 };
enum PlanType { SYMSYM = 2, MIXUNB = 5, PERPER = 6, UNBUNB = 8, EMPTY = 18 };
enum Postpro { PP_NONE = 0, NULL_FIRST = 1, NULL_LAST = 2, ENFORCE_PERIOD = 4 };
enum GreenType {
    CHAT_2 = 0, LGF_2 = 1, HEJ_2 = 2, HEJ_4 = 3, HEJ_6 = 4, HEJ_8 = 5, HEJ_10 = 6, HEJ_0 = 7,
    LGF_4 = 8, LGF_6 = 9, LGF_8 = 10, MEHR_4L = 11, MEHR_6L = 12, MEHR_4F = 13, MEHR_6F = 14
};
enum SolverType { STD = 0, ROT = 1 };
enum DiffType { NOD = 0, SPE = 1, FD2 = 2, FD4 = 4, FD6 = 6 };
constexpr int SIGN_FWD = -1, SIGN_BWD = 1;

// ---- 1-D transforms (unnormalised, FFTW conventions) ----
void fft_c2c(int n, int sign, const cplx* in, cplx* out);
void fft_r2c(int n, const double* in, cplx* out);   // n real -> n/2+1
void fft_c2r(int n, const cplx* in, double* out);   // n/2+1 -> n real
void fft_r2r(int n, R2R kind, const double* in, double* out);

// ---- special functions ----
double bessel_k0(double x);
double bessel_k1(double x);
double expint_e1(double x);
double lgf1d(int order, int n, double c);          // 1-D lattice Green's function
double sine_integral(double x);                     // Si(x)
double ji0c(double x);                              // int_0^x (1 - J0(u)) / u du, the reference's series
// HEJ_2..10 with one unbounded direction: G(k, r) (k > 0) or the k = 0 form
double hej_1unb(GreenType type, double r, double k, double sig, bool k0);
double mehr_1d(int kind, int n, double k1, double k2);   // 1-D Mehrstellen LGF: 0 4L, 1 6L, 2 4F, 3 6F
double lgf_symbol(int order, double kh);

// ---- one direction of the transform ----
struct Plan {
    int  lda = 1, dim = 0, sign = SIGN_FWD;
    bool green = false, node = false;
    std::vector<int> bcl, bcr;   // per component
    PlanType type = EMPTY;
    double volfact = 1, normfact = 1, kfact = 0, koffset = 0, symstart = 0;
    bool   spectral = false, r2c = false;
    int    n_out = 1, fieldstart = 0;
    std::vector<int>  n_in, start_in, start_out, postpro, kind;
    std::vector<bool> imult;

    Plan(int lda, int dim, const double h[3], const double L[3], const int* bl, const int* br, int sign, bool green, bool node);
    void init(const int size[3], bool isComplex);
    bool r2c_by_fft() const { return r2c && !spectral; }
};

// ---- a (single rank) topology: sizes of one pencil stage ----
struct Topo {
    int n[3] = {1, 1, 1};
    int nf = 1;
    int lda = 1;
    int axis = 0;   // pencil direction (only used for the Green's table index order)
    size_t count() const { return (size_t)n[0] * n[1] * n[2] * nf * lda; }
    void to_complex(int ax) { if (nf == 1) { nf = 2; n[ax] /= 2; } }
    void to_real(int ax) { if (nf == 2) { nf = 1; n[ax] *= 2; } }
};

struct Solver {
    int lda = 1, ndim = 3;
    int nglob[3];
    double h[3], L[3];
    bool node = false;
    DiffType odiff = NOD;
    GreenType gtype = CHAT_2;
    double alpha = 2.0;
    std::string kernel_dir;

    std::vector<Plan> fwd, bwd, bwd_diff, gplan;   // sorted by execution order
    Topo hat[3], gtopo[3];
    double normfact = 1, volfact = 1;
    std::vector<double> green;   // on gtopo[ndim-1]
    bool setup_done = false;

    Solver(const int nglob[3], const int* bc /*[3][2][lda]*/, int lda, const double h[3], const double L[3],
           bool node, DiffType odiff);
    void set_green(GreenType t) { gtype = t; }
    void set_alpha(double a) { alpha = a; }
    void setup();
    // rhs and field: x-fastest dense arrays of nglob, lda slowest
    void solve(double* field, const double* rhs, SolverType type);
    // spectral info in the sense of Solver::get_spectralInfo (Solver.hpp:222-229)
    void spectral_info(double kfact[3], double koffset[3], double symstart[3]) const;

    // stages exposed for tests
    std::vector<double> forward(const double* rhs) const;           // returns data on hat[ndim-1]
    void multiply(std::vector<double>& data, SolverType type) const;
    void backward(std::vector<double>& data, double* field, SolverType type) const;

   private:
    void init_topos(std::vector<Plan>& plans, Topo* topos, bool isGreen);
    void compute_green();
};

// validation harness of samples/validation (manufactured solutions)
void validation_case(int center, const int bc6[6], int kernel, int N, int lda, const std::string& kdir, double* err /*[2*lda]*/);
// vortex-tube ROT harness of samples/solve_vtube
void vtube_case(int vdir, int symx, int symy, int N, int kernel, int order, double sigma, const std::string& kdir, double* err /*[6]*/,
                int* bc_out = nullptr, double* rhs_out = nullptr, double* sol_out = nullptr);
void validation_geometry(int center, const int bc6[6], int N, int nglob[3], double h[3]);
void validation_fields(int center, const int bc6[6], int N, int lda, double* rhs, double* sol);

}  // namespace oracle
