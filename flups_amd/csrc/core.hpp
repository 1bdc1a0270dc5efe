// Host-side model of the FLUPS solve path for one GPU.
//
// A solve is a list of pencil passes over dense device arrays ("stages"):
// stage ip holds the field after the transforms of plans 0..ip, dimension 0
// fastest, complex interleaved when complex, components slowest. The pencil
// switches of the reference (SwitchTopoX, src/SwitchTopoX_a2a.cpp:174-247)
// become the gather/scatter index maps of each pass; nothing is exchanged.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace flups {

enum Bc { EVEN = 0, ODD = 1, PER = 3, UNB = 4, NONE = 9 };
enum Kind { REDFT00 = 0, REDFT10, REDFT01, REDFT11, RODFT00, RODFT10, RODFT01, RODFT11 };
enum PType { SYMSYM = 2, MIXUNB = 5, PERPER = 6, UNBUNB = 8, EMPTY = 18 };
enum Post { POST_NONE = 0, POST_FIRST = 1, POST_LAST = 2, POST_PERIOD = 4 };
enum Op { OP_NONE = 0, OP_R2R = 1, OP_R2C = 2, OP_C2R = 3, OP_C2C = 4 };
enum Green {
    CHAT_2 = 0, LGF_2 = 1, HEJ_2 = 2, HEJ_4 = 3, HEJ_6 = 4, HEJ_8 = 5, HEJ_10 = 6, HEJ_0 = 7,
    LGF_4 = 8, LGF_6 = 9, LGF_8 = 10, MEHR_4L = 11, MEHR_6L = 12, MEHR_4F = 13, MEHR_6F = 14
};
enum Diff { NOD = 0, SPE = 1, FD2 = 2, FD4 = 4, FD6 = 6 };
constexpr int FWD = -1, BWD = 1;
// signs of the advanced interface (FLUPS_FORWARD, FLUPS_BACKWARD, FLUPS_BACKWARD_DIFF)
constexpr int SIGN_FORWARD = -1, SIGN_BACKWARD = 1, SIGN_BACKWARD_DIFF = 2;

// failed checks unwind to the C entry point, which records the message
struct Error {
    std::string msg;
};
[[noreturn]] void fail(const char* fmt, ...);
#define FLUPS_REQUIRE(c, ...) \
    do {                      \
        if (!(c)) ::flups::fail(__VA_ARGS__); \
    } while (0)
#define HIP_CHECK(x)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) ::flups::fail("HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// one direction of the transform (reference: FFTW_plan_dim, src/FFTW_plan_dim.hpp:20-162)
struct DirPlan {
    int dim = 0, lda = 1, sign = FWD;
    bool green = false, node = false;
    int bcl[3] = {9, 9, 9}, bcr[3] = {9, 9, 9};
    PType type = EMPTY;
    bool r2c = false, spectral = false;
    int n_out = 1, fieldstart = 0;
    double symstart = 0, normfact = 1, volfact = 1, kfact = 0, koffset = 0;
    int n_in[3] = {1, 1, 1}, start_in[3] = {0, 0, 0}, start_out[3] = {0, 0, 0}, post[3] = {0, 0, 0}, kind[3] = {0, 0, 0};
    bool imult[3] = {false, false, false};
    bool r2c_by_fft() const { return r2c && !spectral; }
};

DirPlan make_dirplan(int dim, int lda, const int* bcl, const int* bcr, const double h[3], const double L[3], int sign,
                     bool green, bool node);
void init_dirplan(DirPlan& p, const int size[3], bool isComplex);
PType plan_type(int l, int r);

// extents of one stage array
struct Shape {
    int n[3] = {1, 1, 1};
    int nf = 1, lda = 1;
    size_t per() const { return (size_t)n[0] * n[1] * n[2] * nf; }   // doubles per component
    size_t count() const { return per() * lda; }
    void to_complex(int ax) { if (nf == 1) { nf = 2; n[ax] /= 2; } }
    void to_real(int ax) { if (nf == 2) { nf = 1; n[ax] *= 2; } }
};

// a strided view of a 3-D (x components) array of doubles
struct View {
    void* p = nullptr;
    int es = 8;                   // bytes of one real: 8 (double) or 4 (float, FP32 stage arrays)
    int n[3] = {1, 1, 1};
    int nf = 1;
    long long s[3] = {0, 0, 0};   // stride in elements (an element is nf reals)
    long long sc = 0;             // component stride in reals
};
View dense_view(void* p, const Shape& sh, int es = 8);

// everything one launch of the generic pencil kernel needs (one component)
struct PassDesc {
    int d = 0;
    int op = OP_NONE, sign = FWD;
    int n = 1, kind = 0, si = 0, so = 0, post = 0, n_out_plan = 1, r2r_stride = 1;
    int pen_in = 1, nf_in = 1, pen_out = 1, nf_out = 1;   // pencil extents (elements) before/after
    int gshift = 0, sshift = 0;
    int M = 1, nst = 0, rad[16] = {0};
    const double* tw = nullptr;     // exp(-2 pi i t / M), t < M, interleaved
    const void* src = nullptr;      // reals of the input array: double, or float with src_f32
    int src_nd = 1, src_nf = 1;
    long long src_s[3] = {0, 0, 0};
    void* dst = nullptr;            // reals of the output array: double, or float with dst_f32
    int dst_nd = 1, dst_nf = 1;
    int src_f32 = 0, dst_f32 = 0;   // storage of the pass's input / output (arithmetic is FP64)
    long long dst_s[3] = {0, 0, 0};
    int n1 = 1, n2 = 1;     // extents of the two other dims ((d+1)%3, (d+2)%3)
    int ppb = 1;            // pencils per block
    int fa = 1, fb = 2;     // pencil grid: dim fa varies fastest between pencils (the one with the smaller stride)
    int na = 1;             // extent along fa
    int alias = 0;          // 1: the second FFT buffer reuses the pencil buffer
    int fastk = 0;          // fast kernels of rowpass.hip instead of the generic one: 1 row r2r, 2 paired real columns,
                            // 3 row r2c / c2r, 4 column r2r, 5 column c2c, 6 row c2c
    int vec = 0;            // fast kernels: pencils start 16-byte aligned in source and destination
    int nt = 0;             // row r2r kernel: its vector accesses are non-temporal (both arrays 128 MiB or more)
    const double* aux = nullptr;   // fast kernels: pre / post twiddle table
};

// the last direction of the general path as one r2c -> normfact G -> c2r launch (rowpass.hip):
// col_rz, real columns along the last dimension, pairs of neighbours along x; row_rz, real rows
// along x (x the last direction); nin real values per column / row, transform length M
struct RzArgs {
    void* d = nullptr;           // field before the transform, in place: (nx, ny, nin) reals, x fastest
    const void* g = nullptr;     // Green's function on the spectral layout (nx, ny, M/2 + 1) complex
    int f32 = 0;                 // data and Green's function stored as float (FP64 arithmetic)
    const double* tw = nullptr;  // M-point twiddle table
    int nx = 1, ny = 1, nin = 1, M = 1, lda = 1;
    long long s_col = 1;         // reals between rows of a column (nx ny)
    long long comp = 0;          // reals between components
    double normfact = 1;
    const double* twn = nullptr; // row_rz: n-point table (n = 2H); tw is then the (H / pf)-point one
    int rows = 1;                // row_rz: rows of the field (ny nz)
};

}  // namespace flups
