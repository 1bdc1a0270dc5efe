// Argument blocks of the Green's function and multiply kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace flups {

struct ExpTerm {   // coef * n1^a n2^b n3^c n^-p log(n)^l
    int a, b, c, p, l;
    double coef;
};

struct GreenArgs {
    int type = 0;
    double length = 0, g0 = 0, hgrid = 0;
    double hfact[3] = {0, 0, 0}, kfact[3] = {0, 0, 0}, koffset[3] = {0, 0, 0}, symstart[3] = {0, 0, 0};
    const double* table = nullptr;   // LGF table near the origin (32^3 or 32^2); HEJ_0, 2 unbounded: Ji0c by s = |i|^2
    int ntable = 0;
    const ExpTerm* exp = nullptr;    // LGF far-field expansion
    int nexp = 0;
    // 1-D LGF stencils, orders 2,4,6,8 at rows 0..3: s_0..s_w and p_0..p_w
    double lgf_s[20] = {0}, lgf_p[20] = {0};
};

struct RotArgs {
    int n[3] = {1, 1, 1};
    int nf = 1, order = 1;
    double normfact = 1, h[3] = {0, 0, 0}, koffset[3] = {0, 0, 0}, symstart[3] = {0, 0, 0};
    double kf[3][3][2] = {};
};

void launch_green_fill(int nunb, double* g, const int n[3], int nf, int axis, const GreenArgs& A, const int* lo, hipStream_t s);
void launch_scale(double* g, size_t n, double f, hipStream_t s);
// f32: data and Green's function stored as float (FP32 solves of the general path), FP64 arithmetic
void launch_mult_std(void* d, const void* g, size_t per, int nf, int lda, double normfact, bool f32, hipStream_t s);
void launch_mult_rot(void* d, const void* g, const RotArgs& R, bool f32, hipStream_t s);
void launch_to_float(const double* a, float* b, size_t n, hipStream_t s);

}  // namespace flups
