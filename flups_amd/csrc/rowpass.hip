// Fast kernels of the general path for its most common pencil shapes, used in
// place of the generic pass kernel (generic.hip) when a pass fits them.
//
// row_r2r: real-to-real transforms of kinds REDFT10/01/11 and RODFT10/01/11
// (src/FFTW_plan_dim.cpp:635-720 runs them through FFTW) along a contiguous
// pencil of even length N = 2H, H a power of two in [16, 1024]. Each is one
// complex FFT of length H held in registers (the column FFT of fft_core.hpp)
// with pre- and post-twiddles, instead of the 2N-point extension of the
// generic kernel:
//   REDFT11 (DCT-IV): v[m] = (x[2m] + i x[N-1-2m]) e^{-i pi (m + 1/4)/N}, V = FFT_H(v),
//                     W[k] = V[k] e^{-i pi k/N}, y[2k] = 2 Re W[k], y[N-1-2k] = -2 Im W[k];
//   REDFT10 (DCT-II): u[m] = x[2m], u[N-1-m] = x[2m+1] (m < H), U = FFT_N(u) from the
//                     packed FFT_H of u[2m] + i u[2m+1], y[k] = 2 Re(e^{-i pi k/(2N)} U[k]);
//   REDFT01 (DCT-III): U[k] = e^{i pi k/(2N)} (x[k] - i x[N-k]) (x[N] = 0), u = N IFFT_N(U)
//                     through the packed inverse FFT_H, y[2m] = u[m], y[2m+1] = u[N-1-m];
//   RODFT10 = reversed REDFT10 of (-1)^n x, RODFT01/11 = (-1)^k REDFT01/11 of reversed x.
// ROWS pencils per block, H/8 threads per pencil; the pencil is staged through
// LDS so that HBM is read and written in 16-byte pieces when rows are aligned.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "core.hpp"
#include "fft_core.hpp"

namespace flups {

namespace {

template <int H, int ROWS, int KIND>
__global__ void __launch_bounds__(ROWS * H / 8) row_r2r(PassDesc P) {
    constexpr int T = H / 8, N = 2 * H, HP = H + H / 8;
    __shared__ C2<double> lds[HP * ROWS];
    constexpr int DCT = KIND >= RODFT10 ? KIND - 4 : KIND;
    constexpr bool ALT = KIND == RODFT10;                       // input (-1)^n x, output reversed
    constexpr bool REV = KIND == RODFT01 || KIND == RODFT11;    // input reversed, output (-1)^k
    const int rl = threadIdx.x / T, j = threadIdx.x % T;
    const int g = blockIdx.x * ROWS + rl;
    const bool live = g < P.n1 * P.n2;
    const int gg = live ? g : 0;   // dead rows read pencil 0 and store nothing
    const int ja = gg % P.na, jb = gg / P.na;
    const double* src = P.src + (long long)ja * P.src_s[P.fa] + (long long)jb * P.src_s[P.fb];
    double* dst = P.dst + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
    C2<double>* row = lds + rl * HP;
    double* rd = reinterpret_cast<double*>(row);   // the pencil as N reals
    const C2<double>* A = reinterpret_cast<const C2<double>*>(P.aux);
    const C2<double>* B = A + (H + 1);
    const C2<double>* tw = reinterpret_cast<const C2<double>*>(P.tw);

    // 1. stage the pencil in LDS
    if (P.vec) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = j + t * T;
            const double2 p = reinterpret_cast<const double2*>(src)[q];
            rd[2 * q] = p.x;
            rd[2 * q + 1] = p.y;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t) rd[j + t * T] = src[j + t * T];
    }
    __syncthreads();
    auto xin = [&](int q) -> double {
        const double v = rd[REV ? N - 1 - q : q];
        return (ALT && (q & 1)) ? -v : v;
    };

    // 2. pre-map to the complex sequence of length H
    C2<double> v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int m = j + i * T;
        if (DCT == REDFT11) {
            v[i] = cmul(C2<double>{xin(2 * m), xin(N - 1 - 2 * m)}, A[m]);
        } else if (DCT == REDFT10) {
            auto u = [&](int q) { return q < H ? xin(2 * q) : xin(2 * (N - 1 - q) + 1); };
            v[i] = {u(2 * m), u(2 * m + 1)};
        } else {   // REDFT01
            auto U = [&](int k) {
                const double a = k < N ? xin(k) : 0.0, b = (N - k) < N ? xin(N - k) : 0.0;
                return cmul(A[k], C2<double>{a, -b});
            };
            const C2<double> uk = U(m), uc = cconj(U(H - m));
            const C2<double> e = cadd(uk, uc), f = csub(uk, uc);
            const C2<double> t = cmul(B[m], f);
            v[i] = {e.x - t.y, e.y + t.x};   // e + i t
        }
    }

    // 3. FFT of length H (forward for DCT-II / IV, inverse for DCT-III)
    column_fft<double, H, DCT != REDFT01, 0>(v, row, j, tw);
    __syncthreads();   // every thread is done with the FFT's LDS before the buffer is reused

    // 4. post-map into the LDS row
    if (DCT == REDFT11) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = j + i * T;
            const C2<double> w = cmul(v[i], B[k]);
            rd[2 * k] = 2.0 * w.x;
            rd[N - 1 - 2 * k] = -2.0 * w.y;
        }
    } else if (DCT == REDFT10) {
#pragma unroll
        for (int i = 0; i < 8; ++i) row[j + i * T] = v[i];
        __syncthreads();
        C2<double> U[8], UH = {0, 0};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = j + i * T;
            const C2<double> zk = row[k], zc = cconj(row[(H - k) & (H - 1)]);
            const C2<double> e = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
            const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
            U[i] = cadd(e, cmul(A[k], C2<double>{df.y, -df.x}));   // e + A[k] (zk - zc) / (2i)
        }
        if (j == 0) {   // U[H] = Re Z[0] - Im Z[0]
            const C2<double> z0 = row[0];
            UH = {z0.x - z0.y, 0.0};
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = j + i * T;
            const C2<double> c = cmul(B[k], U[i]);
            rd[k] = 2.0 * c.x;
            if (k > 0) rd[N - k] = -2.0 * c.y;
        }
        if (j == 0) rd[H] = 2.0 * cmul(B[H], UH).x;
    } else {   // REDFT01: v[i] = z[m], u[2m] = Re, u[2m+1] = Im
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int m = j + i * T;
            const int q0 = 2 * m, q1 = 2 * m + 1;
            rd[q0 < H ? 2 * q0 : 2 * (N - 1 - q0) + 1] = v[i].x;
            rd[q1 < H ? 2 * q1 : 2 * (N - 1 - q1) + 1] = v[i].y;
        }
    }
    __syncthreads();

    // 5. store (DST output reflections)
    if (!live) return;
    auto out = [&](int k) -> double {
        const double y = rd[ALT ? N - 1 - k : k];
        return (REV && (k & 1)) ? -y : y;
    };
    if (P.vec) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = j + t * T;
            reinterpret_cast<double2*>(dst)[q] = double2{out(2 * q), out(2 * q + 1)};
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t) dst[j + t * T] = out(j + t * T);
    }
}

template <int H, int ROWS>
void launch_row_h(const PassDesc& P, hipStream_t s) {
    const int npen = P.n1 * P.n2;
    dim3 grid((npen + ROWS - 1) / ROWS);
    dim3 block(ROWS * H / 8);
    switch (P.kind) {
        case REDFT10: hipLaunchKernelGGL((row_r2r<H, ROWS, REDFT10>), grid, block, 0, s, P); break;
        case REDFT01: hipLaunchKernelGGL((row_r2r<H, ROWS, REDFT01>), grid, block, 0, s, P); break;
        case REDFT11: hipLaunchKernelGGL((row_r2r<H, ROWS, REDFT11>), grid, block, 0, s, P); break;
        case RODFT10: hipLaunchKernelGGL((row_r2r<H, ROWS, RODFT10>), grid, block, 0, s, P); break;
        case RODFT01: hipLaunchKernelGGL((row_r2r<H, ROWS, RODFT01>), grid, block, 0, s, P); break;
        case RODFT11: hipLaunchKernelGGL((row_r2r<H, ROWS, RODFT11>), grid, block, 0, s, P); break;
        default: fail("row r2r: kind %d", P.kind);
    }
}

}  // namespace

bool row_r2r_fits(const PassDesc& P) {
    const int k = P.kind;
    const int n = P.n, h = n / 2;
    if (P.op != OP_R2R || !(k == REDFT10 || k == REDFT01 || k == REDFT11 || k == RODFT10 || k == RODFT01 || k == RODFT11))
        return false;
    if (n % 2 || h < 16 || h > 1024 || (h & (h - 1))) return false;
    const int d = P.d;
    return P.src_s[d] == 1 && P.dst_s[d] == 1 && P.src_nf == 1 && P.dst_nf == 1 && P.nf_in == 1 && P.nf_out == 1 &&
           P.r2r_stride == 1 && P.si == 0 && P.so == 0 && P.gshift == 0 && P.sshift == 0 && P.src_nd == n &&
           P.dst_nd == n && P.pen_in == n && P.pen_out == n && P.post == 0;
}

// pre / post twiddles of the row r2r kernels, (H + 1) complex each, exact to rounding
std::vector<double> row_r2r_table(int kind, int n) {
    const int dct = kind >= RODFT10 ? kind - 4 : kind;
    const int h = n / 2;
    const long double pi = 3.14159265358979323846264338327950288L, N = n;
    std::vector<double> t(4 * (size_t)(h + 1));
    auto put = [&](size_t i, long double angle) {
        t[2 * i] = (double)cosl(angle);
        t[2 * i + 1] = (double)sinl(angle);
    };
    for (int k = 0; k <= h; ++k) {
        if (dct == REDFT11) {
            put(k, -pi * (k + 0.25L) / N);
            put(h + 1 + k, -pi * k / N);
        } else if (dct == REDFT10) {
            put(k, -2.0L * pi * k / N);
            put(h + 1 + k, -pi * k / (2.0L * N));
        } else {
            put(k, pi * k / (2.0L * N));
            put(h + 1 + k, 2.0L * pi * k / N);
        }
    }
    return t;
}

void launch_row_r2r(const PassDesc& P, hipStream_t s) {
    switch (P.n / 2) {
        case 16: launch_row_h<16, 128>(P, s); break;
        case 32: launch_row_h<32, 64>(P, s); break;
        case 64: launch_row_h<64, 32>(P, s); break;
        case 128: launch_row_h<128, 16>(P, s); break;
        case 256: launch_row_h<256, 8>(P, s); break;
        case 512: launch_row_h<512, 4>(P, s); break;
        case 1024: launch_row_h<1024, 2>(P, s); break;
        default: fail("row r2r: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

}  // namespace flups
