// Kernels and launch templates of the general path's fast pass kernels (rowpass.hip), for
// storage types TI (pass input) and TO (pass output): double for the caller's arrays and the
// stage arrays of FP64 solves, float for the stage arrays of FP32 solves. Arithmetic, twiddle
// and pre / post tables are FP64 throughout; values are converted on load and rounded once on
// store. Each instantiation of the launch templates lives in one translation unit (rowpass.hip:
// double -> double; rowpass_df.hip, rowpass_ff.hip, rowpass_fd.hip: the FP32 combinations).
#pragma once
// The kernels: fast kernels of the general path for its most common pencil shapes, used in
// place of the generic pass kernel (generic.hip) when a pass fits them.
//
// Lengths: a power of two, or 3, 5, 7 times a power of two (mixed radix, see
// column_fft_pq and fft_pq_team in fft_core.hpp); the ranges are in the fit tests.
//
// row_r2r (pen_r2r, rows): real-to-real transforms of kinds REDFT10/01/11 and RODFT10/01/11
// (src/FFTW_plan_dim.cpp:635-720 runs them through FFTW) along a contiguous
// pencil of even length N = 2H, H in [16, 1024]. Each is one
// complex FFT of length H held in registers (the column FFT of fft_core.hpp)
// with pre- and post-twiddles, instead of the 2N-point extension of the
// generic kernel:
//   REDFT11 (DCT-IV): v[m] = (x[2m] + i x[N-1-2m]) e^{-i pi (m + 1/4)/N}, V = FFT_H(v),
//                     W[k] = V[k] e^{-i pi k/N}, y[2k] = 2 Re W[k], y[N-1-2k] = -2 Im W[k];
//   REDFT10 (DCT-II): u[m] = x[2m], u[N-1-m] = x[2m+1] (m < H), U = FFT_N(u) from the
//                     packed FFT_H of u[2m] + i u[2m+1], y[k] = 2 Re(e^{-i pi k/(2N)} U[k]);
//   REDFT01 (DCT-III): U[k] = e^{i pi k/(2N)} (x[k] - i x[N-k]) (x[N] = 0), u = N IFFT_N(U)
//                     through the packed inverse FFT_H, y[2m] = u[m], y[2m+1] = u[N-1-m];
//   RODFT10 = reversed REDFT10 of (-1)^n x, RODFT01/11 = (-1)^k REDFT01/11 of reversed x;
//   REDFT00 (DCT-I, n = H + 1) and RODFT00 (DST-I, n = H - 1): the even / odd extension of
//                     2H points through the packed real FFT of length H, y = Re X[k] resp. -Im X[k+1].
// ROWS pencils per block, H/8 threads per pencil; the pencil is staged through
// LDS so that HBM is read and written in 16-byte pieces when rows are aligned.
// col_r2r runs the same transforms along a strided direction: R pencils that are
// neighbours along the unit-stride dimension share a block, interleaved in LDS,
// and each row of the tile is one contiguous run of R doubles in HBM.
// The pencil switch and plan semantics of the generic kernel are kept: gather
// with the field shift and zero padding, input / output offsets of 0 or 1,
// the first / last point zeroed after the transform, truncated scatter.
//
// col_real: r2c / c2r of length n (in [32, 1024]) along a strided
// direction whose pencils are neighbours along a unit-stride dimension (the
// y or z passes of the general path, x fastest). Two neighbouring real pencils
// a, b are one complex column a + i b: Z = FFT_n(a + i b), A[k] = (Z[k] + conj
// Z[n-k]) / 2, B[k] = (Z[k] - conj Z[n-k]) / (2i) for k <= n/2; the inverse
// builds Z[k] = A[k] + i B[k] (conj A[n-k] + i conj B[n-k] above n/2; the
// imaginary parts of the DC and Nyquist terms are ignored, as FFTW's c2r does)
// and a + i b = IFFT_n(Z). TX pairs side by side per block: each row of a
// tile is one contiguous run of 2 TX reals or complex values.
//
// col_c2c: complex-to-complex transforms of length M (in [32, 2048]) along a
// strided direction, TX neighbouring pencils per block, with the zero padding of the gather
// and the truncation of the scatter (the second and third transforms of unbounded and
// periodic directions in the general path). col_c2c_team: the same for M = 5 or 7 times a
// power of two, the columns staged in LDS first.
//
// row_c2c: c2c of length M (in [16, 1024]) along a contiguous pencil (x after an r2c
// direction), staged in LDS, PF teams of M/(8 PF) threads.
//
// col_rz: the general path's last direction when it is an r2c one, with the multiply: r2c,
// normfact G, c2r of paired real columns in one in-place launch; row_rz: the same along x.
//
// row_rc: r2c / c2r of even length n = 2H (H in [16, 1024]) along a
// contiguous pencil through the packed complex FFT of length H: z[m] = x[2m] + i x[2m+1],
// X[k] = (Z[k] + conj Z[H-k])/2 + e^{-2 pi i k/n} (Z[k] - conj Z[H-k])/(2i), k <= H; the
// inverse Z[k] = (X[k] + conj X[H-k]) + i e^{2 pi i k/n} (X[k] - conj X[H-k]), x[2m] + i x[2m+1]
// = IFFT_H(Z) (the imaginary parts of X[0] and X[H] ignored).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "core.hpp"
#include "fft_core.hpp"
#include "group.hpp"

namespace flups {

int r2r_h(int kind, int n);
bool split_pq(int n, int& pf, int& q);

namespace {


// COL = false: R contiguous pencils per block (H/8 threads each, one padded LDS row each);
// COL = true: R strided pencils, neighbours along the unit-stride dimension fa, interleaved
// in LDS (element k of pencil c at slot k * R + c) and read / written as tile rows.
// H = PF Q: PF = 1 (a power of two) or 3, 5, 7 (PF teams of Q/8 threads, fft_pq_team);
// H/8 threads per pencil either way
template <int PF, int Q, int R, int KIND, bool COL, typename TI, typename TO>
struct pen_r2r {
    using Args = PassDesc;
    static constexpr int NT = R * PF * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int H = PF * Q, T = Q / 8, TT = H / 8, N = 2 * H, HP = H + H / 8;
        constexpr int S = COL ? R : 1;   // LDS stride between consecutive elements of a pencil
        __shared__ C2<double> lds[COL ? (H + 1) * R : HP * R];
        constexpr int DCT = KIND >= RODFT10 ? KIND - 4 : KIND;
        // DCT-I / DST-I (node-centred symmetric directions): the even / odd extension of length 2H
        // through the packed real FFT; n = H + 1 (DCT-I) or H - 1 (DST-I) points in and out
        constexpr bool E00 = KIND == REDFT00 || KIND == RODFT00;
        constexpr int NT = KIND == REDFT00 ? H + 1 : KIND == RODFT00 ? H - 1 : N;   // transform outputs
        constexpr bool ALT = KIND == RODFT10;                       // input (-1)^n x, output reversed
        constexpr bool REV = KIND == RODFT01 || KIND == RODFT11;    // input reversed, output (-1)^k
        int rl, tt, ja, jb;
        bool live;
        if (COL) {
            rl = threadIdx.x % R;
            tt = threadIdx.x / R;
            const int ntile = (P.na + R - 1) / R;
            const int tile = blk.x % ntile;
            jb = blk.x / ntile;
            ja = tile * R + rl;
            live = ja < P.na;
            if (!live) ja = P.na - 1;   // dead pencils read the last one and store nothing
        } else {
            rl = threadIdx.x / TT;
            tt = threadIdx.x % TT;
            const int g = blk.x * R + rl;
            live = g < P.n1 * P.n2;
            const int gg = live ? g : 0;   // dead rows read pencil 0 and store nothing
            ja = gg % P.na;
            jb = gg / P.na;
        }
        const int k1 = tt / T, j = tt % T;   // team and thread in the team
        const TI* src = reinterpret_cast<const TI*>(P.src) + (long long)ja * P.src_s[P.fa] + (long long)jb * P.src_s[P.fb];
        TO* dst = reinterpret_cast<TO*>(P.dst) + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
        const long long sd = P.src_s[P.d], dd = P.dst_s[P.d];
        C2<double>* row = COL ? lds + rl : lds + rl * HP;                                    // complex view
        double* rdb = COL ? reinterpret_cast<double*>(lds) + rl : reinterpret_cast<double*>(row);   // real view
        auto RD = [&](int k) -> double& { return rdb[k * S]; };
        auto RW = [&](int k) -> C2<double>& { return row[k * S]; };
        const C2<double>* A = reinterpret_cast<const C2<double>*>(P.aux);
        const C2<double>* B = A + (H + 1);
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(P.tw);   // Q-point table
        const C2<double>* twh = B + (H + 1);                                 // H-point table (PF > 1)
        const int pin = P.pen_in;   // pencil buffer length (the same after the transform), at most N + 2

        // 1. gather the pencil into LDS (field shift, zero padding)
        if (!COL && P.vec) {
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int q = tt + t * TT;
                const vec2_t<TI>* a = reinterpret_cast<const vec2_t<TI>*>(src) + q;
                const vec2_t<TI> p = P.nt ? ld_nt2(a) : *a;
                RD(2 * q) = (double)p.x;
                RD(2 * q + 1) = (double)p.y;
            }
        } else {
            const int last = max(P.src_nd - 1, 0);
            staged_gather<8>(tt, TT, pin, [&](int k) {
                const int ks = k - P.gshift;
                const double x = (double)src[(long long)min(max(ks, 0), last) * sd];
                return (ks >= 0 && ks < P.src_nd) ? x : 0.0;
            }, [&](int k, double x) { RD(k) = x; });
        }
        __syncthreads();
        const int si = P.si;
        auto xin = [&](int q) -> double {
            const double v = RD(si + (REV ? N - 1 - q : q));
            return (ALT && (q & 1)) ? -v : v;
        };

        // 2. pre-map to the complex sequence of length H: element m
        auto z01 = [&](int m) -> C2<double> {   // REDFT01
            auto U = [&](int k) {
                const double a = k < N ? xin(k) : 0.0, b = (N - k) < N ? xin(N - k) : 0.0;
                return cmul(A[k], C2<double>{a, -b});
            };
            const C2<double> uk = U(m), uc = cconj(U(H - m));
            const C2<double> e = cadd(uk, uc), f = csub(uk, uc);
            const C2<double> t = cmul(B[m], f);
            return {e.x - t.y, e.y + t.x};   // e + i t
        };
        // mixed radix (PF > 1) with one thread per team (Q = 8) reads every input PF times in one
        // unrolled loop: the REDFT01 sequence is built in LDS first (all of it read into registers,
        // then written over the pencil), so that the transform's loads are plain LDS reads;
        // inlined there, the construction held 256 registers and more live (one wave per SIMD).
        // With larger teams the inlined form needs no more registers than the staged one.
        constexpr bool PRE = DCT == REDFT01 && PF > 1 && Q == 8;
        if constexpr (PRE) {
            C2<double> zt[8];
    #pragma unroll
            for (int i = 0; i < 8; ++i) zt[i] = z01(tt + i * TT);
            __syncthreads();
    #pragma unroll
            for (int i = 0; i < 8; ++i) RW(tt + i * TT) = zt[i];
            __syncthreads();
        }
        auto load = [&](int m) -> C2<double> {
            if (PRE) return RW(m);
            if (KIND == REDFT00) {   // e[q] = x[q] (q <= H), x[2H - q] (q > H)
                auto e = [&](int q) { return xin(q <= H ? q : N - q); };
                return {e(2 * m), e(2 * m + 1)};
            } else if (KIND == RODFT00) {   // o[0] = o[H] = 0, o[q] = x[q-1] (q < H), -x[2H-q-1] (q > H)
                auto o = [&](int q) { return (q == 0 || q == H) ? 0.0 : q < H ? xin(q - 1) : -xin(N - q - 1); };
                return {o(2 * m), o(2 * m + 1)};
            } else if (DCT == REDFT11) {
                return cmul(C2<double>{xin(2 * m), xin(N - 1 - 2 * m)}, A[m]);
            } else if (DCT == REDFT10) {
                auto u = [&](int q) { return q < H ? xin(2 * q) : xin(2 * (N - 1 - q) + 1); };
                return {u(2 * m), u(2 * m + 1)};
            } else {   // REDFT01
                return z01(m);
            }
        };

        // 3. FFT of length H (forward for DCT-II / IV, inverse for DCT-III): v[i] = Z[k1 + PF (j + i T)]
        C2<double> v[8];
        fft_pq_team<double, PF, Q, DCT != REDFT01, COL ? R : 0>(v, row, k1, j, load, tw, twh, 1);
        __syncthreads();   // every thread is done with the FFT's LDS before the buffer is reused

        // 4. post-map into the LDS pencil
        if (E00) {
            // X[k] = (Z[k] + conj Z[H-k])/2 + e^{-i pi k/H} (Z[k] - conj Z[H-k])/(2i), the real DFT of the
            // extension; DCT-I y[k] = Re X[k] (k <= H), DST-I y[k] = -Im X[k+1] (k < H - 1)
    #pragma unroll
            for (int i = 0; i < 8; ++i) RW(k1 + PF * (j + i * T)) = v[i];
            __syncthreads();
            double XH = 0.0;
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = tt + i * TT;
                const C2<double> zk = RW(k), zc = cconj(RW(k == 0 ? 0 : H - k));
                const C2<double> e = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
                const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
                v[i] = cadd(e, cmul(A[k], C2<double>{df.y, -df.x}));
            }
            if (tt == 0) {
                const C2<double> z0 = RW(0);
                XH = z0.x - z0.y;
            }
            __syncthreads();
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = tt + i * TT;
                if (KIND == REDFT00) RD(k) = v[i].x;
                else if (k >= 1) RD(k - 1) = -v[i].y;
            }
            if (KIND == REDFT00 && tt == 0) RD(H) = XH;
        } else if (DCT == REDFT11) {
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = k1 + PF * (j + i * T);
                const C2<double> w = cmul(v[i], B[k]);
                RD(2 * k) = 2.0 * w.x;
                RD(N - 1 - 2 * k) = -2.0 * w.y;
            }
        } else if (DCT == REDFT10) {
    #pragma unroll
            for (int i = 0; i < 8; ++i) RW(k1 + PF * (j + i * T)) = v[i];
            __syncthreads();
            C2<double> U[8], UH = {0, 0};
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = tt + i * TT;
                const C2<double> zk = RW(k), zc = cconj(RW(k == 0 ? 0 : H - k));
                const C2<double> e = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
                const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
                U[i] = cadd(e, cmul(A[k], C2<double>{df.y, -df.x}));   // e + A[k] (zk - zc) / (2i)
            }
            if (tt == 0) {   // U[H] = Re Z[0] - Im Z[0]
                const C2<double> z0 = RW(0);
                UH = {z0.x - z0.y, 0.0};
            }
            __syncthreads();
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = tt + i * TT;
                const C2<double> c = cmul(B[k], U[i]);
                RD(k) = 2.0 * c.x;
                if (k > 0) RD(N - k) = -2.0 * c.y;
            }
            if (tt == 0) RD(H) = 2.0 * cmul(B[H], UH).x;
        } else {   // REDFT01: v[i] = z[m], u[2m] = Re, u[2m+1] = Im
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int m = k1 + PF * (j + i * T);
                const int q0 = 2 * m, q1 = 2 * m + 1;
                RD(q0 < H ? 2 * q0 : 2 * (N - 1 - q0) + 1) = v[i].x;
                RD(q1 < H ? 2 * q1 : 2 * (N - 1 - q1) + 1) = v[i].y;
            }
        }
        __syncthreads();

        // 5. scatter: output pencil position p holds transform value p - so (DST reflections
        // applied), the first / last point zeroed by the plan's post-correction
        if (!live) return;
        const int so = P.so, post = P.post;
        auto val = [&](int p) -> double {
            if (((post & POST_FIRST) && p == 0) || ((post & POST_LAST) && p == pin - 1)) return 0.0;
            const int k = p - so;
            if (k < 0 || k >= NT) return 0.0;
            const double y = RD(ALT ? N - 1 - k : k);
            return (REV && (k & 1)) ? -y : y;
        };
        if (!COL && P.vec) {
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int q = tt + t * TT;
                const vec2_t<TO> o{(TO)val(2 * q), (TO)val(2 * q + 1)};
                if (P.nt) st_nt2(reinterpret_cast<vec2_t<TO>*>(dst) + q, o);
                else reinterpret_cast<vec2_t<TO>*>(dst)[q] = o;
            }
        } else {
            for (int k = tt; k < P.dst_nd; k += TT) {
                const int p = k + P.sshift;
                if (p < pin) dst[(long long)k * dd] = (TO)val(p);
            }
        }
    }
};

// H = PF Q: PF = 1 (a power of two) or 3, 5, 7 (PF teams of Q/8 threads, fft_pq_team);
// H/8 threads per row either way
template <int PF, int Q, int ROWS, bool FWD, typename TI, typename TO>
struct row_rc {
    using Args = PassDesc;
    static constexpr int NT = ROWS * PF * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int H = PF * Q, T = Q / 8, TT = H / 8, HP = H + H / 8;
        __shared__ C2<double> lds[HP * ROWS];
        const int rl = threadIdx.x / TT, tt = threadIdx.x % TT;
        const int k1 = tt / T, j = tt % T;   // team and thread in the team
        const int g = blk.x * ROWS + rl;
        const bool live = g < P.n1 * P.n2;
        const int gg = live ? g : 0;   // dead rows read pencil 0 and store nothing
        const int ja = gg % P.na, jb = gg / P.na;
        const TI* src = reinterpret_cast<const TI*>(P.src) + (long long)ja * P.src_s[P.fa] * P.src_nf + (long long)jb * P.src_s[P.fb] * P.src_nf;
        TO* dst = reinterpret_cast<TO*>(P.dst) + (long long)ja * P.dst_s[P.fa] * P.dst_nf + (long long)jb * P.dst_s[P.fb] * P.dst_nf;
        C2<double>* row = lds + rl * HP;
        double* rd = reinterpret_cast<double*>(row);
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(P.tw);    // Q-point table
        const C2<double>* twn = reinterpret_cast<const C2<double>*>(P.aux);  // n-point table
        // 1. gather: the pencil's doubles (reals for r2c, interleaved complex for c2r)
        const int nfi = P.nf_in, len = P.pen_in * nfi;
        const int last = max(P.src_nd - 1, 0);
        staged_gather<8>(tt, TT, len, [&](int q) {
            const int k = q / nfi, ks = k - P.gshift;
            const double x = (double)src[(long long)min(max(ks, 0), last) * nfi + (q - k * nfi)];
            return (ks >= 0 && ks < P.src_nd) ? x : 0.0;
        }, [&](int q, double x) { rd[q] = x; });
        __syncthreads();
        // element m of the packed complex sequence: z[m] = x[2m] + i x[2m+1] (r2c), or built from
        // X[m] and X[H-m] (c2r)
        auto c2r_z = [&](int m) -> C2<double> {
            C2<double> a = {rd[2 * m], rd[2 * m + 1]}, b = {rd[2 * (H - m)], rd[2 * (H - m) + 1]};
            if (m == 0) {   // DC and Nyquist are real
                a.y = 0.0;
                b.y = 0.0;
            }
            const C2<double> bc = cconj(b);
            const C2<double> e = cadd(a, bc), f = csub(a, bc);
            const C2<double> t = cmul(cconj(twn[m]), f);   // e^{+2 pi i m/n} (a - conj b)
            return {e.x - t.y, e.y + t.x};
        };
        // mixed radix (PF > 1) reads every input PF times: build z in the row first, the pairs
        // (m, H - m) by one thread, so that the transform's loads are plain LDS reads (inlined into
        // its PF-fold unrolled loads, the construction held several hundred registers live)
        constexpr bool PRE = !FWD && PF > 1;
        if constexpr (PRE) {
            for (int m = tt; m <= H / 2; m += TT) {
                const C2<double> zm = c2r_z(m);
                const C2<double> zh = (m > 0 && 2 * m != H) ? c2r_z(H - m) : C2<double>{0, 0};
                row[m] = zm;
                if (m > 0 && 2 * m != H) row[H - m] = zh;
            }
            __syncthreads();
        }
        auto load = [&](int m) -> C2<double> {
            if (FWD || PRE) return {rd[2 * m], rd[2 * m + 1]};
            return c2r_z(m);
        };
        C2<double> v[8];
        fft_pq_team<double, PF, Q, FWD, 0>(v, row, k1, j, load, tw, twn, 2);   // w_H^e = twn[2 e]
        __syncthreads();
        const int nfo = P.nf_out, plen = P.pen_out;
        // v[i] = Z[k1 + PF (j + i T)]
    #pragma unroll
        for (int i = 0; i < 8; ++i) row[k1 + PF * (j + i * T)] = v[i];
        if (FWD) {
            // X[k], k <= H, into the row as complex values
            __syncthreads();
            C2<double> XH = {0, 0};
    #pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = tt + i * TT;
                const C2<double> zk = row[k], zc = cconj(row[k == 0 ? 0 : H - k]);
                const C2<double> e = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
                const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
                v[i] = cadd(e, cmul(twn[k], C2<double>{df.y, -df.x}));
            }
            if (tt == 0) {
                const C2<double> z0 = row[0];
                XH = {z0.x - z0.y, 0.0};
            }
            __syncthreads();
    #pragma unroll
            for (int i = 0; i < 8; ++i) row[tt + i * TT] = v[i];
            if (tt == 0) row[H] = XH;
            // positions past H + 1 complex values keep nothing (zero)
            for (int q = H + 1 + tt; q < plen; q += TT) row[q] = {0.0, 0.0};
            // periodic node-centred direction: the first value again at the end (n_out - 1)
            if ((P.post & POST_PERIOD) && tt == 0) row[P.n_out_plan - 1] = v[0];
        } else {
            // x[2m] = Re z[m], x[2m+1] = Im z[m] (written above)
            for (int q = 2 * H + tt; q < plen; q += TT) rd[q] = 0.0;
            if ((P.post & POST_PERIOD) && tt == 0) rd[2 * H] = v[0].x;   // x[n] = x[0]
        }
        __syncthreads();
        if (!live) return;
        // 2. scatter (truncation), element k of the destination = pencil position k + sshift
        const int dn = P.dst_nd;
        for (int q = tt; q < dn * nfo; q += TT) {
            const int k = q / nfo, p = k + P.sshift;
            if (p < plen) dst[(long long)k * nfo + (q - k * nfo)] = (TO)rd[p * nfo + (q - k * nfo)];
        }
    }
};

// M = PF Q: PF = 1 (a power of two) or 3, 5 (8 PF values per thread, column_fft_pq)
template <int PF, int Q, int TX, bool FWD, typename TI, typename TO>
struct col_real {
    using Args = PassDesc;
    static constexpr int NT = TX * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int M = PF * Q, T = Q / 8, NH = M / 2, NV = 8 * PF;
        __shared__ C2<double> lds[(M + 2) * TX];
        const int c = threadIdx.x % TX, j = threadIdx.x / TX;
        const int npair = (P.na + 1) / 2;
        const int ntile = (npair + TX - 1) / TX;
        const int tile = blk.x % ntile, jb = blk.x / ntile;
        const int pr = tile * TX + c;
        const bool la = 2 * pr < P.na, lb = 2 * pr + 1 < P.na;
        const int xa = la ? 2 * pr : 2 * (npair - 1);   // dead pairs read the last pair and store nothing
        const int d = P.d;
        const long long sr = FWD ? P.src_s[d] : P.dst_s[d];       // row stride of the real array
        const long long sc = FWD ? P.dst_s[d] : P.src_s[d];       // row stride of the complex array
        const TI* rsrc = reinterpret_cast<const TI*>(P.src) + (long long)jb * P.src_s[P.fb] + xa;          // FWD: real source
        TO* rdst = reinterpret_cast<TO*>(P.dst) + (long long)jb * P.dst_s[P.fb] + xa;                    // !FWD: real destination
        const TI* csrc = reinterpret_cast<const TI*>(P.src) + 2 * ((long long)jb * P.src_s[P.fb] + xa);   // !FWD: complex source
        TO* cdst = reinterpret_cast<TO*>(P.dst) + 2 * ((long long)jb * P.dst_s[P.fb] + xa);             // FWD: complex destination
        const bool hasb = lb;   // a real partner column exists
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(P.tw);    // Q-point table
        const C2<double>* twm = reinterpret_cast<const C2<double>*>(P.aux);  // M-point table (PF > 1)
        C2<double> v[NV];
        if (FWD) {
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                const TI* p = rsrc + (long long)row * sr;
                if (row >= P.src_nd) {   // zero padding of an unbounded direction
                    v[i] = {0.0, 0.0};
                } else if (P.vec) {
                    const vec2_t<TI> q = *reinterpret_cast<const vec2_t<TI>*>(p);
                    v[i] = {(double)q.x, (double)q.y};
                } else {
                    v[i] = {(double)p[0], hasb ? (double)p[1] : 0.0};
                }
            }
            column_fft_pq<double, PF, Q, true, TX>(v, lds + c, j, tw, twm, 1);
            __syncthreads();   // the FFT's last LDS reads are done
    #pragma unroll
            for (int i = 0; i < NV; ++i) lds[(j + i * T) * TX + c] = v[i];
            __syncthreads();
            if (!la) return;
    #pragma unroll
            for (int i = 0; i <= NV / 2; ++i) {
                const int k = j + i * T;
                if (i == NV / 2 && j != 0) break;   // k = n/2 once
                const C2<double> zk = lds[k * TX + c], zc = cconj(lds[(k == 0 ? 0 : M - k) * TX + c]);
                const C2<double> A = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
                const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
                TO* o = cdst + 2 * (long long)k * sc;
                *reinterpret_cast<vec2_t<TO>*>(o) = vec2_t<TO>{(TO)A.x, (TO)A.y};
                if (hasb) *reinterpret_cast<vec2_t<TO>*>(o + 2) = vec2_t<TO>{(TO)df.y, (TO)-df.x};   // (zk - zc) / (2i)
                if ((P.post & POST_PERIOD) && k == 0) {   // node-centred periodic: X[0] again at n/2 + 1
                    TO* o2 = cdst + 2 * (long long)(NH + 1) * sc;
                    *reinterpret_cast<vec2_t<TO>*>(o2) = vec2_t<TO>{(TO)A.x, (TO)A.y};
                    if (hasb) *reinterpret_cast<vec2_t<TO>*>(o2 + 2) = vec2_t<TO>{(TO)df.y, (TO)-df.x};
                }
            }
        } else {
            // rows 0..n/2 of the two spectra into LDS: A at row k, B at row n/2 + 1 + k
    #pragma unroll
            for (int i = 0; i <= NV / 2; ++i) {
                const int k = j + i * T;
                if (i == NV / 2 && j != 0) break;
                const TI* p = csrc + 2 * (long long)k * sc;
                const vec2_t<TI> qa = *reinterpret_cast<const vec2_t<TI>*>(p);
                const vec2_t<TI> qb = hasb ? *reinterpret_cast<const vec2_t<TI>*>(p + 2) : vec2_t<TI>{0, 0};
                C2<double> A = {(double)qa.x, (double)qa.y}, B = {(double)qb.x, (double)qb.y};
                if (k == 0 || k == NH) {   // DC and Nyquist are real
                    A.y = 0.0;
                    B.y = 0.0;
                }
                lds[k * TX + c] = A;
                lds[(NH + 1 + k) * TX + c] = B;
            }
            __syncthreads();
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int k = j + i * T;
                const int m = k <= NH ? k : M - k;
                const C2<double> A = lds[m * TX + c], B = lds[(NH + 1 + m) * TX + c];
                v[i] = k <= NH ? C2<double>{A.x - B.y, A.y + B.x} : C2<double>{A.x + B.y, B.x - A.y};
            }
            column_fft_pq<double, PF, Q, false, TX>(v, lds + c, j, tw, twm, 1);   // syncs before its first LDS write
            if (!la) return;
    #pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int row = j + i * T;
                if (row >= P.dst_nd) continue;   // truncation back to the field's extent
                TO* p = rdst + (long long)row * sr;
                if (P.vec) {
                    *reinterpret_cast<vec2_t<TO>*>(p) = vec2_t<TO>{(TO)v[i].x, (TO)v[i].y};
                } else {
                    p[0] = (TO)v[i].x;
                    if (hasb) p[1] = (TO)v[i].y;
                }
            }
            if ((P.post & POST_PERIOD) && j == 0 && M < P.dst_nd) {   // node-centred periodic: x[n] = x[0]
                TO* p = rdst + (long long)M * sr;
                p[0] = (TO)v[0].x;
                if (hasb) p[1] = (TO)v[0].y;
            }
        }
    }
};

// The general path's last direction when it is an r2c one (symmetric x and y, unbounded or
// periodic z): forward r2c, normfact G, inverse c2r in one launch, in place. Two neighbouring
// real columns a, b are one complex column a + i b (as in col_real): with Z = FFT(a + i b) and
// Z' the product spectrum, A' = Ga (Z[k] + conj Z[M-k])/2 and B' = Gb (Z[k] - conj Z[M-k])/(2i)
// recombine to Z'[k] = A' + i B' = (Ga + Gb)/2 Z[k] + (Ga - Gb)/2 conj Z[M-k] (G real and even
// in kz), and IFFT(Z') = a' + i b'. Rows from nin on are zero (padding) and are not stored.
template <int M, int TX, typename S>
struct col_rz {
    using Args = RzArgs;
    static constexpr int NT = TX * M / 8;
    static __device__ __forceinline__ void run(const RzArgs& A, Blk blk) {
        constexpr int T = M / 8;
        __shared__ C2<double> lds[M * TX];
        const int c = threadIdx.x % TX, j = threadIdx.x / TX;
        const int npair = (A.nx + 1) / 2;
        const int ntile = (npair + TX - 1) / TX;
        const int tile = blk.x % ntile, y = blk.x / ntile;
        const int pr = tile * TX + c;
        const bool la = 2 * pr < A.nx, lb = 2 * pr + 1 < A.nx;
        const int xa = la ? 2 * pr : 2 * (npair - 1);   // dead pairs read the last pair and store nothing
        const bool hb = 2 * pr + 1 < A.nx || (!la && xa + 1 < A.nx);   // partner column present (for the loads)
        S* base = reinterpret_cast<S*>(A.d) + (long long)blk.y * A.comp + (long long)y * A.nx + xa;
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(A.tw);
        C2<double> v[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = j + i * T;
            const S* p = base + (long long)row * A.s_col;
            v[i] = row < A.nin ? C2<double>{(double)p[0], hb ? (double)p[1] : 0.0} : C2<double>{0.0, 0.0};
        }
        column_fft<double, M, true, TX>(v, lds + c, j, tw);
        __syncthreads();   // the FFT's last LDS reads are done
    #pragma unroll
        for (int i = 0; i < 8; ++i) lds[(j + i * T) * TX + c] = v[i];
        __syncthreads();
        const S* g = reinterpret_cast<const S*>(A.g) + 2 * ((long long)y * A.nx + xa);
        const long long gs = 2LL * A.nx * A.ny;   // doubles between kz rows of G
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = j + i * T;
            const C2<double> zc = cconj(lds[(k == 0 ? 0 : M - k) * TX + c]);
            const int kz = k <= M / 2 ? k : M - k;
            const double ga = (double)g[kz * gs], gb = hb ? (double)g[kz * gs + 2] : ga;
            const double p = 0.5 * (ga + gb) * A.normfact, q = 0.5 * (ga - gb) * A.normfact;
            v[i] = {p * v[i].x + q * zc.x, p * v[i].y + q * zc.y};
        }
        column_fft<double, M, false, TX>(v, lds + c, j, tw);   // syncs before its first LDS write
        if (!la) return;
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = j + i * T;
            if (row >= A.nin) continue;
            S* p = base + (long long)row * A.s_col;
            p[0] = (S)v[i].x;
            if (lb) p[1] = (S)v[i].y;
        }
    }
};

// The general path's last direction when it is an r2c along x (every other direction real-
// to-real): r2c of each row (length n = 2H, rows zero-padded from nin), X[k] normfact G[k]
// (G real, one row of the spectral layout per field row), c2r, the first nin values stored,
// all in LDS, in place. PF teams of Q/8 threads as row_rc.
template <int PF, int Q, int ROWS, typename S>
struct row_rz {
    using Args = RzArgs;
    static constexpr int NT = ROWS * PF * Q / 8;
    static __device__ __forceinline__ void run(const RzArgs& A, Blk blk) {
        constexpr int H = PF * Q, T = Q / 8, TT = H / 8, HP = H + H / 8;
        __shared__ C2<double> lds[HP * ROWS];
        const int rl = threadIdx.x / TT, tt = threadIdx.x % TT;
        const int k1 = tt / T, j = tt % T;
        const int g = blk.x * ROWS + rl;
        const bool live = g < A.rows;
        const int r = live ? g : 0;   // dead rows read row 0 and store nothing
        S* d = reinterpret_cast<S*>(A.d) + (long long)blk.y * A.comp + (long long)r * A.nin;
        const S* gr = reinterpret_cast<const S*>(A.g) + 2LL * r * (H + 1);   // real parts at even offsets
        C2<double>* row = lds + rl * HP;
        double* rd = reinterpret_cast<double*>(row);
        const C2<double>* tw = reinterpret_cast<const C2<double>*>(A.tw);    // Q-point table
        const C2<double>* twn = reinterpret_cast<const C2<double>*>(A.twn);  // n-point table
        staged_gather<8>(tt, TT, 2 * H, [&](int q) {
            const double x = (double)d[min(q, A.nin - 1)];
            return q < A.nin ? x : 0.0;
        }, [&](int q, double x) { rd[q] = x; });
        __syncthreads();
        C2<double> v[8];
        auto loadf = [&](int m) -> C2<double> { return {rd[2 * m], rd[2 * m + 1]}; };
        fft_pq_team<double, PF, Q, true, 0>(v, row, k1, j, loadf, tw, twn, 2);
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) row[k1 + PF * (j + i * T)] = v[i];
        __syncthreads();
        // X[k] = (Z[k] + conj Z[H-k])/2 + e^{-2 pi i k/n} (Z[k] - conj Z[H-k])/(2i), times normfact G[k]
        C2<double> XH = {0, 0};
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = tt + i * TT;
            const C2<double> zk = row[k], zc = cconj(row[k == 0 ? 0 : H - k]);
            const C2<double> e = {0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y)};
            const C2<double> df = {0.5 * (zk.x - zc.x), 0.5 * (zk.y - zc.y)};
            const C2<double> X = cadd(e, cmul(twn[k], C2<double>{df.y, -df.x}));
            const double gk = A.normfact * (double)gr[2 * k];
            v[i] = {X.x * gk, X.y * gk};
        }
        if (tt == 0) {
            const C2<double> z0 = row[0];
            XH = {(z0.x - z0.y) * A.normfact * (double)gr[2 * H], 0.0};
        }
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) row[tt + i * TT] = v[i];
        if (tt == 0) row[H] = XH;
        __syncthreads();
        // c2r: Z[m] = (X[m] + conj X[H-m]) + i e^{2 pi i m/n} (X[m] - conj X[H-m]), z = IFFT_H(Z)
        auto c2r_z = [&](int m) -> C2<double> {
            C2<double> a = {rd[2 * m], rd[2 * m + 1]}, b = {rd[2 * (H - m)], rd[2 * (H - m) + 1]};
            if (m == 0) {   // DC and Nyquist are real
                a.y = 0.0;
                b.y = 0.0;
            }
            const C2<double> bc = cconj(b);
            const C2<double> e = cadd(a, bc), f = csub(a, bc);
            const C2<double> t = cmul(cconj(twn[m]), f);
            return {e.x - t.y, e.y + t.x};
        };
        // mixed radix: Z built in the row first (as in row_rc's c2r), the transform reads it plainly
        constexpr bool PRE = PF > 1;
        if constexpr (PRE) {
            for (int m = tt; m <= H / 2; m += TT) {
                const C2<double> zm = c2r_z(m);
                const C2<double> zh = (m > 0 && 2 * m != H) ? c2r_z(H - m) : C2<double>{0, 0};
                row[m] = zm;
                if (m > 0 && 2 * m != H) row[H - m] = zh;
            }
            __syncthreads();
        }
        auto loadb = [&](int m) -> C2<double> {
            if (PRE) return {rd[2 * m], rd[2 * m + 1]};
            return c2r_z(m);
        };
        fft_pq_team<double, PF, Q, false, 0>(v, row, k1, j, loadb, tw, twn, 2);
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) row[k1 + PF * (j + i * T)] = v[i];   // x[2m] = Re z[m], x[2m+1] = Im z[m]
        __syncthreads();
        if (!live) return;
        for (int q = tt; q < A.nin; q += TT) d[q] = (S)rd[q];
    }
};

// Strided complex pencils, TX neighbours along the unit-stride dimension fa per block,
// interleaved in LDS; the gather zero-pads (rows below gshift or from src_nd on read as
// zero) and the scatter truncates (pencil position p lands in row p - sshift when that
// row is below dst_nd). HIN: rows M/2 and up are padding (gshift 0, src_nd <= M/2), the
// first stage skips them; HOUT: only rows below M/2 are stored (sshift 0, dst_nd <= M/2).
// M = PF Q: PF = 1 (a power of two) or 3 (mixed radix, 8 PF values per thread, column_fft_pq)
template <int PF, int Q, int TX, bool FWD, bool HIN, bool HOUT, typename TI, typename TO>
struct col_c2c {
    using Args = PassDesc;
    static constexpr int NT = TX * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int M = PF * Q, T = Q / 8, NV = 8 * PF;
        __shared__ C2<double> lds[M * TX];
        const int c = threadIdx.x % TX, j = threadIdx.x / TX;
        const int ntile = (P.na + TX - 1) / TX;
        const int tile = blk.x % ntile, jb = blk.x / ntile;
        int ja = tile * TX + c;
        const bool live = ja < P.na;
        if (!live) ja = P.na - 1;   // dead columns read the last one and store nothing
        const C2<TI>* src = reinterpret_cast<const C2<TI>*>(P.src) + (long long)ja * P.src_s[P.fa] + (long long)jb * P.src_s[P.fb];
        C2<TO>* dst = reinterpret_cast<C2<TO>*>(P.dst) + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
        const long long sd = P.src_s[P.d], dd = P.dst_s[P.d];
        C2<double> v[NV];
    #pragma unroll
        for (int i = 0; i < NV; ++i) {
            if (HIN && i >= NV / 2) {
                v[i] = {0.0, 0.0};
                continue;
            }
            const int ks = j + i * T - P.gshift;   // clamped address, zero selected: no branch
            const C2<double> x = c2d(src[min(max(ks, 0), max(P.src_nd - 1, 0)) * sd]);
            v[i] = (ks >= 0 && ks < P.src_nd) ? x : C2<double>{0.0, 0.0};
        }
        column_fft_pq<double, PF, Q, FWD, TX, HIN && PF == 1>(v, lds + c, j, reinterpret_cast<const C2<double>*>(P.tw),
                                                             reinterpret_cast<const C2<double>*>(P.aux), 1);
        if (!live) return;
    #pragma unroll
        for (int i = 0; i < (HOUT ? NV / 2 : NV); ++i) {
            const int k = j + i * T - P.sshift;
            if (k >= 0 && k < P.dst_nd) dst[k * dd] = c2t<TO>(v[i]);
        }
        if ((P.post & POST_PERIOD) && j == 0) {   // node-centred periodic: X[0] again at position n
            const int k = M - P.sshift;
            if (k >= 0 && k < P.dst_nd) dst[k * dd] = c2t<TO>(v[0]);
        }
    }
};

// Contiguous complex pencils of M = PF Q points (a c2c direction along x, after an r2c one),
// ROWS pencils per block, M/8 threads per pencil in PF teams (fft_pq_team); each pencil is
// staged in LDS so that HBM is read and written in whole rows, with the zero padding of the
// gather and the truncation of the scatter.
template <int PF, int Q, int ROWS, bool FWD, typename TI, typename TO>
struct row_c2c {
    using Args = PassDesc;
    static constexpr int NT = ROWS * PF * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int M = PF * Q, T = Q / 8, TT = M / 8, HP = M + M / 8;
        __shared__ C2<double> lds[HP * ROWS];
        const int rl = threadIdx.x / TT, tt = threadIdx.x % TT;
        const int k1 = tt / T, j = tt % T;
        const int g = blk.x * ROWS + rl;
        const bool live = g < P.n1 * P.n2;
        const int gg = live ? g : 0;   // dead rows read pencil 0 and store nothing
        const int ja = gg % P.na, jb = gg / P.na;
        const C2<TI>* src = reinterpret_cast<const C2<TI>*>(P.src) + (long long)ja * P.src_s[P.fa] + (long long)jb * P.src_s[P.fb];
        C2<TO>* dst = reinterpret_cast<C2<TO>*>(P.dst) + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
        C2<double>* row = lds + rl * HP;
        const int last = max(P.src_nd - 1, 0);
        staged_gather<8>(tt, TT, M, [&](int k) {
            const int ks = k - P.gshift;
            const C2<double> x = c2d(src[min(max(ks, 0), last)]);
            return (ks >= 0 && ks < P.src_nd) ? x : C2<double>{0.0, 0.0};
        }, [&](int k, C2<double> x) { row[k] = x; });
        __syncthreads();
        auto load = [&](int m) -> C2<double> { return row[m]; };
        C2<double> v[8];
        fft_pq_team<double, PF, Q, FWD, 0>(v, row, k1, j, load, reinterpret_cast<const C2<double>*>(P.tw),
                                            reinterpret_cast<const C2<double>*>(P.aux), 1);
        __syncthreads();
    #pragma unroll
        for (int i = 0; i < 8; ++i) row[k1 + PF * (j + i * T)] = v[i];
        if ((P.post & POST_PERIOD) && k1 == 0 && j == 0) row[M] = v[0];   // node-centred periodic copy
        __syncthreads();
        if (!live) return;
        for (int k = tt; k < P.dst_nd; k += TT) {
            const int p = k + P.sshift;
            if (p < P.pen_out) dst[k] = c2t<TO>(row[p]);
        }
    }
};

// Strided complex pencils of M = PF Q points through thread teams (fft_pq_team): the TX
// columns are staged in LDS (rows read as contiguous runs of TX values), M/8 threads per
// column form the PF twiddled inputs and run the Q-point FFTs. Same plan semantics as col_c2c.
template <int PF, int Q, int TX, bool FWD, typename TI, typename TO>
struct col_c2c_team {
    using Args = PassDesc;
    static constexpr int NT = TX * PF * Q / 8;
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        constexpr int M = PF * Q, T = Q / 8, TT = M / 8;
        __shared__ C2<double> lds[M * TX];
        const int c = threadIdx.x % TX, tt = threadIdx.x / TX;
        const int k1 = tt / T, j = tt % T;
        const int ntile = (P.na + TX - 1) / TX;
        const int tile = blk.x % ntile, jb = blk.x / ntile;
        int ja = tile * TX + c;
        const bool live = ja < P.na;
        if (!live) ja = P.na - 1;   // dead columns read the last one and store nothing
        const C2<TI>* src = reinterpret_cast<const C2<TI>*>(P.src) + (long long)ja * P.src_s[P.fa] + (long long)jb * P.src_s[P.fb];
        C2<TO>* dst = reinterpret_cast<C2<TO>*>(P.dst) + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
        const long long sd = P.src_s[P.d], dd = P.dst_s[P.d];
        // all eight loads (clamped addresses, zero selected afterwards) before the LDS writes
        const int last = max(P.src_nd - 1, 0);
        C2<double> u[8];
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int ks = tt + i * TT - P.gshift;
            const C2<double> x = c2d(src[min(max(ks, 0), last) * sd]);
            u[i] = (ks >= 0 && ks < P.src_nd) ? x : C2<double>{0.0, 0.0};
        }
    #pragma unroll
        for (int i = 0; i < 8; ++i) lds[(tt + i * TT) * TX + c] = u[i];
        __syncthreads();
        auto load = [&](int m) -> C2<double> { return lds[m * TX + c]; };
        C2<double> v[8];
        fft_pq_team<double, PF, Q, FWD, TX>(v, lds + c, k1, j, load, reinterpret_cast<const C2<double>*>(P.tw),
                                             reinterpret_cast<const C2<double>*>(P.aux), 1);
        if (!live) return;
    #pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = k1 + PF * (j + i * T) - P.sshift;
            if (k >= 0 && k < P.dst_nd) dst[k * dd] = c2t<TO>(v[i]);
        }
        if ((P.post & POST_PERIOD) && k1 == 0 && j == 0) {   // node-centred periodic: X[0] again at position n
            const int k = M - P.sshift;
            if (k >= 0 && k < P.dst_nd) dst[k * dd] = c2t<TO>(v[0]);
        }
    }
};

template <int PF, int Q, int R, bool COL, typename TI, typename TO>
void launch_pen_h(const PassDesc& P, hipStream_t s) {
    const int npen = P.n1 * P.n2;
    const int nb = COL ? ((P.na + R - 1) / R) * (npen / P.na) : (npen + R - 1) / R;
    dim3 grid(nb);
    dim3 block(R * PF * Q / 8);
    switch (P.kind) {
        case REDFT10: launch_k<pen_r2r<PF, Q, R, REDFT10, COL, TI, TO>>(grid, block, 0, s, P); break;
        case REDFT01: launch_k<pen_r2r<PF, Q, R, REDFT01, COL, TI, TO>>(grid, block, 0, s, P); break;
        case REDFT11: launch_k<pen_r2r<PF, Q, R, REDFT11, COL, TI, TO>>(grid, block, 0, s, P); break;
        case RODFT10: launch_k<pen_r2r<PF, Q, R, RODFT10, COL, TI, TO>>(grid, block, 0, s, P); break;
        case RODFT01: launch_k<pen_r2r<PF, Q, R, RODFT01, COL, TI, TO>>(grid, block, 0, s, P); break;
        case RODFT11: launch_k<pen_r2r<PF, Q, R, RODFT11, COL, TI, TO>>(grid, block, 0, s, P); break;
        case REDFT00: launch_k<pen_r2r<PF, Q, R, REDFT00, COL, TI, TO>>(grid, block, 0, s, P); break;
        case RODFT00: launch_k<pen_r2r<PF, Q, R, RODFT00, COL, TI, TO>>(grid, block, 0, s, P); break;
        default: fail("r2r pencil kernel: kind %d", P.kind);
    }
}
// H = PF Q (3 <= PF <= 7): rows, 2048 / H per block (about 256 threads, 36 KB of LDS);
// columns, 32 pencils per block up to H = 64, then 16, 8, 4 (LDS (H + 1) R complex <= 64 KB)
template <int PF, bool COL, typename TI, typename TO>
void launch_pen_pf(int q, const PassDesc& P, hipStream_t s) {
    constexpr int H8 = PF * 8, H16 = PF * 16, H32 = PF * 32, H64 = PF * 64, H128 = PF * 128;
    switch (q) {
        case 8: launch_pen_h<PF, 8, COL ? (H8 <= 64 ? 32 : 16) : 2048 / H8, COL, TI, TO>(P, s); break;
        case 16: launch_pen_h<PF, 16, COL ? (H16 <= 64 ? 32 : 16) : 2048 / H16, COL, TI, TO>(P, s); break;
        case 32: launch_pen_h<PF, 32, COL ? (H32 <= 64 ? 32 : H32 <= 256 ? 16 : 8) : 2048 / H32, COL, TI, TO>(P, s); break;
        case 64: launch_pen_h<PF, 64, COL ? (H64 <= 256 ? 16 : H64 <= 512 ? 8 : 4) : 2048 / H64, COL, TI, TO>(P, s); break;
        case 128: launch_pen_h<PF, 128, COL ? (H128 <= 512 ? 8 : 4) : 2048 / H128, COL, TI, TO>(P, s); break;
        default: fail("r2r pencil kernel: length %d", P.n);
    }
}

template <int PF, int Q, int ROWS, typename TI, typename TO>
void launch_row_rc_h(const PassDesc& P, hipStream_t s) {
    dim3 grid((P.n1 * P.n2 + ROWS - 1) / ROWS);
    dim3 block(ROWS * PF * Q / 8);
    if (P.op == OP_R2C) launch_k<row_rc<PF, Q, ROWS, true, TI, TO>>(grid, block, 0, s, P);
    else launch_k<row_rc<PF, Q, ROWS, false, TI, TO>>(grid, block, 0, s, P);
}
// H = PF Q, 3 <= PF <= 7: about 256 threads (2048 / H rows) and 36 KB of LDS per block
template <int PF, typename TI, typename TO>
void launch_row_rc_pf(int q, const PassDesc& P, hipStream_t s) {
    switch (q) {
        case 8: launch_row_rc_h<PF, 8, 2048 / (PF * 8), TI, TO>(P, s); break;
        case 16: launch_row_rc_h<PF, 16, 2048 / (PF * 16), TI, TO>(P, s); break;
        case 32: launch_row_rc_h<PF, 32, 2048 / (PF * 32), TI, TO>(P, s); break;
        case 64: launch_row_rc_h<PF, 64, 2048 / (PF * 64), TI, TO>(P, s); break;
        case 128: launch_row_rc_h<PF, 128, 2048 / (PF * 128), TI, TO>(P, s); break;
        default: fail("row rc: length %d", P.n);
    }
}

template <int PF, int Q, int TX, typename TI, typename TO>
void launch_col_real_m(const PassDesc& P, hipStream_t s) {
    const int npair = (P.na + 1) / 2;
    const int ntile = (npair + TX - 1) / TX;
    const int nb = P.n1 * P.n2 / P.na;
    dim3 grid(ntile * nb);
    if (P.op == OP_R2C) launch_k<col_real<PF, Q, TX, true, TI, TO>>(grid, dim3(TX * Q / 8), 0, s, P);
    else launch_k<col_real<PF, Q, TX, false, TI, TO>>(grid, dim3(TX * Q / 8), 0, s, P);
}
// mixed radix: TX Q = 1024 (PF 3) or 512 (PF 5), LDS about 48 / 40 KB
template <int PF, typename TI, typename TO>
void launch_col_real_pf(int q, const PassDesc& P, hipStream_t s) {
    constexpr int TQ = PF == 3 ? 1024 : 512;
    switch (q) {
        case 8: launch_col_real_m<PF, 8, TQ / 8, TI, TO>(P, s); break;
        case 16: launch_col_real_m<PF, 16, TQ / 16, TI, TO>(P, s); break;
        case 32: launch_col_real_m<PF, 32, TQ / 32, TI, TO>(P, s); break;
        case 64: launch_col_real_m<PF, 64, TQ / 64, TI, TO>(P, s); break;
        case 128: launch_col_real_m<PF, 128, TQ / 128, TI, TO>(P, s); break;
        default: fail("col real: length %d", P.n);
    }
}

template <int PF, int Q, int TX, typename TI, typename TO>
void launch_col_c2c_m(const PassDesc& P, hipStream_t s) {
    constexpr int M = PF * Q;
    const int ntile = (P.na + TX - 1) / TX;
    dim3 grid(ntile * (P.n1 * P.n2 / P.na));
    dim3 block(TX * Q / 8);
    const bool hin = P.gshift == 0 && P.src_nd <= M / 2, hout = P.sshift == 0 && P.dst_nd <= M / 2;
    if (P.sign == FWD) {
        if (hin) launch_k<col_c2c<PF, Q, TX, true, true, false, TI, TO>>(grid, block, 0, s, P);
        else launch_k<col_c2c<PF, Q, TX, true, false, false, TI, TO>>(grid, block, 0, s, P);
    } else {
        if (hout) launch_k<col_c2c<PF, Q, TX, false, false, true, TI, TO>>(grid, block, 0, s, P);
        else launch_k<col_c2c<PF, Q, TX, false, false, false, TI, TO>>(grid, block, 0, s, P);
    }
}
template <int PF, int Q, int TX, typename TI, typename TO>
void launch_col_c2c_t(const PassDesc& P, hipStream_t s) {
    const int ntile = (P.na + TX - 1) / TX;
    dim3 grid(ntile * (P.n1 * P.n2 / P.na));
    dim3 block(TX * PF * Q / 8);
    if (P.sign == FWD) launch_k<col_c2c_team<PF, Q, TX, true, TI, TO>>(grid, block, 0, s, P);
    else launch_k<col_c2c_team<PF, Q, TX, false, TI, TO>>(grid, block, 0, s, P);
}
template <int PF, int Q, typename TI, typename TO>
void launch_row_c2c_h(const PassDesc& P, hipStream_t s) {
    constexpr int ROWS = 2048 / (PF * Q) > 0 ? 2048 / (PF * Q) : 1;   // about 256 threads, 36 KB of LDS
    dim3 grid((P.n1 * P.n2 + ROWS - 1) / ROWS);
    dim3 block(ROWS * PF * Q / 8);
    if (P.sign == FWD) launch_k<row_c2c<PF, Q, ROWS, true, TI, TO>>(grid, block, 0, s, P);
    else launch_k<row_c2c<PF, Q, ROWS, false, TI, TO>>(grid, block, 0, s, P);
}
template <int PF, typename TI, typename TO>
void launch_row_c2c_pf(int q, const PassDesc& P, hipStream_t s) {
    switch (q) {
        case 8: launch_row_c2c_h<PF, 8, TI, TO>(P, s); break;
        case 16: launch_row_c2c_h<PF, 16, TI, TO>(P, s); break;
        case 32: launch_row_c2c_h<PF, 32, TI, TO>(P, s); break;
        case 64: launch_row_c2c_h<PF, 64, TI, TO>(P, s); break;
        case 128: launch_row_c2c_h<PF, 128, TI, TO>(P, s); break;
        default: fail("row c2c: length %d", P.n);
    }
}
// thread teams: TX M <= 4096 (64 KB of LDS), TX >= 4
template <int PF, typename TI, typename TO>
void launch_col_c2c_team_pf(int q, const PassDesc& P, hipStream_t s) {
    switch (q) {
        case 8: launch_col_c2c_t<PF, 8, 16, TI, TO>(P, s); break;
        case 16: launch_col_c2c_t<PF, 16, 16, TI, TO>(P, s); break;
        case 32: launch_col_c2c_t<PF, 32, (PF * 32 <= 128 ? 16 : 8), TI, TO>(P, s); break;
        case 64: launch_col_c2c_t<PF, 64, (PF * 64 <= 256 ? 16 : PF * 64 <= 512 ? 8 : 4), TI, TO>(P, s); break;
        case 128: launch_col_c2c_t<PF, 128, (PF * 128 <= 512 ? 8 : 4), TI, TO>(P, s); break;
        default: fail("column c2c: length %d", P.n);
    }
}
// mixed radix, 8 PF values per thread: TX Q = 1024 (PF 3) or 512, LDS 48 / 40 KB
template <int PF, typename TI, typename TO>
void launch_col_c2c_pf(int q, const PassDesc& P, hipStream_t s) {
    constexpr int TQ = PF == 3 ? 1024 : 512;
    switch (q) {
        case 8: launch_col_c2c_m<PF, 8, TQ / 8, TI, TO>(P, s); break;
        case 16: launch_col_c2c_m<PF, 16, TQ / 16, TI, TO>(P, s); break;
        case 32: launch_col_c2c_m<PF, 32, TQ / 32, TI, TO>(P, s); break;
        case 64: launch_col_c2c_m<PF, 64, TQ / 64, TI, TO>(P, s); break;
        case 128: launch_col_c2c_m<PF, 128, TQ / 128, TI, TO>(P, s); break;
        default: fail("column c2c: length %d", P.n);
    }
}


template <int M, int TX, typename S>
void launch_col_rz_m(const RzArgs& A, hipStream_t s) {
    const int ntile = ((A.nx + 1) / 2 + TX - 1) / TX;
    launch_k<col_rz<M, TX, S>>(dim3(ntile * A.ny, A.lda), dim3(TX * M / 8), 0, s, A);
}

template <int PF, int Q, typename S>
void launch_row_rz_h(const RzArgs& A, hipStream_t s) {
    constexpr int ROWS = 2048 / (PF * Q) > 0 ? 2048 / (PF * Q) : 1;   // about 256 threads, 36 KB of LDS
    launch_k<row_rz<PF, Q, ROWS, S>>(dim3((A.rows + ROWS - 1) / ROWS, A.lda), dim3(ROWS * PF * Q / 8), 0, s, A);
}

template <int PF, typename S>
void launch_row_rz_pf(int q, const RzArgs& A, hipStream_t s) {
    switch (q) {
        case 8: launch_row_rz_h<PF, 8, S>(A, s); break;
        case 16: launch_row_rz_h<PF, 16, S>(A, s); break;
        case 32: launch_row_rz_h<PF, 32, S>(A, s); break;
        case 64: launch_row_rz_h<PF, 64, S>(A, s); break;
        case 128: launch_row_rz_h<PF, 128, S>(A, s); break;
        default: fail("fused r2c row: length %d", A.M);
    }
}

template <typename TI, typename TO>
void launch_row_rc_t(const PassDesc& P, hipStream_t s) {
    int pf = 1, q = 0;
    split_pq(P.n / 2, pf, q);
    if (pf != 1) {
        if (pf == 3) launch_row_rc_pf<3, TI, TO>(q, P, s);
        else if (pf == 5) launch_row_rc_pf<5, TI, TO>(q, P, s);
        else launch_row_rc_pf<7, TI, TO>(q, P, s);
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (P.n / 2) {
        case 16: launch_row_rc_h<1, 16, 128, TI, TO>(P, s); break;
        case 32: launch_row_rc_h<1, 32, 64, TI, TO>(P, s); break;
        case 64: launch_row_rc_h<1, 64, 32, TI, TO>(P, s); break;
        case 128: launch_row_rc_h<1, 128, 16, TI, TO>(P, s); break;
        case 256: launch_row_rc_h<1, 256, 8, TI, TO>(P, s); break;
        case 512: launch_row_rc_h<1, 512, 4, TI, TO>(P, s); break;
        case 1024: launch_row_rc_h<1, 1024, 2, TI, TO>(P, s); break;
        default: fail("row rc: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename TI, typename TO>
void launch_col_real_t(const PassDesc& P, hipStream_t s) {
    int pf = 1, q = 0;
    split_pq(P.n, pf, q);
    if (pf != 1) {
        if (pf == 3) launch_col_real_pf<3, TI, TO>(q, P, s);
        else launch_col_real_pf<5, TI, TO>(q, P, s);
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (P.n) {
        case 32: launch_col_real_m<1, 32, 16, TI, TO>(P, s); break;
        case 64: launch_col_real_m<1, 64, 16, TI, TO>(P, s); break;
        case 128: launch_col_real_m<1, 128, 8, TI, TO>(P, s); break;
        case 256: launch_col_real_m<1, 256, 8, TI, TO>(P, s); break;
        case 512: launch_col_real_m<1, 512, 4, TI, TO>(P, s); break;     // LDS (M + 2) TX complex: 33 KB at most
        case 1024: launch_col_real_m<1, 1024, 2, TI, TO>(P, s); break;
        default: fail("col real: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename TI, typename TO>
void launch_col_c2c_t(const PassDesc& P, hipStream_t s) {
    int pf = 1, q = 0;
    split_pq(P.n, pf, q);
    if (pf != 1) {
        // 24 values per thread for PF = 3 (HBM loads straight into registers); thread teams over
        // the LDS-staged columns for 5 and 7 (40 or 56 values would crowd the registers)
        if (pf == 3) launch_col_c2c_pf<3, TI, TO>(q, P, s);
        else if (pf == 5) launch_col_c2c_team_pf<5, TI, TO>(q, P, s);
        else launch_col_c2c_team_pf<7, TI, TO>(q, P, s);
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (P.n) {
        case 32: launch_col_c2c_m<1, 32, 16, TI, TO>(P, s); break;
        case 64: launch_col_c2c_m<1, 64, 16, TI, TO>(P, s); break;
        case 128: launch_col_c2c_m<1, 128, 8, TI, TO>(P, s); break;
        case 256: launch_col_c2c_m<1, 256, 8, TI, TO>(P, s); break;
        case 512: launch_col_c2c_m<1, 512, 8, TI, TO>(P, s); break;
        case 1024: launch_col_c2c_m<1, 1024, 4, TI, TO>(P, s); break;
        case 2048: launch_col_c2c_m<1, 2048, 2, TI, TO>(P, s); break;
        default: fail("column c2c: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename TI, typename TO>
void launch_row_c2c_t(const PassDesc& P, hipStream_t s) {
    int pf = 1, q = 0;
    split_pq(P.n, pf, q);
    if (pf == 1) {
        switch (q) {
            case 16: launch_row_c2c_h<1, 16, TI, TO>(P, s); break;
            case 32: launch_row_c2c_h<1, 32, TI, TO>(P, s); break;
            case 64: launch_row_c2c_h<1, 64, TI, TO>(P, s); break;
            case 128: launch_row_c2c_h<1, 128, TI, TO>(P, s); break;
            case 256: launch_row_c2c_h<1, 256, TI, TO>(P, s); break;
            case 512: launch_row_c2c_h<1, 512, TI, TO>(P, s); break;
            case 1024: launch_row_c2c_h<1, 1024, TI, TO>(P, s); break;
            default: fail("row c2c: length %d", P.n);
        }
    } else if (pf == 3) {
        launch_row_c2c_pf<3, TI, TO>(q, P, s);
    } else if (pf == 5) {
        launch_row_c2c_pf<5, TI, TO>(q, P, s);
    } else {
        launch_row_c2c_pf<7, TI, TO>(q, P, s);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename S>
void launch_col_rz_t(const RzArgs& A, hipStream_t s) {
    switch (A.M) {
        case 32: launch_col_rz_m<32, 16, S>(A, s); break;
        case 64: launch_col_rz_m<64, 16, S>(A, s); break;
        case 128: launch_col_rz_m<128, 8, S>(A, s); break;
        case 256: launch_col_rz_m<256, 8, S>(A, s); break;
        case 512: launch_col_rz_m<512, 4, S>(A, s); break;
        case 1024: launch_col_rz_m<1024, 2, S>(A, s); break;
        default: fail("fused r2c column: length %d", A.M);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename S>
void launch_row_rz_t(const RzArgs& A, hipStream_t s) {
    int pf = 1, q = 0;
    split_pq(A.M / 2, pf, q);
    if (pf == 1) {
        switch (q) {
            case 16: launch_row_rz_h<1, 16, S>(A, s); break;
            case 32: launch_row_rz_h<1, 32, S>(A, s); break;
            case 64: launch_row_rz_h<1, 64, S>(A, s); break;
            case 128: launch_row_rz_h<1, 128, S>(A, s); break;
            case 256: launch_row_rz_h<1, 256, S>(A, s); break;
            case 512: launch_row_rz_h<1, 512, S>(A, s); break;
            case 1024: launch_row_rz_h<1, 1024, S>(A, s); break;
            default: fail("fused r2c row: length %d", A.M);
        }
    } else if (pf == 3) {
        launch_row_rz_pf<3, S>(q, A, s);
    } else if (pf == 5) {
        launch_row_rz_pf<5, S>(q, A, s);
    } else {
        launch_row_rz_pf<7, S>(q, A, s);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename TI, typename TO>
void launch_row_r2r_t(const PassDesc& P, hipStream_t s) {
    const int h = r2r_h(P.kind, P.n);
    int pf = 1, q = 0;
    split_pq(h, pf, q);
    if (pf != 1) {
        if (pf == 3) launch_pen_pf<3, false, TI, TO>(q, P, s);
        else if (pf == 5) launch_pen_pf<5, false, TI, TO>(q, P, s);
        else launch_pen_pf<7, false, TI, TO>(q, P, s);
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (h) {
        case 16: launch_pen_h<1, 16, 128, false, TI, TO>(P, s); break;
        case 32: launch_pen_h<1, 32, 64, false, TI, TO>(P, s); break;
        case 64: launch_pen_h<1, 64, 32, false, TI, TO>(P, s); break;
        case 128: launch_pen_h<1, 128, 16, false, TI, TO>(P, s); break;
        case 256: launch_pen_h<1, 256, 8, false, TI, TO>(P, s); break;
        case 512: launch_pen_h<1, 512, 4, false, TI, TO>(P, s); break;
        case 1024: launch_pen_h<1, 1024, 2, false, TI, TO>(P, s); break;
        default: fail("row r2r: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename TI, typename TO>
void launch_col_r2r_t(const PassDesc& P, hipStream_t s) {
    const int h = r2r_h(P.kind, P.n);
    int pf = 1, q = 0;
    split_pq(h, pf, q);
    if (pf != 1) {
        if (pf == 3) launch_pen_pf<3, true, TI, TO>(q, P, s);
        else if (pf == 5) launch_pen_pf<5, true, TI, TO>(q, P, s);
        else launch_pen_pf<7, true, TI, TO>(q, P, s);
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (h) {
        case 16: launch_pen_h<1, 16, 32, true, TI, TO>(P, s); break;
        case 32: launch_pen_h<1, 32, 16, true, TI, TO>(P, s); break;
        case 64: launch_pen_h<1, 64, 16, true, TI, TO>(P, s); break;
        case 128: launch_pen_h<1, 128, 16, true, TI, TO>(P, s); break;
        case 256: launch_pen_h<1, 256, 8, true, TI, TO>(P, s); break;
        case 512: launch_pen_h<1, 512, 4, true, TI, TO>(P, s); break;
        case 1024: launch_pen_h<1, 1024, 2, true, TI, TO>(P, s); break;
        default: fail("column r2r: length %d", P.n);
    }
    HIP_CHECK(hipGetLastError());
}

}  // namespace

// the passes of one (TI, TO) storage combination, compiled in its own translation unit
enum RowpassKernel { RP_ROW_R2R, RP_COL_R2R, RP_ROW_RC, RP_COL_REAL, RP_COL_C2C, RP_ROW_C2C };
void rowpass_dd(int which, const PassDesc& P, hipStream_t s);
void rowpass_df(int which, const PassDesc& P, hipStream_t s);
void rowpass_ff(int which, const PassDesc& P, hipStream_t s);
void rowpass_fd(int which, const PassDesc& P, hipStream_t s);
void rz_d(int row, const RzArgs& A, hipStream_t s);
void rz_f(int row, const RzArgs& A, hipStream_t s);

#define FLUPS_ROWPASS_IO(NAME, TI, TO)                                  \
    void NAME(int which, const PassDesc& P, hipStream_t s) {            \
        switch (which) {                                                \
            case RP_ROW_R2R: launch_row_r2r_t<TI, TO>(P, s); break;     \
            case RP_COL_R2R: launch_col_r2r_t<TI, TO>(P, s); break;     \
            case RP_ROW_RC: launch_row_rc_t<TI, TO>(P, s); break;       \
            case RP_COL_REAL: launch_col_real_t<TI, TO>(P, s); break;   \
            case RP_COL_C2C: launch_col_c2c_t<TI, TO>(P, s); break;     \
            default: launch_row_c2c_t<TI, TO>(P, s); break;             \
        }                                                               \
    }
#define FLUPS_RZ(NAME, S)                                    \
    void NAME(int row, const RzArgs& A, hipStream_t s) {     \
        if (row) launch_row_rz_t<S>(A, s);                   \
        else launch_col_rz_t<S>(A, s);                       \
    }

}  // namespace flups
