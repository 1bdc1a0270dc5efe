// General pencil pass for every boundary condition and length.
//
// One launch transforms one component along one direction: each block takes
// `ppb` pencils, gathers them from the source array into LDS (applying the
// pencil-switch shift and zero padding of src/SwitchTopoX_a2a.cpp:174-247),
// runs the 1-D transform of the plan with its input/output offsets
// (src/FFTW_plan_dim.cpp:635-720) through a mixed-radix Stockham FFT held in
// LDS, applies the post-correction (src/FFTW_plan_dim.cpp:497-622) and
// scatters the pencil to the destination array (truncating on the way back).
//
// r2r kinds are computed from one complex FFT of an extended sequence:
//   REDFT00 length 2(n-1) even extension, RODFT00 length 2(n+1) odd extension,
//   REDFT10/01/11 length 2n with quarter-sample twiddles, and
//   RODFT10/01/11 through DST(x) = reflections of DCT of (-1)^j x or reversed x.
// r2c/c2r of even length n use the half-length packed complex FFT.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "core.hpp"
#include "fft_core.hpp"
#include "group.hpp"

namespace flups {

namespace {

struct cd {
    double x, y;
};
__device__ __forceinline__ cd mk(double a, double b) { return {a, b}; }
__device__ __forceinline__ cd add(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd sub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd mul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd conj(cd a) { return {a.x, -a.y}; }
__device__ __forceinline__ cd scal(cd a, double s) { return {a.x * s, a.y * s}; }
// exp(-i pi t), accurate for any t
__device__ __forceinline__ cd expmpi(double t) {
    double s, c;
    sincospi(t, &s, &c);
    return {c, -s};
}

// twiddle omega_M^e, 0 <= e < M, with the pass sign (table holds exp(-2 pi i t/M))
__device__ __forceinline__ cd twid(const double* tw, unsigned e, int sign) {
    cd w = {tw[2 * e], tw[2 * e + 1]};
    return sign > 0 ? conj(w) : w;
}

// One radix-p butterfly of a Stockham stage: inputs a[j + q mp] (q < p, j = b span + k),
// twiddled by w_sp^(q k), then a p-point DFT whose output u lands at o[u span]:
//   out[b sp + u span + k] = sum_q a[j + q mp] w_sp^(q k) w_p^(q u),  sp = span p, mp = M / p.
template <int P>
__device__ __forceinline__ void butterfly(const cd* a, cd* o, unsigned j, unsigned mp, unsigned span, unsigned k,
                                          unsigned scale, const double* tw, int sign, unsigned M) {
    cd x[P];
    x[0] = a[j];
#pragma unroll
    for (int q = 1; q < P; ++q) x[q] = mul(a[j + q * mp], twid(tw, q * k * scale, sign));   // q k scale < M
    if (P == 2) {
        o[0] = add(x[0], x[1]);
        o[span] = sub(x[0], x[1]);
    } else if (P == 4) {
        const cd s0 = add(x[0], x[2]), d0 = sub(x[0], x[2]), s1 = add(x[1], x[3]), t = sub(x[1], x[3]);
        const cd d1 = sign > 0 ? mk(-t.y, t.x) : mk(t.y, -t.x);   // w_4 (x1 - x3), w_4 = +-i
        o[0] = add(s0, s1);
        o[span] = add(d0, d1);
        o[2 * span] = sub(s0, s1);
        o[3 * span] = sub(d0, d1);
    } else {
        cd w[P];
        w[0] = mk(1.0, 0.0);
#pragma unroll
        for (int m = 1; m < P; ++m) w[m] = twid(tw, m * (M / P), sign);   // w_p^m
#pragma unroll
        for (int u = 0; u < P; ++u) {
            cd acc = x[0];
#pragma unroll
            for (int q = 1; q < P; ++q) acc = add(acc, mul(x[q], w[(q * u) % P]));
            o[u * span] = acc;
        }
    }
}

// one Stockham stage, all butterflies of all pencils of the block (radix 2, 3, 4, 5, 7
// unrolled; any other prime one output at a time)
__device__ void stockham_stage(const cd* in, cd* out, int M, int p, int span, const double* tw, int sign, int npen) {
    const unsigned um = M, mp = M / p, sp = span * p, scale = M / sp, us = span;
    if (p == 2 || p == 3 || p == 4 || p == 5 || p == 7) {
        for (unsigned i = threadIdx.x; i < (unsigned)npen * mp; i += blockDim.x) {
            const unsigned pen = i / mp, j = i - pen * mp;
            const unsigned b = j / us, k = j - b * us;
            const cd* a = in + (size_t)pen * M;
            cd* o = out + (size_t)pen * M + b * sp + k;
            switch (p) {
                case 2: butterfly<2>(a, o, j, mp, us, k, scale, tw, sign, um); break;
                case 3: butterfly<3>(a, o, j, mp, us, k, scale, tw, sign, um); break;
                case 4: butterfly<4>(a, o, j, mp, us, k, scale, tw, sign, um); break;
                case 5: butterfly<5>(a, o, j, mp, us, k, scale, tw, sign, um); break;
                default: butterfly<7>(a, o, j, mp, us, k, scale, tw, sign, um); break;
            }
        }
        return;
    }
    for (unsigned i = threadIdx.x; i < (unsigned)(npen * M); i += blockDim.x) {
        const unsigned pen = i / um, o = i - pen * um;
        const unsigned b = o / sp, r = o - b * sp;          // r = k + u * span
        const unsigned k = r % us;
        const unsigned j = b * us + k;
        const cd* a = in + (size_t)pen * M;
        const unsigned base = r * scale;                    // < M
        cd acc = a[j];
        unsigned e = 0;
        for (int q = 1; q < p; ++q) {
            e += base;
            if (e >= um) e -= um;
            acc = add(acc, mul(a[j + q * mp], twid(tw, e, sign)));
        }
        out[(size_t)pen * M + o] = acc;
    }
}

}  // namespace

// TI / TO: storage of the source / destination reals (double, or float for the stage arrays of
// FP32 solves); the pencil buffer and the arithmetic are FP64
template <typename TI, typename TO>
struct generic_pass {
    using Args = PassDesc;
    static constexpr int NT = 256;
    static constexpr int kWaves = 2;   // at most 256 registers per lane: two waves per SIMD
    static __device__ __forceinline__ void run(const PassDesc& P, Blk blk) {
        extern __shared__ double smem[];
        const int len = P.pen_in * P.nf_in;   // doubles per pencil (same before and after)
        const int npen_total = P.n1 * P.n2;
        const int pen0 = blk.x * P.ppb;
        const int npen = min(P.ppb, npen_total - pen0);
        if (npen <= 0) return;
        double* pb = smem;                                       // pencils, ppb * len doubles
        cd* A = reinterpret_cast<cd*>(smem + (size_t)P.ppb * len + ((P.ppb * len) & 1));
        cd* B = P.alias ? reinterpret_cast<cd*>(smem) : A + (size_t)P.ppb * P.M;   // free while the FFT runs
        const int d = P.d;

        // ---- gather (pencil switch: shift, zero padding) ----
        // Neighbouring threads take neighbouring elements of a pencil when the pencil
        // is contiguous in the source, neighbouring pencils (adjacent along fa) otherwise.
        {
            const bool efast = P.src_s[d] == 1;
            const int nfi = P.nf_in;
            const int per = P.pen_in;   // elements per pencil
            // element i: pencil and position; nf_in is 1 (real) or 2 (complex)
            auto at = [&](int i, int& pen, int& k) {
                pen = efast ? i / per : i % npen;
                k = efast ? i - pen * per : i / npen;
            };
            const int last = max(P.src_nd - 1, 0);
            staged_gather<4>(threadIdx.x, blockDim.x, npen * per, [&](int i) -> cd {
                int pen, k;
                at(i, pen, k);
                const int g = pen0 + pen;
                const int ja = g % P.na, jb = g / P.na;
                const int ks = k - P.gshift;
                const long long idx = (long long)min(max(ks, 0), last) * P.src_s[d] + (long long)ja * P.src_s[P.fa] +
                                      (long long)jb * P.src_s[P.fb];
                const TI* sp = reinterpret_cast<const TI*>(P.src) + idx * P.src_nf;
                const double x0 = (double)sp[0], x1 = nfi > 1 ? (double)sp[1] : 0.0;
                const bool in = ks >= 0 && ks < P.src_nd;
                return cd{in ? x0 : 0.0, in ? x1 : 0.0};
            }, [&](int i, cd x) {
                int pen, k;
                at(i, pen, k);
                double* o = pb + (size_t)pen * len + (size_t)k * nfi;
                o[0] = x.x;
                if (nfi > 1) o[1] = x.y;
            });
        }
        __syncthreads();

        // ---- pre-map: pencil -> complex sequence of length M ----
        const int n = P.n, M = P.M, st = P.r2r_stride;
        int kind = P.kind;
        // DST kinds are computed through the matching DCT
        const bool dst = (kind >= RODFT10);
        const int dct = dst ? kind - 4 : kind;   // RODFT10 -> REDFT10, RODFT01 -> REDFT01, RODFT11 -> REDFT11
        for (int i = threadIdx.x; i < npen * M; i += blockDim.x) {
            const int pen = i / M, j = i - pen * M;
            const double* x = pb + (size_t)pen * len;
            cd v = {0, 0};
            if (P.op == OP_C2C) {
                v = mk(x[2 * (P.si + j)], x[2 * (P.si + j) + 1]);
            } else if (P.op == OP_R2C) {
                if ((n & 1) == 0) v = mk(x[P.si + 2 * j], x[P.si + 2 * j + 1]);
                else v = mk(x[P.si + j], 0.0);
            } else if (P.op == OP_C2R) {
                if ((n & 1) == 0) {
                    const int h = n / 2;
                    cd a = mk(x[2 * (P.si + j)], x[2 * (P.si + j) + 1]);
                    cd b = mk(x[2 * (P.si + h - j)], x[2 * (P.si + h - j) + 1]);
                    if (j == 0) { a.y = 0.0; b.y = 0.0; }   // DC and Nyquist are real
                    const cd bc = conj(b);
                    const cd e = add(a, bc), f = sub(a, bc);
                    const cd w = conj(expmpi(2.0 * j / n));          // exp(+2 pi i j/n)
                    const cd t = mul(w, f);
                    v = add(e, mk(-t.y, t.x));                          // e + i * w * f
                } else {
                    if (j == 0) v = mk(x[2 * P.si], 0.0);
                    else if (j <= n / 2) v = mk(x[2 * (P.si + j)], x[2 * (P.si + j) + 1]);
                    else v = conj(mk(x[2 * (P.si + n - j)], x[2 * (P.si + n - j) + 1]));
                }
            } else if (P.op == OP_R2R) {
                // input element jj of the (possibly reflected) real sequence
                auto xin = [&](int jj) -> double {
                    if (kind == RODFT10) return (jj & 1) ? -x[P.si + st * jj] : x[P.si + st * jj];
                    if (kind == RODFT01 || kind == RODFT11) return x[P.si + st * (n - 1 - jj)];
                    return x[P.si + st * jj];
                };
                if (kind == REDFT00) {
                    const int jj = (j < n) ? j : M - j;
                    v = mk(xin(jj), 0.0);
                } else if (kind == RODFT00) {
                    if (j >= 1 && j <= n) v = mk(xin(j - 1), 0.0);
                    else if (j >= n + 2) v = mk(-xin(M - 1 - j), 0.0);
                } else if (dct == REDFT10) {
                    v = mk(xin(j < n ? j : M - 1 - j), 0.0);
                } else if (dct == REDFT01) {
                    if (j < n) v = scal(conj(expmpi(0.5 * j / n)), (j == 0 ? 1.0 : 2.0) * xin(j));
                } else {   // REDFT11
                    if (j < n) v = scal(expmpi(0.5 * j / n), xin(j));
                }
            }
            A[(size_t)pen * M + j] = v;
        }
        __syncthreads();

        // ---- FFT ----
        int sgn = P.sign;
        if (P.op == OP_R2C) sgn = FWD;
        if (P.op == OP_C2R) sgn = BWD;
        if (P.op == OP_R2R) sgn = (dct == REDFT01) ? BWD : FWD;
        cd* in = A;
        cd* out = B;
        int span = 1;
        for (int s = 0; s < P.nst; ++s) {
            stockham_stage(in, out, M, P.rad[s], span, P.tw, sgn, npen);
            span *= P.rad[s];
            __syncthreads();
            cd* t = in; in = out; out = t;
        }
        if (P.alias && in != A) {   // the result sits in the pencil buffer, which the post-map writes: move it
            for (int i = threadIdx.x; i < npen * M; i += blockDim.x) A[i] = in[i];
            __syncthreads();
            in = A;
        }
        const cd* R = in;

        // ---- post-map: sequence -> pencil ----
        int nwrite = 0;
        if (P.op == OP_C2C) nwrite = n;
        else if (P.op == OP_R2C) nwrite = n / 2 + 1;
        else if (P.op == OP_C2R) nwrite = ((n & 1) == 0) ? n / 2 : n;
        else if (P.op == OP_R2R) nwrite = n;
        for (int i = threadIdx.x; i < npen * nwrite; i += blockDim.x) {
            const int pen = i / nwrite, k = i - pen * nwrite;
            double* x = pb + (size_t)pen * len;
            const cd* r = R + (size_t)pen * M;
            if (P.op == OP_C2C) {
                x[2 * (P.so + k)] = r[k].x;
                x[2 * (P.so + k) + 1] = r[k].y;
            } else if (P.op == OP_R2C) {
                cd X;
                if ((n & 1) == 0) {
                    const int h = n / 2;
                    const cd zk = r[k % h], zc = conj(r[(h - k) % h]);
                    const cd e = scal(add(zk, zc), 0.5);
                    const cd df = scal(sub(zk, zc), 0.5);
                    const cd o = mk(df.y, -df.x);                       // (zk - zc) / (2 i)
                    X = add(e, mul(expmpi(2.0 * k / n), o));
                } else {
                    X = r[k];
                }
                x[2 * (P.so + k)] = X.x;
                x[2 * (P.so + k) + 1] = X.y;
            } else if (P.op == OP_C2R) {
                if ((n & 1) == 0) {
                    x[P.so + 2 * k] = r[k].x;
                    x[P.so + 2 * k + 1] = r[k].y;
                } else {
                    x[P.so + k] = r[k].x;
                }
            } else {
                double y;
                int kk = k;
                if (kind == RODFT10) kk = n - 1 - k;   // DST-II(x)[k] = DCT-II((-1)^j x)[n-1-k]
                if (kind == REDFT00) y = r[kk].x;
                else if (kind == RODFT00) y = -r[kk + 1].y;
                else if (dct == REDFT10) y = mul(expmpi(0.5 * kk / n), r[kk]).x;
                else if (dct == REDFT01) y = r[kk].x;
                else y = 2.0 * mul(expmpi((2.0 * kk + 1.0) / (4.0 * n)), r[kk]).x;
                if ((kind == RODFT01 || kind == RODFT11) && (k & 1)) y = -y;   // (-1)^k reflection
                x[P.so + st * k] = y;
            }
        }
        __syncthreads();

        // ---- post-correction ----
        if (threadIdx.x < npen) {
            double* x = pb + (size_t)threadIdx.x * len;
            const int pp = P.post;
            const bool f = pp & POST_FIRST, l = pp & POST_LAST;
            if (f) x[0] = 0.0;
            if (l) x[P.pen_out - 1] = 0.0;
            if (!f && !l && (pp & POST_PERIOD)) {
                const int nf = (P.sign == FWD) ? P.n_out_plan - 1 : P.n;
                if (P.op == OP_C2R) x[nf] = x[0];
                else { x[2 * nf] = x[0]; x[2 * nf + 1] = x[1]; }
            }
        }
        __syncthreads();

        // ---- scatter (truncation on the way back) ----
        {
            const bool efast = P.dst_s[d] == 1;
            const int nfo = P.dst_nf;
            const int per = P.dst_nd;
            for (int i = threadIdx.x; i < npen * per; i += blockDim.x) {
                const int pen = efast ? i / per : i % npen;
                const int k = efast ? i - pen * per : i / npen;
                const int kt = k + P.sshift;
                if (kt >= P.pen_out) continue;
                const int g = pen0 + pen;
                const int ja = g % P.na, jb = g / P.na;
                const double* v = pb + (size_t)pen * len + (size_t)kt * P.nf_out;
                const long long idx = (long long)k * P.dst_s[d] + (long long)ja * P.dst_s[P.fa] + (long long)jb * P.dst_s[P.fb];
                TO* dp = reinterpret_cast<TO*>(P.dst) + idx * nfo;
                for (int c = 0; c < nfo; ++c) dp[c] = (TO)v[c];
            }
        }
    }
};

// B aliases the pencil buffer when that is at least as large (c2c and r2c/c2r pencils)
bool generic_alias(const PassDesc& P) { return (size_t)P.pen_in * P.nf_in >= 2 * (size_t)P.M; }

size_t generic_lds_bytes(const PassDesc& P, int ppb) {
    const size_t len = (size_t)P.pen_in * P.nf_in;
    const size_t nbuf = generic_alias(P) ? 1 : 2;
    return (ppb * len + ((ppb * len) & 1)) * sizeof(double) + nbuf * (size_t)ppb * P.M * 2 * sizeof(double);
}

void launch_generic(const PassDesc& P, hipStream_t s) {
    const int npen = P.n1 * P.n2;
    const int blocks = (npen + P.ppb - 1) / P.ppb;
    const size_t lds = generic_lds_bytes(P, P.ppb);
    // threads: enough for one element (or butterfly) each of the block's pencils, whole waves, at most 256
    const int work = P.ppb * std::max(P.M, P.pen_in);
    const int threads = std::min(256, std::max(64, (work + 63) / 64 * 64));
    if (!P.src_f32 && !P.dst_f32) launch_k<generic_pass<double, double>>(dim3(blocks), dim3(threads), lds, s, P);
    else if (!P.src_f32) launch_k<generic_pass<double, float>>(dim3(blocks), dim3(threads), lds, s, P);
    else if (P.dst_f32) launch_k<generic_pass<float, float>>(dim3(blocks), dim3(threads), lds, s, P);
    else launch_k<generic_pass<float, double>>(dim3(blocks), dim3(threads), lds, s, P);
    HIP_CHECK(hipGetLastError());
}

}  // namespace flups
