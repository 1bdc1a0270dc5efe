// Register-resident FFT building blocks shared by the fused pipeline (fast.hip)
// and the fast kernels of the general path (rowpass.hip): complex helpers,
// radix-2/4/8 butterflies and the column FFT of length M held by M/8 threads
// (8 values each) with Stockham exchanges through LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace flups {

template <typename R>
struct C2 {
    R x, y;
};
// storage conversions: complex values loaded from T and computed in double, rounded once on store
template <typename T>
__device__ __forceinline__ C2<double> c2d(C2<T> a) { return {(double)a.x, (double)a.y}; }
template <typename T>
__device__ __forceinline__ C2<T> c2t(C2<double> a) { return {(T)a.x, (T)a.y}; }
// complex value stored as S at byte address p, computed as R
template <typename R, typename S>
__device__ __forceinline__ C2<R> ld2(const char* p) {
    const C2<S> a = *reinterpret_cast<const C2<S>*>(p);
    return {(R)a.x, (R)a.y};
}
template <typename R, typename S>
__device__ __forceinline__ void st2(char* p, C2<R> v) {
    *reinterpret_cast<C2<S>*>(p) = C2<S>{(S)v.x, (S)v.y};
}
// two consecutive reals of type T as one vector access (16 bytes for double, 8 for float)
template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};
template <typename T>
using vec2_t = typename Vec2<T>::type;
// non-temporal vector accesses (lines not kept in the Infinity Cache): for arrays read or
// written once per solve that are too large for the cache anyway
__device__ __forceinline__ double2 ld_nt2(const double2* p) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 a = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return double2{a.x, a.y};
}
__device__ __forceinline__ float2 ld_nt2(const float2* p) {
    typedef float v2 __attribute__((ext_vector_type(2)));
    const v2 a = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return float2{a.x, a.y};
}
__device__ __forceinline__ void st_nt2(double2* p, double2 v) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2*>(p));
}
__device__ __forceinline__ void st_nt2(float2* p, float2 v) {
    typedef float v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2*>(p));
}

template <typename R>
__device__ __forceinline__ C2<R> cadd(C2<R> a, C2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ C2<R> csub(C2<R> a, C2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename R>
__device__ __forceinline__ C2<R> cconj(C2<R> a) { return {a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename R, bool FWD>
__device__ __forceinline__ C2<R> rot90(C2<R> a) { return FWD ? C2<R>{a.y, -a.x} : C2<R>{-a.y, a.x}; }

template <typename R, bool FWD>
__device__ __forceinline__ void dft2(C2<R>& a, C2<R>& b) {
    const C2<R> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
template <typename R, bool FWD>
__device__ __forceinline__ void dft4(C2<R>& a0, C2<R>& a1, C2<R>& a2, C2<R>& a3) {
    const C2<R> s0 = cadd(a0, a2), d0 = csub(a0, a2);
    const C2<R> s1 = cadd(a1, a3), d1 = rot90<R, FWD>(csub(a1, a3));
    a0 = cadd(s0, s1);
    a2 = csub(s0, s1);
    a1 = cadd(d0, d1);
    a3 = csub(d0, d1);
}
// HIN: v[4..7] are zero (zero-padded input) and are not read
template <typename R, bool FWD, bool HIN = false>
__device__ __forceinline__ void dft8(C2<R>* v) {
    const R h = (R)0.70710678118654752440;
    // radix-2 split into even/odd halves, then 4-point DFTs
    C2<R> e0, e1, e2, e3, o0, o1, o2, o3;
    if (HIN) {
        e0 = o0 = v[0];
        e1 = o1 = v[1];
        e2 = o2 = v[2];
        e3 = o3 = v[3];
    } else {
        e0 = cadd(v[0], v[4]), o0 = csub(v[0], v[4]);
        e1 = cadd(v[1], v[5]), o1 = csub(v[1], v[5]);
        e2 = cadd(v[2], v[6]), o2 = csub(v[2], v[6]);
        e3 = cadd(v[3], v[7]), o3 = csub(v[3], v[7]);
    }
    // twiddles w8^k on the odd half
    o1 = FWD ? C2<R>{h * (o1.x + o1.y), h * (o1.y - o1.x)} : C2<R>{h * (o1.x - o1.y), h * (o1.y + o1.x)};
    o2 = rot90<R, FWD>(o2);
    o3 = FWD ? C2<R>{h * (o3.y - o3.x), -h * (o3.x + o3.y)} : C2<R>{-h * (o3.x + o3.y), h * (o3.x - o3.y)};
    dft4<R, FWD>(e0, e1, e2, e3);
    dft4<R, FWD>(o0, o1, o2, o3);
    v[0] = e0; v[2] = e1; v[4] = e2; v[6] = e3;
    v[1] = o0; v[3] = o1; v[5] = o2; v[7] = o3;
}

// twiddle exp(-+2 pi i e / M) from the table tw[t] = exp(-2 pi i t / M)
template <typename R, bool FWD>
__device__ __forceinline__ C2<R> twl(const C2<R>* tw, int e) {
    const C2<R> w = tw[e];
    return FWD ? w : cconj(w);
}

// Gather of n values by the threads t0, t0 + step, t0 + 2 step, ...: element q is get(q) and
// is handed to put(q, value) (an LDS write). get must load from a valid address for every
// q < n (clamp the index, select zero afterwards) so that its load is unconditional. The U
// loads of a thread's next U elements are all issued before the first of their LDS writes:
// a load behind a branch followed by its LDS write would wait for that one load, leaving a
// single load of the thread in flight.
template <int U, typename Get, typename Put>
__device__ __forceinline__ void staged_gather(int t0, int step, int n, Get&& get, Put&& put) {
    for (int q0 = t0; q0 < n; q0 += U * step) {
        decltype(get(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = get(min(q0 + u * step, n - 1));
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u * step < n) put(q0 + u * step, v[u]);
    }
}

// LDS slot of column element e: e*ES for interleaved columns (ES columns side by
// side), or e + e/8 for a single contiguous column (ES = 0): the pad slot every 8
// elements keeps the stride-8 writes of the first stage on distinct banks.
template <int ES>
__device__ __forceinline__ int lslot(int e) { return ES > 0 ? e * ES : e + (e >> 3); }

// Full column FFT of length M held by a team of T = M/8 threads (j = 0..T-1).
// On entry v[i] = x[j + i*T]; on exit v[i] = X[j + i*T]. lds points at the
// column (element e at lds[lslot<ES>(e)]); the whole block must call this together.
// HIN: the upper half of the input is zero (v[4..7] not read). Outputs the caller
// never uses (the upper half of a truncated inverse) are removed by the compiler.
// LTW: tw is a copy of the table in LDS, and each radix-8 stage reads its seven
// twiddles w^q from it (seven LDS reads instead of three table loads and four
// complex products: fewer VALU instructions for kernels bound by them).
template <typename R, int M, bool FWD, int ES, bool HIN = false, bool LTW = false>
__device__ __forceinline__ void column_fft(C2<R>* v, C2<R>* lds, int j, const C2<R>* tw) {
    constexpr int T = M / 8;
    // table entries w, w^2, w^4 of the next radix-8 stage, loaded one stage ahead
    // so that their latency hides behind the current stage and its LDS exchange
    // (loading all seven w^q instead of forming w^3, w^5, w^6, w^7 by products
    // competes with the data loads: FP32 z pass 184 -> 236 us at 256^3)
    C2<R> r1 = {1, 0}, r2 = {1, 0}, r4 = {1, 0};
    int span = 1;
#pragma unroll
    for (int st = 0; st < 8; ++st) {
        // stages of radix 8 while 8*span <= M, then one radix 4 or 2 stage
        if (span >= M) break;
        const int P = (span * 8 <= M) ? 8 : (M / span);
        if (span > 1) {
            __syncthreads();   // previous stage's writes are visible
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = lds[lslot<ES>(j + i * T)];
        }
        if (P == 8) {
            const int k = j % span;
            C2<R> n1 = {1, 0}, n2 = {1, 0}, n4 = {1, 0};
            const int nspan = span * 8;
            if (!LTW && nspan * 8 <= M) {
                const int e = (j % nspan) * (M / (nspan * 8));
                n1 = twl<R, FWD>(tw, e);
                n2 = twl<R, FWD>(tw, 2 * e);
                n4 = twl<R, FWD>(tw, 4 * e);
            }
            if (span > 1 && LTW) {
                const int e = k * (M / (span * 8));   // q e < M
#pragma unroll
                for (int q = 1; q < 8; ++q) v[q] = cmul(v[q], twl<R, FWD>(tw, q * e));
            } else if (span > 1) {
                // w^q, q = 1..7, from w, w^2, w^4 and four products
                C2<R> tw8[8];
                tw8[1] = r1;
                tw8[2] = r2;
                tw8[4] = r4;
                tw8[3] = cmul(tw8[1], tw8[2]);
                tw8[5] = cmul(tw8[1], tw8[4]);
                tw8[6] = cmul(tw8[2], tw8[4]);
                tw8[7] = cmul(tw8[3], tw8[4]);
#pragma unroll
                for (int q = 1; q < 8; ++q) v[q] = cmul(v[q], tw8[q]);
            }
            if (HIN && span == 1) dft8<R, FWD, true>(v);
            else dft8<R, FWD>(v);
            if (span * 8 < M) {
                __syncthreads();   // everyone has read this stage's inputs
                const int base = (j / span) * span * 8 + k;
#pragma unroll
                for (int u = 0; u < 8; ++u) lds[lslot<ES>(base + u * span)] = v[u];
            }
            r1 = n1;
            r2 = n2;
            r4 = n4;
        } else if (P == 4) {
            // two butterflies b = j + t*T (t = 0, 1), inputs v[t + 2q]
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int b = j + t * T, k = b % span, stride = M / (span * 4);
                C2<R> a[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) a[q] = v[t + 2 * q];
#pragma unroll
                for (int q = 1; q < 4; ++q) a[q] = cmul(a[q], twl<R, FWD>(tw, q * k * stride));
                dft4<R, FWD>(a[0], a[1], a[2], a[3]);
#pragma unroll
                for (int u = 0; u < 4; ++u) v[t + 2 * u] = a[u];
            }
        } else {   // P == 2, four butterflies b = j + t*T, inputs v[t], v[t+4]
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int b = j + t * T, k = b % span, stride = M / (span * 2);
                C2<R> a0 = v[t], a1 = cmul(v[t + 4], twl<R, FWD>(tw, k * stride));
                dft2<R, FWD>(a0, a1);
                v[t] = a0;
                v[t + 4] = a1;
            }
        }
        span *= P;
    }
}

// Column FFT of length H = P Q (P = 1, 3, 5 or 7; Q a power of two, at least 8) held by
// T = Q/8 threads with 8P values each: on entry v[i] = x[j + i T], on exit v[i] = X[j + i T]
// (i < 8P). For P > 1 (Cooley-Tukey with n = Q n1 + n2, k = k1 + P k2): the P-point DFTs
// over n1 (a thread holds every n1 of its n2 = j + i T), the twiddles w_H^(n2 k1), the P
// Q-point column FFTs over n2 one after the other on the same LDS, and an LDS exchange back
// to the natural order. twq: Q-point table; twh[e * hs] = exp(-2 pi i e / H). The whole block
// calls this together; it ends with LDS reads (sync before the next LDS write).
template <typename R, int P, int Q, bool FWD, int ES, bool HIN = false, bool LTW = false>
__device__ __forceinline__ void column_fft_pq(C2<R>* v, C2<R>* lds, int j, const C2<R>* twq, const C2<R>* twh, int hs) {
    static_assert(P == 1 || !LTW, "LDS twiddles for powers of two only");
    if constexpr (P == 1) {
        column_fft<R, Q, FWD, ES, HIN, LTW>(v, lds, j, twq);
    } else {
        constexpr int T = Q / 8;
        C2<R> wp[P];   // w_P^m = w_H^(m Q)
        wp[0] = {1, 0};
#pragma unroll
        for (int m = 1; m < P; ++m) wp[m] = twl<R, FWD>(twh, m * Q * hs);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            C2<R> a[P];
#pragma unroll
            for (int n1 = 0; n1 < P; ++n1) a[n1] = v[i + 8 * n1];
            const int n2 = j + i * T;
#pragma unroll
            for (int k1 = 0; k1 < P; ++k1) {
                C2<R> acc = a[0];
#pragma unroll
                for (int n1 = 1; n1 < P; ++n1) acc = cadd(acc, cmul(a[n1], wp[(n1 * k1) % P]));
                if (k1 > 0) acc = cmul(acc, twl<R, FWD>(twh, n2 * k1 * hs));   // n2 k1 < H
                v[i + 8 * k1] = acc;
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            if (k1 > 0) __syncthreads();   // the previous column FFT's last LDS reads are done
            column_fft<R, Q, FWD, ES>(v + 8 * k1, lds, j, twq);   // v[i + 8 k1] = X[k1 + P (j + i T)]
        }
        __syncthreads();
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1)
#pragma unroll
            for (int i = 0; i < 8; ++i) lds[lslot<ES>(k1 + P * (j + i * T))] = v[i + 8 * k1];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8 * P; ++i) v[i] = lds[lslot<ES>(j + i * T)];
    }
}

// Column FFT of length H = PF Q by PF teams of T = Q/8 threads (team k1, thread j of the
// team), 8 values per thread: team k1 forms its inputs y[n2] = w_H^(n2 k1) sum_n1 z[Q n1 + n2]
// w_PF^(n1 k1) (n2 = j + i T, z read through load(m)), then runs the Q-point column FFT on its
// own LDS region. On exit v[i] = Z[k1 + PF (j + i T)]. All of load()'s reads happen before the
// column FFT's first LDS write, so the input may sit in the same LDS. ES = 0: regions of
// Q + Q/8 slots; ES > 0: Q ES slots (interleaved columns). PF = 1 is the plain column FFT.
template <typename R, int PF, int Q, bool FWD, int ES, typename Load>
__device__ __forceinline__ void fft_pq_team(C2<R>* v, C2<R>* lds, int k1, int j, Load&& load, const C2<R>* twq,
                                            const C2<R>* twh, int hs) {
    constexpr int T = Q / 8;
    if constexpr (PF == 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = load(j + i * T);
        column_fft<R, Q, FWD, ES>(v, lds, j, twq);
    } else {
        C2<R> wk[PF];   // w_PF^(n1 k1) = w_H^(Q ((n1 k1) mod PF))
        wk[0] = {1, 0};
#pragma unroll
        for (int n1 = 1; n1 < PF; ++n1) wk[n1] = twl<R, FWD>(twh, ((n1 * k1) % PF) * Q * hs);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int n2 = j + i * T;
            C2<R> acc = load(n2);
#pragma unroll
            for (int n1 = 1; n1 < PF; ++n1) acc = cadd(acc, cmul(load(Q * n1 + n2), wk[n1]));
            v[i] = k1 > 0 ? cmul(acc, twl<R, FWD>(twh, n2 * k1 * hs)) : acc;   // n2 k1 < H
        }
        column_fft<R, Q, FWD, ES>(v, lds + k1 * (ES > 0 ? Q * ES : Q + Q / 8), j, twq);
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
// hardware block b runs on XCD b % 8. Renumber so that consecutive logical
// blocks run on the same XCD at about the same time (they share cache lines).
__device__ __forceinline__ int xcd_group(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8, slot = b / 8;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

}  // namespace flups
